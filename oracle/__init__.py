"""CPU oracle of the PlenVDB hot path (test infrastructure only)."""
