/*
 * CPU oracle library for the PlenVDB hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * Built by oracle/Makefile into oracle/liboracle.so (gcc, OpenMP,
 * -ffp-contract=off so every operation rounds as written).  Loaded through
 * ctypes by oracle/oracle.py, which only tests/, the smoke check and
 * bench.py's CPU baseline import.  The product never links or calls it.
 *
 * The function bodies live in pvdb_oracle_impl.h, compiled once in FP32
 * reference order (suffix _f32) and once in FP64 (suffix _f64).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define REAL float
#define SFX _f32
#include "pvdb_oracle_impl.h"
#undef REAL
#undef SFX

#define REAL double
#define SFX _f64
#include "pvdb_oracle_impl.h"
#undef REAL
#undef SFX

#include "pvdb_oracle_render.h"

int pvdb_oracle_abi_version(void) { return 1; }

#ifdef _OPENMP
#include <omp.h>
#endif
/* Threads of the OpenMP loops from here on (the CPU baseline's 1-thread run). */
int pvdb_oracle_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}
