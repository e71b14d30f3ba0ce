"""Pins the oracle's topology and accessor semantics against golden vectors
produced by the reference's own NanoVDB (tests/golden/make_golden.py)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

CASES = ["nanovdb_c1", "nanovdb_sparse"]


def load_case(name):
    return dict(np.load(os.path.join(GOLDEN, name + ".npz")))


def voxels_of(origins, masks):
    """Active voxels (lexicographic order) of a leaf list."""
    bits = np.unpackbits(np.ascontiguousarray(masks).view(np.uint8).reshape(-1, 64), axis=1,
                         bitorder="little").astype(bool)
    off = np.arange(512)
    loc = np.stack([off >> 6, (off >> 3) & 7, off & 7], 1)
    vox = (origins[:, None, :].astype(np.int64) + loc[None])[bits]
    order = np.lexsort((vox[:, 2], vox[:, 1], vox[:, 0]))
    return vox[order]


def slot_values(origins, masks, vals):
    """Slot payload with the fixture values on active voxels and 0 elsewhere."""
    from oracle import oracle

    vox = voxels_of(origins, masks)
    s = oracle.slot_of(origins, vox)
    sv = np.zeros(len(origins) * 512, np.float32)
    sv[s] = vals
    return sv


@pytest.mark.parametrize("name", CASES)
def test_leaf_order_matches_nanovdb(name, oracle_mod):
    g = load_case(name)
    o = g["leaf_origins"].astype(np.int64)
    rng = np.random.default_rng(0)
    perm = rng.permutation(len(o))
    order = oracle_mod.nanovdb_leaf_order(o[perm])
    np.testing.assert_array_equal(o[perm][order], o)


def test_leaves_from_mask_matches_nanovdb(oracle_mod):
    from plenvdb_amd import scenes

    g = load_case("nanovdb_c1")
    origins, vm = oracle_mod.leaves_from_mask(scenes.occupancy_mask(scenes.CONFIGS["c1"]))
    np.testing.assert_array_equal(origins, g["leaf_origins"])
    bits = np.unpackbits(g["leaf_masks"].view(np.uint8).reshape(-1, 64), axis=1,
                         bitorder="little").astype(bool)
    np.testing.assert_array_equal(vm, bits)


@pytest.mark.parametrize("name", CASES)
def test_slots_values_active(name, oracle_mod):
    g = load_case(name)
    o = g["leaf_origins"].astype(np.int64)
    slots = oracle_mod.slot_of(o, g["queries"])
    np.testing.assert_array_equal(slots, g["q_slot"])
    np.testing.assert_array_equal(slots >= 0, g["q_cached"].astype(bool))
    sv = slot_values(o, g["leaf_masks"], g["vals"])
    val = np.where(slots >= 0, sv[np.maximum(slots, 0)], 0).astype(np.float32)
    np.testing.assert_array_equal(val, g["q_value"])
    bits = np.unpackbits(g["leaf_masks"].view(np.uint8).reshape(-1, 64), axis=1,
                         bitorder="little").astype(bool).reshape(-1)
    act = np.where(slots >= 0, bits[np.maximum(slots, 0)], False)
    np.testing.assert_array_equal(act, g["q_active"].astype(bool))


@pytest.mark.parametrize("name", CASES)
def test_trilinear_matches_nanovdb_accessor(name, oracle_mod):
    g = load_case(name)
    o = g["leaf_origins"].astype(np.int64)
    bits = np.unpackbits(g["leaf_masks"].view(np.uint8).reshape(-1, 64), axis=1,
                         bitorder="little").astype(bool)
    dom = oracle_mod.SlotDomain(o, bits)
    sv = slot_values(o, g["leaf_masks"], g["vals"])
    p = g["points"]
    f32 = oracle_mod.trilinear_fwd(dom, sv, p[:, 0], p[:, 1], p[:, 2], 0, "f32")[:, 0]
    # same operation order as densityvdb.cu:115-122 through NanoVDB's accessor: bit-exact
    np.testing.assert_array_equal(f32, g["p_tri_ref"])
    f64 = oracle_mod.trilinear_fwd(dom, sv.astype(np.float64), p[:, 0], p[:, 1], p[:, 2], 0, "f64")[:, 0]
    # NanoVDB's TrilinearSampler evaluates the same interpolant differently
    np.testing.assert_allclose(f64, g["p_tri_sampler"], rtol=1e-5, atol=2e-5)
