"""GPU parity of the stage-per-kernel ops (through the C ABI) against the oracle."""
import os

import numpy as np
import pytest
import torch

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _print_case(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def c1_topo():
    from plenvdb_amd import scenes
    from plenvdb_amd.topology import Topology

    return Topology.from_mask(scenes.occupancy_mask(scenes.CONFIGS["c1"]))


def test_topology_matches_nanovdb(c1_topo):
    g = dict(np.load(os.path.join(GOLDEN, "nanovdb_c1.npz")))
    o, m = c1_topo.leaves()
    np.testing.assert_array_equal(o, g["leaf_origins"])
    np.testing.assert_array_equal(m, g["leaf_masks"])
    s, a = c1_topo.slots(g["queries"])
    np.testing.assert_array_equal(s, g["q_slot"])
    np.testing.assert_array_equal(a, g["q_active"].astype(bool))


def test_sparse_topology_from_leaves_matches_nanovdb():
    from plenvdb_amd.topology import Topology

    g = dict(np.load(os.path.join(GOLDEN, "nanovdb_sparse.npz")))
    rng = np.random.default_rng(0)
    perm = rng.permutation(len(g["leaf_origins"]))
    t = Topology.from_leaves(g["leaf_origins"][perm], g["leaf_masks"][perm], [1, 1, 1])
    assert t.two_level
    o, m = t.leaves()
    np.testing.assert_array_equal(o, g["leaf_origins"])
    np.testing.assert_array_equal(m, g["leaf_masks"])
    s, a = t.slots(g["queries"])
    np.testing.assert_array_equal(s, g["q_slot"])


@pytest.mark.parametrize("name,layout", [("nanovdb_c1", 1), ("nanovdb_c1", 2), ("nanovdb_sparse", 0)])
def test_density_gather_bit_exact_vs_nanovdb(name, layout):
    """DensityVDB.forward equals the reference kernel's arithmetic evaluated
    through NanoVDB's accessor, bit for bit."""
    from plenvdb_amd.plenvdb import DensityVDB
    from plenvdb_amd.topology import Topology
    from test_oracle_golden import slot_values

    g = dict(np.load(os.path.join(GOLDEN, name + ".npz")))
    _print_case("density gather", name, layout)
    t = Topology.from_leaves(g["leaf_origins"], g["leaf_masks"], [1, 1, 1], layout=layout)
    assert t.two_level == (layout == 2 or name == "nanovdb_sparse")
    v = DensityVDB([1, 1, 1], 1, topology=t)
    sv = slot_values(g["leaf_origins"].astype(np.int64), g["leaf_masks"], g["vals"])
    v.grid.copy_(torch.from_numpy(sv).reshape(-1, 1))
    p = g["points"]
    out = v.forward(p[:, 0], p[:, 1], p[:, 2])  # numpy in -> numpy out
    np.testing.assert_array_equal(out, g["p_tri_ref"])


def _random_points(R, n, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1.0, R, size=(n, 3)).astype(np.float32)


@pytest.mark.parametrize("C", [1, 3, 12])
def test_gather_and_scatter_vs_oracle(c1_topo, oracle_mod, C):
    from plenvdb_amd.plenvdb import ColorVDB, DensityVDB

    _print_case("gather/scatter", "C", C)
    o, m = c1_topo.leaves()
    bits = np.unpackbits(m.view(np.uint8).reshape(-1, 64), axis=1, bitorder="little").astype(bool)
    dom = oracle_mod.SlotDomain(o, bits)
    rng = np.random.default_rng(C)
    vals = rng.normal(size=(c1_topo.nslots, C)).astype(np.float32)
    vdb = DensityVDB([128] * 3, 1, topology=c1_topo) if C == 1 else ColorVDB([128] * 3, C, topology=c1_topo)
    order = 0 if C == 1 else 1
    vdb.grid.copy_(torch.from_numpy(vals))
    p = _random_points(128, 20000, 7 + C)
    xs, ys, zs = (torch.from_numpy(np.ascontiguousarray(p[:, i])).cuda() for i in range(3))
    out = vdb.forward(xs, ys, zs).cpu().numpy()
    ref = oracle_mod.trilinear_fwd(dom, vals, p[:, 0], p[:, 1], p[:, 2], order, "f32")
    np.testing.assert_array_equal(out, ref)  # same operation order, no contraction
    g = rng.normal(size=(len(p), C)).astype(np.float32)
    vdb.backward(xs, ys, zs, torch.from_numpy(g).cuda())
    got = vdb.grad.cpu().numpy()
    ref_g = oracle_mod.trilinear_bwd(dom, C, p[:, 0], p[:, 1], p[:, 2], g.astype(np.float64), order, "f64")
    # atomics sum in arbitrary order: FP32 accumulation against the FP64 sum
    scale = np.abs(ref_g).max()
    np.testing.assert_allclose(got, ref_g, rtol=1e-5, atol=1e-6 * scale)
    assert np.array_equal(got != 0, ref_g != 0)


def test_sampler_and_mask_vs_oracle(c1_topo, oracle_mod):
    from plenvdb_amd import render_utils, scenes

    _print_case("sampler c1")
    sc = scenes.make_scene("c1", n_rays=2048)
    ro = torch.from_numpy(sc["rays_o"]).cuda()
    rd = torch.from_numpy(sc["rays_d"]).cuda()
    stepdist = np.float32(sc["stepsize"]) * np.float32(sc["voxel_size"])
    pts, outb, rid, sid, ns, tmin, tmax = render_utils.sample_pts_on_rays(
        ro, rd, sc["xyz_min"], sc["xyz_max"], sc["near"], 1e9, float(stepdist))
    r_tmin, r_tmax, r_ns = oracle_mod.ray_bounds(sc["rays_o"], sc["rays_d"], sc["xyz_min"],
                                                 sc["xyz_max"], sc["near"], 1e9, stepdist)
    np.testing.assert_array_equal(tmin.cpu().numpy(), r_tmin)
    np.testing.assert_array_equal(tmax.cpu().numpy(), r_tmax)
    np.testing.assert_array_equal(ns.cpu().numpy(), r_ns)
    r_pts, r_rid, r_sid, r_inb = oracle_mod.sample_pts(sc["rays_o"], sc["rays_d"], r_tmin, r_ns,
                                                       stepdist, sc["xyz_min"], sc["xyz_max"])
    np.testing.assert_array_equal(pts.cpu().numpy(), r_pts)
    np.testing.assert_array_equal(rid.cpu().numpy(), r_rid)
    np.testing.assert_array_equal(sid.cpu().numpy(), r_sid)
    np.testing.assert_array_equal((~outb).cpu().numpy(), r_inb)
    # occupancy: dense-mask lookup and VDB value-mask lookup agree with the oracle
    reso = np.array(sc["reso"], np.float32)
    scale = ((reso - 1) / (sc["xyz_max"] - sc["xyz_min"])).astype(np.float32)
    shift = (-sc["xyz_min"] * scale).astype(np.float32)
    ref_keep = oracle_mod.mask_lookup(sc["mask"], r_pts, scale, shift)
    keep_dense = render_utils.maskcache_lookup(torch.from_numpy(sc["mask"]).cuda(), pts, scale, shift)
    keep_vdb = render_utils.active_lookup(c1_topo, pts, scale, shift)
    np.testing.assert_array_equal(keep_dense.cpu().numpy(), ref_keep)
    np.testing.assert_array_equal(keep_vdb.cpu().numpy(), ref_keep)


def test_raw2alpha_alpha2weight_vs_oracle(oracle_mod):
    from plenvdb_amd import render_utils

    _print_case("raw2alpha/alpha2weight")
    rng = np.random.default_rng(3)
    n_rays = 300
    counts = rng.integers(0, 200, size=n_rays)
    counts[:5] = 0
    rid = np.repeat(np.arange(n_rays), counts).astype(np.int64)
    sigma = rng.normal(0, 3, size=len(rid)).astype(np.float32)
    shift, interval = float(np.float32(-4.595)), 0.5
    e, a = render_utils.raw2alpha(torch.from_numpy(sigma).cuda(), shift, interval)
    re, ra = oracle_mod.raw2alpha(sigma, shift, interval, "f32")
    np.testing.assert_allclose(e.cpu().numpy(), re, rtol=2e-7)
    np.testing.assert_allclose(a.cpu().numpy(), ra, rtol=0, atol=2e-7)
    # alpha2weight on the oracle's alphas: identical inputs -> identical sequential products
    at = torch.from_numpy(ra).cuda()
    w, T, tl, i_s, i_e = render_utils.alpha2weight(at, torch.from_numpy(rid).cuda(), n_rays)
    rw, rT, rtl, ris, rie = oracle_mod.alpha2weight(ra, rid, n_rays, "f32")
    np.testing.assert_array_equal(w.cpu().numpy(), rw)
    np.testing.assert_array_equal(T.cpu().numpy(), rT)
    np.testing.assert_array_equal(tl.cpu().numpy(), rtl)
    np.testing.assert_array_equal(i_e.cpu().numpy(), rie)
    gw = rng.normal(size=len(rid)).astype(np.float32)
    gl = rng.normal(size=n_rays).astype(np.float32)
    ga = render_utils.alpha2weight_backward(at, w, T, tl, i_s, i_e, n_rays,
                                            torch.from_numpy(gw).cuda(), torch.from_numpy(gl).cuda())
    rga = oracle_mod.alpha2weight_bwd(ra, rw, rT, rtl, ris, rie, gw, gl, "f32")
    np.testing.assert_array_equal(ga.cpu().numpy(), rga)
    gs = render_utils.raw2alpha_backward(e, ga, interval)
    rgs = oracle_mod.raw2alpha_bwd(re, rga, interval, "f32")
    np.testing.assert_allclose(gs.cpu().numpy(), rgs, rtol=1e-6, atol=1e-12)


def test_vdb_adam_bit_exact(c1_topo, oracle_mod):
    from plenvdb_amd.plenvdb import ColorOpt, ColorVDB, DensityOpt, DensityVDB

    _print_case("vdb adam")
    rng = np.random.default_rng(4)
    act = c1_topo.active_slot_mask()
    for cls, ocls, C, group in ((DensityVDB, DensityOpt, 1, 1), (ColorVDB, ColorOpt, 12, 3)):
        vdb = cls([128] * 3, C, topology=c1_topo)
        p = rng.normal(size=(c1_topo.nslots, C)).astype(np.float32)
        g = rng.normal(size=(c1_topo.nslots, C)).astype(np.float32)
        g[rng.random(c1_topo.nslots) < 0.3] = 0
        if C == 12:
            g[rng.random(c1_topo.nslots) < 0.2, 4] = 0
        vdb.grid.copy_(torch.from_numpy(p))
        vdb.grad.copy_(torch.from_numpy(g))
        opt = ocls(vdb, 0.1, 1e-8, 0.9, 0.99)
        rp, rm, rv = p.copy(), np.zeros_like(p), np.zeros_like(p)
        for step in range(1, 4):
            for mode in (1, 0):
                opt.step(mode)
                sz = oracle_mod.adam_stepsz(opt.lr, 0.9, 0.99, opt.getStep())
                oracle_mod.vdb_adam(rp, g, rm, rv, act, group, mode, sz, 1e-8, 0.9, 0.99)
        np.testing.assert_array_equal(vdb.grid.cpu().numpy(), rp)
        np.testing.assert_array_equal(opt.exp_avg.cpu().numpy(), rm)
        np.testing.assert_array_equal(opt.exp_avg_sq.cpu().numpy(), rv)
        opt.zero_grad()
        gz = vdb.grad.cpu().numpy()
        assert (gz[act] == 0).all() and np.array_equal(gz[~act], g[~act])


def test_dense_adam_modes(oracle_mod):
    from plenvdb_amd import render_utils

    rng = np.random.default_rng(5)
    n = 22019
    for mode, fn in enumerate((render_utils.adam_upd, render_utils.masked_adam_upd, None)):
        p = rng.normal(size=n).astype(np.float32)
        g = rng.normal(size=n).astype(np.float32)
        g[::7] = 0
        pl = rng.uniform(size=n).astype(np.float32)
        tp, tg = torch.from_numpy(p.copy()).cuda(), torch.from_numpy(g).cuda()
        tm, tv = torch.zeros_like(tp), torch.zeros_like(tp)
        rp, rm, rv = p.copy(), np.zeros_like(p), np.zeros_like(p)
        for step in (1, 2):
            if mode == 2:
                render_utils.adam_upd_with_perlr(tp, tg, tm, tv, torch.from_numpy(pl).cuda(), step,
                                                 0.9, 0.99, 1e-3, 1e-8)
            else:
                fn(tp, tg, tm, tv, step, 0.9, 0.99, 1e-3, 1e-8)
            sz = oracle_mod.adam_stepsz(1e-3, 0.9, 0.99, step)
            oracle_mod.dense_adam(rp, g, rm, rv, mode, sz, 1e-8, 0.9, 0.99,
                                  per_lr=pl if mode == 2 else None)
        np.testing.assert_array_equal(tp.cpu().numpy(), rp)


@pytest.mark.parametrize("cfg", ["c2", "c4"])
def test_empty_run_map_is_the_chessboard_distance(cfg):
    """The marchers' empty-run map equals the chessboard distance from each
    empty 8^3 cell to the nearest leaf (scipy.ndimage.distance_transform_cdt),
    capped at 64; 0 on leaves."""
    from scipy import ndimage

    from plenvdb_amd import scenes
    from plenvdb_amd.topology import Topology

    topo = Topology.from_mask(scenes.occupancy_mask(scenes.CONFIGS[cfg]))
    got = topo.empty_run()
    assert got is not None and list(got.shape) == topo.leaf_dim
    o, _ = topo.leaves()
    occ = np.zeros(topo.leaf_dim, dtype=bool)
    cells = (o >> 3) - np.asarray(topo.leaf_lo)
    occ[cells[:, 0], cells[:, 1], cells[:, 2]] = True
    want = np.minimum(ndimage.distance_transform_cdt(~occ, metric="chessboard"), 64)
    print(f"[result] {cfg}: leaf grid {topo.leaf_dim}, {int(occ.sum())} leaves, "
          f"max run {int(got.max())}", flush=True)
    np.testing.assert_array_equal(got, want.astype(np.uint8))
    assert got.max() >= 2  # the map does skip something here
