"""Pins the oracle's arithmetic against closed forms, library routines, brute
force and finite differences, chosen so that a dropped term, a wrong sign or
index, or a transposed operand fails one of them."""
import numpy as np
import pytest
import torch


def dense_block(oracle, n_leaf=2, C=1, seed=0, fill=None):
    """Fully dense n_leaf^3-leaf block at the origin with random (or given) values."""
    R = 8 * n_leaf
    mask = np.ones((R, R, R), bool)
    origins, vm = oracle.leaves_from_mask(mask)
    dom = oracle.SlotDomain(origins, vm)
    coords = dom.coords()
    if fill is None:
        rng = np.random.default_rng(seed)
        vals = rng.normal(size=(dom.nslots, C))
    else:
        vals = fill(coords)
    return dom, vals, coords, R


def test_trilinear_reproduces_linear_field(oracle_mod):
    # NanoVDB TestNanoVDB.cc:4277 pattern: trilinear reproduces a linear field.
    a = np.array([0.7, -1.3, 2.1])
    dom, vals, coords, R = dense_block(oracle_mod, fill=lambda c: (c @ a + 0.25)[:, None])
    rng = np.random.default_rng(1)
    p = rng.uniform(0, R - 1, size=(2000, 3)).astype(np.float32)
    for order in (0, 1):
        got = oracle_mod.trilinear_fwd(dom, vals, p[:, 0], p[:, 1], p[:, 2], order, "f64")[:, 0]
        np.testing.assert_allclose(got, p.astype(np.float64) @ a + 0.25, rtol=0, atol=1e-10)
        got32 = oracle_mod.trilinear_fwd(dom, vals.astype(np.float32), p[:, 0], p[:, 1], p[:, 2],
                                         order, "f32")[:, 0]
        np.testing.assert_allclose(got32, p.astype(np.float64) @ a + 0.25, rtol=1e-5, atol=1e-4)


def test_trilinear_equals_grid_sample(oracle_mod):
    # On a fully dense block the interpolant equals F.grid_sample(align_corners=True)
    # (the DenseGrid path, plenvdb/lib/grid.py:151-162).
    C = 3
    dom, vals, coords, R = dense_block(oracle_mod, C=C, seed=3)
    dense = np.zeros((R, R, R, C))
    dense[coords[:, 0], coords[:, 1], coords[:, 2]] = vals
    rng = np.random.default_rng(4)
    p = rng.uniform(0, R - 1, size=(1500, 3)).astype(np.float32)
    got = oracle_mod.trilinear_fwd(dom, vals, p[:, 0], p[:, 1], p[:, 2], 1, "f64")
    g = torch.tensor(dense, dtype=torch.float64).permute(3, 0, 1, 2)[None]
    ind = torch.tensor(p.astype(np.float64) / (R - 1) * 2 - 1).flip(-1).reshape(1, 1, 1, -1, 3)
    ref = torch.nn.functional.grid_sample(g, ind, mode="bilinear", align_corners=True)
    ref = ref.reshape(C, -1).T.numpy()
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12)


def sparse_partial(oracle, seed=5):
    """A few partial leaves with gaps, so points straddle existing and missing leaves."""
    rng = np.random.default_rng(seed)
    R = 32
    lm = rng.random((4, 4, 4)) < 0.5
    vm = rng.random((R, R, R)) < 0.7
    mask = np.repeat(np.repeat(np.repeat(lm, 8, 0), 8, 1), 8, 2) & vm
    origins, vmask = oracle.leaves_from_mask(mask)
    return oracle.SlotDomain(origins, vmask), R


def test_scatter_is_adjoint_of_gather(oracle_mod):
    dom, R = sparse_partial(oracle_mod)
    rng = np.random.default_rng(6)
    C = 12
    G = rng.normal(size=(dom.nslots, C))
    p = rng.uniform(-1.5, R + 0.5, size=(3000, 3)).astype(np.float32)
    g = rng.normal(size=(len(p), C))
    for order in (0, 1):
        Cc = 1 if order == 0 else C
        fwd = oracle_mod.trilinear_fwd(dom, G[:, :Cc], p[:, 0], p[:, 1], p[:, 2], order, "f64")
        bwd = oracle_mod.trilinear_bwd(dom, Cc, p[:, 0], p[:, 1], p[:, 2], g[:, :Cc], order, "f64")
        lhs = np.sum(fwd * g[:, :Cc])
        rhs = np.sum(G[:, :Cc] * bwd)
        assert abs(lhs - rhs) <= 1e-9 * max(1.0, abs(lhs))


def test_scatter_leaf_must_exist(oracle_mod):
    # Leaf at origin exists (all voxels inactive except one), its +x neighbour does not.
    vm = np.zeros((1, 512), bool)
    vm[0, 0] = True
    dom = oracle_mod.SlotDomain(np.array([[0, 0, 0]]), vm)
    p = np.array([[7.25, 3.5, 2.0]], np.float32)
    grad = oracle_mod.trilinear_bwd(dom, 1, p[:, 0], p[:, 1], p[:, 2], np.ones((1, 1)), 0, "f64")
    touched = np.nonzero(grad[:, 0])[0]
    coords = dom.coords()[touched]
    # only x=7 corners survive; inactive voxels of the existing leaf receive the add
    assert set(map(tuple, coords.tolist())) == {(7, 3, 2), (7, 4, 2)}
    np.testing.assert_allclose(grad[touched, 0], [0.75 * 0.5, 0.75 * 0.5])


def test_ray_bounds_hand_cases(oracle_mod):
    mn, mx = np.array([-1.5] * 3, np.float32), np.array([1.5] * 3, np.float32)
    o = np.array([[-4, 0, 0], [0, 0, 4], [0.5, 0.2, -4]], np.float32)
    d = np.array([[1, 0, 0], [0, 0, -1], [0, 0, 2]], np.float32)
    tmin, tmax, ns = oracle_mod.ray_bounds(o, d, mn, mx, 0.0, 1e9, 0.5, prec="f64")
    np.testing.assert_allclose(tmin, [2.5, 2.5, 1.25])
    np.testing.assert_allclose(tmax, [5.5, 5.5, 2.75])
    np.testing.assert_array_equal(ns, [6, 6, 6])  # ceil((tmax-tmin)*|d|/0.5)
    # a miss gives tmin > tmax clamped to near; the count is at least 1
    tmin, tmax, ns = oracle_mod.ray_bounds(np.array([[5, 5, 5]], np.float32),
                                           np.array([[1, 0, 0]], np.float32), mn, mx, 2.0, 1e9, 0.5)
    assert ns[0] == 1 and tmin[0] == 2.0 and tmax[0] == 2.0


def test_sample_points_and_mask_rounding(oracle_mod):
    mn, mx = np.array([-1.5] * 3, np.float32), np.array([1.5] * 3, np.float32)
    o = np.array([[-4, 0.1, 0.2]], np.float32)
    d = np.array([[2, 0, 0]], np.float32)
    tmin, tmax, ns = oracle_mod.ray_bounds(o, d, mn, mx, 0.0, 1e9, 0.25, prec="f64")
    pts, rid, sid, inb = oracle_mod.sample_pts(o, d, tmin, ns, 0.25, mn, mx, prec="f64")
    np.testing.assert_allclose(pts[:, 0], -1.5 + 0.25 * np.arange(ns[0]))
    assert inb.all() and len(pts) == 12
    # round half away from zero: 2.5 -> 3, -0.5 -> -1 (outside)
    mask = np.zeros((4, 4, 4), bool)
    mask[3, 0, 0] = True
    got = oracle_mod.mask_lookup(mask, np.array([[2.5, 0, 0], [2.4999, 0, 0]]), [1, 1, 1], [0, 0, 0])
    np.testing.assert_array_equal(got, [True, False])


def test_raw2alpha_closed_form_and_derivative(oracle_mod):
    s = np.linspace(-12, 8, 201)
    shift = float(np.float32(-4.595))  # the kernels take shift as a float
    e, a = oracle_mod.raw2alpha(s, shift, 0.5, "f64")
    sp = np.log1p(np.exp(s + shift))
    np.testing.assert_allclose(a, 1 - np.exp(-sp * 0.5), rtol=1e-12, atol=1e-15)
    g = oracle_mod.raw2alpha_bwd(e, np.ones_like(s), 0.5, "f64")
    h = 1e-6
    fd = (oracle_mod.raw2alpha(s + h, shift, 0.5, "f64")[1]
          - oracle_mod.raw2alpha(s - h, shift, 0.5, "f64")[1]) / (2 * h)
    # FD of 1-(1+e)^-k cancels badly for tiny alpha: loose absolute floor
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-9)
    # d/ds [1-(1+e^(s+b))^-k] = k e (1+e)^(-k-1)
    ee = np.exp(s + shift)
    np.testing.assert_allclose(g, 0.5 * ee * (1 + ee) ** -1.5, rtol=1e-12)


def test_alpha2weight_invariants_and_gradient(oracle_mod):
    rng = np.random.default_rng(8)
    n_rays = 40
    counts = rng.integers(0, 30, size=n_rays)
    rid = np.repeat(np.arange(n_rays), counts)
    a = rng.uniform(0, 0.15, size=len(rid))
    w, T, tl, i_s, i_e = oracle_mod.alpha2weight(a, rid, n_rays, "f64")
    # no early stop here: sum(w) + T_last = 1 and T is the exclusive product
    np.testing.assert_array_less(1e-3, tl)
    for r in range(n_rays):
        seg = slice(i_s[r], i_e[r])
        assert abs(w[seg].sum() + tl[r] - 1) < 1e-12
        ex = np.concatenate([[1.0], np.cumprod(1 - a[seg])[:-1]]) if counts[r] else []
        np.testing.assert_allclose(T[seg], ex, rtol=1e-12)
    # gradient of L = <gw, w> + <gl, T_last> against central differences
    gw = rng.normal(size=len(a))
    gl = rng.normal(size=n_rays)
    ga = oracle_mod.alpha2weight_bwd(a, w, T, tl, i_s, i_e, gw, gl, "f64")

    def L(x):
        w_, _, tl_, _, _ = oracle_mod.alpha2weight(x, rid, n_rays, "f64")
        return gw @ w_ + gl @ tl_

    h = 1e-7
    idx = rng.choice(len(a), 40, replace=False)
    for i in idx:
        e = np.zeros_like(a)
        e[i] = h
        fd = (L(a + e) - L(a - e)) / (2 * h)
        assert abs(fd - ga[i]) < 1e-6 * max(1, abs(fd))


def test_alpha2weight_early_stop(oracle_mod):
    a = np.array([0.5, 0.9, 0.99, 0.3, 0.2])
    rid = np.zeros(5, np.int64)
    w, T, tl, i_s, i_e = oracle_mod.alpha2weight(a, rid, 1, "f64")
    # T after samples: 0.5, 0.05, 5e-4 (<1e-3: stop after the third sample)
    assert i_e[0] == 3
    np.testing.assert_allclose(w[:3], [0.5, 0.45, 0.0495])
    np.testing.assert_allclose(w[3:], 0)
    np.testing.assert_allclose(T[3:], 1)
    np.testing.assert_allclose(tl, [5e-4])


def test_vdb_adam_closed_form(oracle_mod):
    rng = np.random.default_rng(9)
    n = 64
    p = rng.normal(size=(n, 3))
    g = rng.normal(size=(n, 3))
    g[5] = 0
    g[7, 0] = 0
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    act = np.ones(n, np.uint8)
    act[3] = 0
    p0 = p.copy()
    sz = oracle_mod.adam_stepsz(0.1, 0.9, 0.99, 1)
    np.testing.assert_allclose(sz, 0.1 * np.sqrt(1 - 0.99) / (1 - 0.9), rtol=1e-6)
    oracle_mod.vdb_adam(p, g, m, v, act, 3, 1, sz, 1e-8, 0.9, 0.99, prec="f64")
    f = lambda x: float(np.float32(x))  # the kernels take the hyper-parameters as floats
    exp = p0 - f(sz) * ((1 - f(0.9)) * g) / (f(1e-8) + np.sqrt((1 - f(0.99)) * g * g))
    exp[3] = p0[3]  # inactive
    exp[5] = p0[5]  # whole Vec3 group has zero grad: skipped
    np.testing.assert_allclose(p, exp, rtol=1e-12, atol=1e-12)
    assert m[7, 0] == 0 and m[7, 1] != 0  # the group with one zero component is updated


def test_reference_step_gradient_matches_fd(oracle_mod):
    """Central differences of the composed FP64 reference step (MLP included)
    w.r.t. 256 random slot values with a non-zero gradient: 128 density and 128
    colour entries (fast_color_thres 0, so the culls are frozen)."""
    from plenvdb_amd import scenes

    sc = scenes.make_scene("c2", n_rays=24)
    origins, vm = oracle_mod.leaves_from_mask(sc["mask"])
    rng = np.random.default_rng(1)
    nslots = len(origins) * 512
    den = rng.normal(0, 2, nslots)
    col = rng.normal(0, 0.1, (nslots, 12))
    sc.update(origins=origins, vmask=vm, density=den, color=col, fast_color_thres=0.0)
    out = oracle_mod.reference_train_step(sc, prec="f64")
    gd, gc = out["grad_density"], out["grad_color"]
    cand_d = np.nonzero(np.abs(gd) > 1e-6)[0]
    cand_c = np.argwhere(np.abs(gc) > 1e-6)
    assert len(cand_d) > 10 and len(cand_c) > 10
    h = 1e-5
    for i in rng.choice(cand_d, 128, replace=False):
        sc["density"] = den.copy()
        sc["density"][i] += h
        lp = oracle_mod.reference_train_step(sc, prec="f64")["loss"]
        sc["density"][i] -= 2 * h
        lm = oracle_mod.reference_train_step(sc, prec="f64")["loss"]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - gd[i]) < 1e-5 * max(abs(fd), 1e-6) + 1e-11
    sc["density"] = den
    for s, c in cand_c[rng.choice(len(cand_c), 128, replace=False)]:
        sc["color"] = col.copy()
        sc["color"][s, c] += h
        lp = oracle_mod.reference_train_step(sc, prec="f64")["loss"]
        sc["color"][s, c] -= 2 * h
        lm = oracle_mod.reference_train_step(sc, prec="f64")["loss"]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - gc[s, c]) < 1e-5 * max(abs(fd), 1e-6) + 1e-11


def test_reference_step_loss_extras(oracle_mod):
    """The fine stage's extra loss terms (run.py:553-573) in the FP64 reference
    step: the entropy term equals scipy's binary entropy of the clamped
    alphainv_last; the rgbper term is positive and, its weights being detached,
    leaves the density gradient unchanged; central differences match the slot
    gradients (64 density entries with the entropy term, whose gradient is the
    loss's true derivative, and 64 colour entries with both terms: the colour
    does not move the weights, so there the detach is exact)."""
    from scipy.special import entr

    from plenvdb_amd import scenes

    sc = scenes.make_scene("c2", n_rays=24)
    origins, vm = oracle_mod.leaves_from_mask(sc["mask"])
    rng = np.random.default_rng(2)
    nslots = len(origins) * 512
    den = rng.normal(0, 2, nslots)
    col = rng.normal(0, 0.1, (nslots, 12))
    sc.update(origins=origins, vmask=vm, density=den, color=col, fast_color_thres=0.0)
    base = oracle_mod.reference_train_step(sc, prec="f64")
    ent = oracle_mod.reference_train_step(dict(sc, weight_entropy_last=0.3), prec="f64")
    p = np.clip(base["alphainv_last"], 1e-6, 1 - 1e-6)
    np.testing.assert_allclose(ent["loss"] - base["loss"], 0.3 * np.mean(entr(p) + entr(1 - p)),
                               rtol=1e-10)
    assert np.abs(ent["grad_density"] - base["grad_density"]).max() > 1e-6
    rp = oracle_mod.reference_train_step(dict(sc, weight_rgbper=0.5), prec="f64")
    assert rp["loss"] > base["loss"]
    np.testing.assert_array_equal(rp["grad_density"], base["grad_density"])
    assert np.abs(rp["grad_color"] - base["grad_color"]).max() > 1e-6
    h = 1e-5

    def fd_check(scene, key, idx, g):
        v0 = scene[key]
        vp, vm_ = v0.copy(), v0.copy()
        vp[idx] += h
        vm_[idx] -= h
        lp = oracle_mod.reference_train_step(dict(scene, **{key: vp}), prec="f64")["loss"]
        lm = oracle_mod.reference_train_step(dict(scene, **{key: vm_}), prec="f64")["loss"]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g) < 1e-5 * max(abs(fd), 1e-6) + 1e-11, (key, idx, fd, g)

    sc_e = dict(sc, weight_entropy_last=0.3)
    gd = ent["grad_density"]
    for i in rng.choice(np.nonzero(np.abs(gd) > 1e-6)[0], 64, replace=False):
        fd_check(sc_e, "density", i, gd[i])
    sc_b = dict(sc, weight_entropy_last=0.3, weight_rgbper=0.5)
    gc = oracle_mod.reference_train_step(sc_b, prec="f64")["grad_color"]
    cand_c = np.argwhere(np.abs(gc) > 1e-6)
    for s_, c in cand_c[rng.choice(len(cand_c), 64, replace=False)]:
        fd_check(sc_b, "color", (s_, c), gc[s_, c])


def _rotation(seed):
    q, _ = np.linalg.qr(np.random.default_rng(seed).normal(size=(3, 3)))
    return q * np.sign(np.linalg.det(q))


def test_rays_of_view_closed_forms(oracle_mod):
    """get_rays / ndc_rays (dvgo.py:471-527) in the oracle: the identity pose gives
    the pinhole directions; a rotation pose rotates them (FP64 product); flips
    reverse an axis; inverse_y negates y and z; viewdirs are unit vectors; NDC
    origins lie on z = -1, origin + direction on z = 1, and the NDC image of any
    point of the camera ray lies on the NDC ray."""
    H, W = 37, 53
    K = np.array([[120.5, 0, 26.25], [0, 118.0, 17.75], [0, 0, 1]])
    eye = np.eye(4)[:3]
    ro, rd, vd = oracle_mod.rays_of_view(H, W, K, eye, mode="center")
    jj, ii = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    np.testing.assert_allclose(rd[..., 0], (ii + 0.5 - 26.25) / 120.5, rtol=1e-6)
    np.testing.assert_allclose(rd[..., 1], -(jj + 0.5 - 17.75) / 118.0, rtol=1e-6, atol=1e-7)
    assert np.all(rd[..., 2] == -1) and np.all(ro == 0)
    lt = oracle_mod.rays_of_view(H, W, K, eye, mode="lefttop")[1]
    np.testing.assert_allclose(lt[..., 0], (ii - 26.25) / 120.5, rtol=1e-6)
    fx = oracle_mod.rays_of_view(H, W, K, eye, flip_x=True)[1]
    np.testing.assert_array_equal(fx[..., 0], rd[:, ::-1, 0])
    fy = oracle_mod.rays_of_view(H, W, K, eye, flip_y=True)[1]
    np.testing.assert_array_equal(fy[..., 1], rd[::-1, :, 1])
    iy = oracle_mod.rays_of_view(H, W, K, eye, inverse_y=True)[1]
    np.testing.assert_array_equal(iy[..., 1], -rd[..., 1])
    assert np.all(iy[..., 2] == 1)
    R = _rotation(3)
    c2w = np.concatenate([R, [[0.3], [-1.2], [2.5]]], 1)
    ro, rdR, vdR = oracle_mod.rays_of_view(H, W, K, c2w)
    np.testing.assert_allclose(rdR, np.einsum("km,hwm->hwk", R, rd.astype(np.float64)),
                               rtol=0, atol=1e-6)
    np.testing.assert_allclose(np.linalg.norm(vdR, axis=-1), 1.0, atol=1e-6)
    np.testing.assert_array_equal(ro, np.broadcast_to(np.float32(c2w[:, 3]), ro.shape))
    # NDC: camera looking down -z from z = 3
    c2w = np.concatenate([np.eye(3), [[0.1], [0.2], [3.0]]], 1)
    rd_w = oracle_mod.rays_of_view(H, W, K, c2w)[1].astype(np.float64)
    o_n, d_n, _ = oracle_mod.rays_of_view(H, W, K, c2w, ndc=True)
    np.testing.assert_allclose(o_n[..., 2], -1.0, atol=1e-6)
    np.testing.assert_allclose(o_n[..., 2] + d_n[..., 2], 1.0, atol=1e-6)
    f = K[0][0]

    def proj(p):
        return np.stack([-(2 * f / W) * p[..., 0] / p[..., 2], -(2 * f / H) * p[..., 1] / p[..., 2],
                         1 + 2 * 1.0 / p[..., 2]], -1)

    o_w = np.broadcast_to(c2w[:, 3], rd_w.shape)
    t = -(1.0 + o_w[..., 2]) / rd_w[..., 2]
    for s_ in (0.5, 4.0):
        q = proj(o_w + (t + s_)[..., None] * rd_w)  # a point beyond the near plane
        lam = (q[..., 2] - o_n[..., 2]) / d_n[..., 2]
        np.testing.assert_allclose(q, o_n + lam[..., None] * d_n, rtol=0, atol=1e-5)


def _pe27(vd):
    """Per-pixel PE of the renderer in the order get_rays writes it
    (renderer.cu:153-165): v, then sin(v_x 2^k), sin(v_y 2^k), sin(v_z 2^k),
    cos(...) for k = 0..3."""
    out = [vd[:, 0], vd[:, 1], vd[:, 2]]
    for fn in (np.sin, np.cos):
        for a in range(3):
            for k in range(4):
                out.append(fn(vd[:, a] * 2.0 ** k))
    return np.stack(out, 1)


@pytest.mark.parametrize("sigma,t_stop", [(6.0, 1e-3), (6.0, 1e-4), (5.0, 1e-4), (-1.0, 1e-3)])
def test_render_constant_field_closed_form(oracle_mod, sigma, t_stop):
    """All voxels active with one density and one colour: every in-box sample has
    the same alpha and every kept sample the colour c = sigmoid(MLP) of the
    constant feature and the pixel's PE.  Along a ray of n in-box samples
    (t_k = tmin + k*step, k >= 1, from the slab test) a sample is kept when its
    weight exceeds fast_color_thres, the march stops after a kept sample that
    leaves T < t_stop, and the pixel is c * sum(kept weights) + T * bg.  The
    cases cover a stop at 1e-3, a stop at 1e-4 (needs alpha > 0.5), weights that
    fall below the threshold before T does (no stop: T keeps falling to the box
    exit) and a thin medium."""
    from plenvdb_amd import scenes

    R, H, W = 24, 20, 22
    idx = np.arange(1, R ** 3 + 1, dtype=np.int32).reshape(R, R, R)
    den = np.full(R ** 3 + 1, sigma, np.float32)
    den[0] = 0
    rng = np.random.default_rng(5)
    cvec = rng.normal(0, 0.5, 12).astype(np.float32)
    col = np.tile(cvec, (R ** 3 + 1, 1)).astype(np.float32)
    col[0] = 0
    mlp = scenes.mlp_params(scenes.CONFIGS["c2"])
    mlp["b2"] = rng.normal(0, 0.3, 3).astype(np.float32)
    pose = scenes.orbit_poses(8, radius=4.0)[3]
    K = scenes.intrinsics(H, W, 14.0)  # wide enough that some rays miss the box
    mn, mx = np.full(3, -1.5, np.float32), np.full(3, 1.5, np.float32)
    shift, interval, stepdist, bg, thres = -4.595, 0.5, 3.0 / R * 0.5, 0.25, 1e-4
    img, info = oracle_mod.reference_render(idx, den, col, mlp, pose, K, mn, mx, H, W, 0.5,
                                            stepdist, shift, interval, thres, bg, t_stop=t_stop,
                                            prec="f64")
    alpha = 1.0 - (1.0 + np.exp(sigma + shift)) ** (-interval)
    # pixel rays and PE, independently of the oracle
    i, j = np.meshgrid(np.arange(W) + 0.5, np.arange(H) + 0.5)
    d = np.stack([(i - K[0, 2]) / K[0, 0], -(j - K[1, 2]) / K[1, 1], -np.ones_like(i)], -1)
    rd = d.reshape(-1, 3) @ pose[:3, :3].T.astype(np.float64)
    vd = rd / np.linalg.norm(rd, axis=1, keepdims=True)
    x = np.concatenate([np.tile(cvec.astype(np.float64), (H * W, 1)), _pe27(vd)], 1)
    h = np.maximum(x @ mlp["w0"] + mlp["b0"], 0)
    h = np.maximum(h @ mlp["w1"] + mlp["b1"], 0)
    c = 1 / (1 + np.exp(-(h @ mlp["w2"] + mlp["b2"])))
    # in-box sample count from the slab test
    o = (pose[:3, 3].astype(np.float64) - mn) / (mx - mn)
    dn = rd / (mx - mn)
    step = stepdist / np.linalg.norm(rd, axis=1)
    with np.errstate(divide="ignore"):
        ta, tb = (1 - o) / dn, -o / dn
    tmin = np.maximum(np.minimum(ta, tb).max(1), 0.5)
    tmax = np.maximum(np.maximum(ta, tb).min(1), 0.5)
    span = (tmax - tmin) / step
    n_in = np.maximum(np.floor(span + 1e-9), 0).astype(np.int64)
    clear = np.abs(span - np.round(span)) > 1e-3  # away from a sample on the box face
    want = np.zeros((H * W, 3))
    kept = np.zeros(H * W, np.int64)
    for p in range(H * W):
        T, S = 1.0, 0.0
        for _ in range(n_in[p]):
            w = T * alpha
            T = T * (1 - alpha)
            if w <= thres:
                continue
            S += w
            kept[p] += 1
            if T < t_stop:
                break
        want[p] = c[p] * S + T * bg
    assert (kept[clear] == 0).any() and (kept[clear] > 0).sum() > H * W // 4
    assert np.array_equal(info["count"][clear], kept[clear])
    assert np.abs(img.reshape(-1, 3)[clear] - want[clear]).max() < 1e-6


# ---- total-variation gradient: closed forms pin the dense restatement ----
def _tv(p, g, w, dense=True):
    from oracle import oracle

    return oracle.tv_add_grad_dense(p, g, *w, dense)


def test_tv_constant_grid_adds_nothing():
    p = np.full((2, 5, 6, 7), 0.37, np.float32)
    g = np.random.default_rng(0).normal(size=p.shape).astype(np.float32)
    np.testing.assert_array_equal(_tv(p, g, (1.0, 2.0, 3.0)), g)


@pytest.mark.parametrize("h", [0.3, -0.25, 5.0])
def test_tv_spike_closed_form(h):
    """One interior spike of height h on zeros: the spike gains 2 (wz + wy +
    wz)/6 clamp(h) and each face neighbour -w/6 clamp(h) with w = wz on the
    z and x axes (the reference's x-axis weight) and wy on y."""
    wx, wy, wz = 0.7, 1.1, 1.3
    p = np.zeros((1, 5, 5, 5), np.float32)
    p[0, 2, 2, 2] = h
    out = _tv(p, np.zeros_like(p), (wx, wy, wz))
    c = float(np.clip(h, -1, 1))
    exp = np.zeros_like(p)
    exp[0, 2, 2, 2] = (4 * wz + 2 * wy) / 6 * c
    for (dx, dy, dz), w in (((1, 0, 0), wz), ((0, 1, 0), wy), ((0, 0, 1), wz)):
        exp[0, 2 + dx, 2 + dy, 2 + dz] = -w / 6 * c
        exp[0, 2 - dx, 2 - dy, 2 - dz] = -w / 6 * c
    np.testing.assert_allclose(out, exp, rtol=1e-6, atol=1e-7)
    assert _tv(p, np.zeros_like(p), (99.0, wy, wz))[0, 3, 2, 2] == out[0, 3, 2, 2]  # wx unused


def test_tv_corner_spike_and_gate():
    """A corner spike has three neighbours; without dense mode only elements
    with a nonzero gradient change."""
    p = np.zeros((1, 3, 3, 3), np.float32)
    p[0, 0, 0, 0] = 0.5
    out = _tv(p, np.zeros_like(p), (1.0, 1.0, 1.0))
    assert abs(out[0, 0, 0, 0] - 3 * 0.5 / 6) < 1e-7
    assert abs(out[0, 1, 0, 0] + 0.5 / 6) < 1e-7 and out[0, 1, 1, 0] == 0
    g = np.zeros_like(p)
    g[0, 1, 0, 0] = 2.0
    gated = _tv(p, g, (1.0, 1.0, 1.0), dense=False)
    assert gated[0, 0, 0, 0] == 0 and abs(gated[0, 1, 0, 0] - (2.0 - 0.5 / 6)) < 1e-6
    assert np.count_nonzero(gated) == 1


def test_tv_brute_force_small():
    """Element loop over a small random grid (border, clamp and gate all hit)."""
    rng = np.random.default_rng(3)
    p = (rng.normal(size=(2, 3, 4, 5)) * 0.8).astype(np.float32)
    g = rng.normal(size=p.shape).astype(np.float32) * (rng.random(p.shape) < 0.5)
    w = (0.3, 0.6, 0.9)
    out = _tv(p, g, w, dense=False)
    exp = g.astype(np.float64).copy()
    for c, i, j, k in np.ndindex(*p.shape):
        if g[c, i, j, k] == 0:
            continue
        acc = 0.0
        for (di, dj, dk), ww in (((0, 0, 1), w[2]), ((0, 1, 0), w[1]), ((1, 0, 0), w[2])):
            for sgn in (-1, 1):
                ii, jj, kk = i + sgn * di, j + sgn * dj, k + sgn * dk
                if 0 <= ii < p.shape[1] and 0 <= jj < p.shape[2] and 0 <= kk < p.shape[3]:
                    acc += ww / 6 * np.clip(float(p[c, i, j, k]) - float(p[c, ii, jj, kk]), -1, 1)
        exp[c, i, j, k] += acc
    np.testing.assert_allclose(out, exp, rtol=1e-6, atol=1e-6)


def test_tv_vdb_view_matches_dense_on_full_cover():
    """A topology covering the box with leaves: the VDB form equals the dense
    form slot by slot (pins the slot <-> voxel mapping of the VDB form)."""
    from oracle import oracle

    R = (16, 8, 12)
    mask = np.ones(R, bool)
    origins, vmask = oracle.leaves_from_mask(mask)
    dom = oracle.SlotDomain(origins, vmask)
    rng = np.random.default_rng(5)
    vals = (rng.normal(size=(dom.nslots, 3)) * 0.7).astype(np.float32)
    grad = rng.normal(size=(dom.nslots, 3)).astype(np.float32)
    got = oracle.tv_add_grad_vdb(dom, vals, grad, R, 1.0, 0.5, 2.0, True)
    xyz = dom.coords()
    inside = np.all(xyz < np.array(R), axis=1)
    dense = np.zeros((3, *R), np.float32)
    gd = np.zeros((3, *R), np.float32)
    x, y, z = xyz[inside].T
    dense[:, x, y, z] = vals[inside].T
    gd[:, x, y, z] = grad[inside].T
    ref = oracle.tv_add_grad_dense(dense, gd, 1.0, 0.5, 2.0, True)
    np.testing.assert_array_equal(got[inside], ref[:, x, y, z].T)
    np.testing.assert_array_equal(got[~inside], grad[~inside])
