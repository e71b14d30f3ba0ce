"""Independent checks of the oracle functions that the other CPU tests leave
open: the world-to-index map, the dense (MLP) Adam in its three modes, the
VDB Adam in modes 0 and 2 and over several steps, and the merged renderer's
per-sample rules (box and active tests, corner cell, corners that read the
index-0 row).  Each check comes from a library routine, a closed form or an
invariant, chosen so that a dropped term, swapped hyper-parameters, a wrong
eps placement, a transposed axis or a wrong corner rule fails it."""
import numpy as np
import pytest
import torch

B0, B1 = 0.9, 0.99
F = lambda x: float(np.float32(x))  # the kernels take hyper-parameters as floats


# ---------------------------------------------------------------------------
# wld2idx (plenvdb/lib/grid.py:77-78)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_wld2idx_endpoints_affinity_and_axes(oracle_mod, prec):
    reso = np.array([11, 17, 23], np.int32)
    mn = np.array([-1.5, -0.5, 0.25], np.float32)
    mx = np.array([1.5, 2.0, 3.0], np.float32)
    # the box corners map to 0 and R-1 on every axis, exactly
    xs, ys, zs = oracle_mod.wld2idx(np.stack([mn, mx]), mn, mx, reso, prec)
    assert list(xs) == [0, 10] and list(ys) == [0, 16] and list(zs) == [0, 22]
    # the box centre maps to (R-1)/2
    xs, ys, zs = oracle_mod.wld2idx(((mn.astype(np.float64) + mx) / 2)[None], mn, mx, reso, prec)
    np.testing.assert_allclose([xs[0], ys[0], zs[0]], (reso - 1) / 2, rtol=1e-6)
    # affine: f(a p + (1-a) q) = a f(p) + (1-a) f(q)
    rng = np.random.default_rng(0)
    p = rng.uniform(-3, 4, (500, 3))
    q = rng.uniform(-3, 4, (500, 3))
    a = rng.uniform(0, 1, (500, 1))
    fp = np.stack(oracle_mod.wld2idx(p, mn, mx, reso, prec), 1).astype(np.float64)
    fq = np.stack(oracle_mod.wld2idx(q, mn, mx, reso, prec), 1).astype(np.float64)
    fm = np.stack(oracle_mod.wld2idx(a * p + (1 - a) * q, mn, mx, reso, prec), 1)
    np.testing.assert_allclose(fm, a * fp + (1 - a) * fq, rtol=0, atol=2e-5 * 23)
    # moving along one world axis moves only that index axis, at (R-1)/(max-min) per unit
    base = np.zeros((1, 3))
    f0 = np.stack(oracle_mod.wld2idx(base, mn, mx, reso, prec), 1)[0]
    for ax in range(3):
        d = base.copy()
        d[0, ax] = 0.5
        f1 = np.stack(oracle_mod.wld2idx(d, mn, mx, reso, prec), 1)[0]
        want = np.zeros(3)
        want[ax] = 0.5 * (reso[ax] - 1) / (float(mx[ax]) - float(mn[ax]))
        np.testing.assert_allclose(f1 - f0, want, atol=1e-5)


# ---------------------------------------------------------------------------
# Dense Adam (plenvdb/lib/cuda/adam_upd_kernel.cu:8-58)
# ---------------------------------------------------------------------------
def _torch_adam(p0, grads, lr):
    """torch.optim.Adam with eps=0 in FP64.  The reference's update,
    p -= lr*sqrt(1-b1^t)/(1-b0^t) * m / (sqrt(v) + eps), equals torch's
    bias-corrected update exactly when eps = 0."""
    p = torch.tensor(p0, dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([p], lr=lr, betas=(F(B0), F(B1)), eps=0.0)
    out = []
    for g in grads:
        opt.zero_grad()
        p.grad = torch.tensor(g, dtype=torch.float64)
        opt.step()
        out.append(p.detach().numpy().copy())
    return out


def _stepsz(oracle, lr, t):
    return oracle.adam_stepsz(lr, B0, B1, t)


def _stepsz64(lr, t):
    """The step size in FP64 from the float-valued betas; the kernels then
    round it to float once (relative error <= 6e-8)."""
    return lr * np.sqrt(1 - F(B1) ** t) / (1 - F(B0) ** t)


def test_adam_stepsz_closed_form(oracle_mod):
    for t in (1, 2, 7, 1000):
        want = 0.1 * np.sqrt(1 - F(B1) ** t) / (1 - F(B0) ** t)
        np.testing.assert_allclose(oracle_mod.adam_stepsz(0.1, B0, B1, t), want, rtol=2e-6)
    # beta1 (the square's rate) sits under the root: swapping the betas changes step 1
    assert abs(oracle_mod.adam_stepsz(0.1, B1, B0, 1) - oracle_mod.adam_stepsz(0.1, B0, B1, 1)) > 0.1


def test_dense_adam_mode0_equals_torch_adam_without_eps(oracle_mod):
    rng = np.random.default_rng(1)
    n = 301
    p0 = rng.normal(size=n)
    grads = [rng.normal(size=n) for _ in range(4)]
    want = _torch_adam(p0, grads, lr=1e-3)
    p, m, v = p0.copy(), np.zeros(n), np.zeros(n)
    prev = [p0] + want
    for t, g in enumerate(grads, 1):
        before = p.copy()
        oracle_mod.dense_adam(p, g.copy(), m, v, 0, _stepsz64(1e-3, t), 0.0, B0, B1, prec="f64")
        # each step's update (it depends on the moments only); the step size
        # is rounded to float once, hence 2e-7
        np.testing.assert_allclose(p - before, prev[t] - prev[t - 1], rtol=2e-7, atol=1e-15)


def test_dense_adam_eps_sits_beside_sqrt_v(oracle_mod):
    # first step from m = v = 0 with tiny gradients: the update is
    # s*(1-b0)*g / (sqrt(1-b1)*|g| + eps); eps beside the bias-corrected root
    # (torch's placement) would give s*(1-b0)*g / (sqrt(1-b1)*|g| + eps*sqrt(1-b1))
    g = np.array([1e-9, -3e-9, 2e-8, 5e-10])
    p, m, v = np.zeros(4), np.zeros(4), np.zeros(4)
    s = _stepsz(oracle_mod, 1e-3, 1)
    oracle_mod.dense_adam(p, g.copy(), m, v, 0, s, 1e-8, B0, B1, prec="f64")
    want = -F(s) * (1 - F(B0)) * g / (np.sqrt(1 - F(B1)) * np.abs(g) + F(1e-8))
    np.testing.assert_allclose(p, want, rtol=1e-6)
    torch_like = -F(s) * (1 - F(B0)) * g / (np.sqrt(1 - F(B1)) * (np.abs(g) + F(1e-8)))
    assert np.abs(p - torch_like).max() > 0.1 * np.abs(want).max()
    np.testing.assert_allclose(m, (1 - F(B0)) * g, rtol=1e-12)
    np.testing.assert_allclose(v, (1 - F(B1)) * g * g, rtol=1e-12)


def test_dense_adam_mode1_skips_zero_gradients(oracle_mod):
    rng = np.random.default_rng(2)
    n = 64
    p0 = rng.normal(size=n)
    g1, g2, g3 = (rng.normal(size=n) for _ in range(3))
    z = np.zeros(n, bool)
    z[::5] = True
    g2[z] = 0.0
    st = []
    p, m, v = p0.copy(), np.zeros(n), np.zeros(n)
    for t, g in enumerate((g1, g2, g3), 1):
        oracle_mod.dense_adam(p, g.copy(), m, v, 1, _stepsz(oracle_mod, 1e-3, t), 1e-8, B0, B1,
                              prec="f64")
        st.append((p.copy(), m.copy(), v.copy()))
    # mode 0 on the same gradients
    q, m0, v0 = p0.copy(), np.zeros(n), np.zeros(n)
    for t, g in enumerate((g1, g2, g3), 1):
        oracle_mod.dense_adam(q, g.copy(), m0, v0, 0, _stepsz(oracle_mod, 1e-3, t), 1e-8, B0, B1,
                              prec="f64")
    # entries that never saw a zero gradient follow mode 0 exactly
    np.testing.assert_array_equal(st[2][0][~z], q[~z])
    # a zero gradient leaves parameter and both moments untouched at that step ...
    for k in range(3):
        np.testing.assert_array_equal(st[1][k][z], st[0][k][z])
    # ... while mode 0 decays the moments and still moves the parameter
    assert np.all(q[z] != st[2][0][z])
    # step 3 of a skipped entry is one Adam step from the step-1 state with s_3
    p1, m1, v1 = st[0]
    s3 = F(_stepsz(oracle_mod, 1e-3, 3))
    mm = F(B0) * m1[z] + (1 - F(B0)) * g3[z]
    vv = F(B1) * v1[z] + (1 - F(B1)) * g3[z] * g3[z]
    np.testing.assert_allclose(st[2][0][z], p1[z] - s3 * mm / (np.sqrt(vv) + F(1e-8)),
                               rtol=1e-12)


def test_dense_adam_mode2_scales_every_step_by_the_per_element_rate(oracle_mod):
    rng = np.random.default_rng(3)
    n = 97
    p0 = rng.normal(size=n)
    per = rng.uniform(0.25, 4.0, n)
    a, b = p0.copy(), p0.copy()
    ma, va, mb, vb = (np.zeros(n) for _ in range(4))
    for t in range(1, 4):
        g = rng.normal(size=n)
        g[t::9] = 0.0
        pa, pb = a.copy(), b.copy()
        s = _stepsz(oracle_mod, 1e-2, t)
        oracle_mod.dense_adam(a, g.copy(), ma, va, 0, s, 1e-8, B0, B1, prec="f64")
        oracle_mod.dense_adam(b, g.copy(), mb, vb, 2, s, 1e-8, B0, B1, per_lr=per, prec="f64")
        # the moments do not depend on the rate; each step is per * the mode-0 step
        np.testing.assert_array_equal(ma, mb)
        np.testing.assert_array_equal(va, vb)
        np.testing.assert_allclose(b - pb, per * (a - pa), rtol=1e-9, atol=1e-14)


# ---------------------------------------------------------------------------
# VDB Adam (plenvdb/lib/vdb/densityvdb.cu:185-270, colorvdb.cu:179-371)
# ---------------------------------------------------------------------------
def _vdb_state(seed, nslots, C):
    rng = np.random.default_rng(seed)
    p0 = rng.normal(size=(nslots, C))
    act = (rng.random(nslots) < 0.8).astype(np.uint8)
    per = rng.uniform(0.25, 4.0, nslots)
    grads = [rng.normal(size=(nslots, C)) for _ in range(3)]
    return p0, act, per, grads


@pytest.mark.parametrize("C,group", [(1, 1), (12, 3)])
def test_vdb_adam_mode0_equals_torch_adam_on_active_slots(oracle_mod, C, group):
    p0, act, _, grads = _vdb_state(4, 200, C)
    act[7] = 1
    grads[1][7] = 0.0  # mode 0 updates a group whose gradient is all zero
    want = _torch_adam(p0, grads, lr=0.1)
    p, m, v = p0.copy(), np.zeros_like(p0), np.zeros_like(p0)
    on = act.astype(bool)
    prev = [p0] + want
    for t, g in enumerate(grads, 1):
        before = p.copy()
        oracle_mod.vdb_adam(p, g.copy(), m, v, act, group, 0, _stepsz64(0.1, t), 0.0, B0, B1,
                            prec="f64")
        d, w = p - before, prev[t] - prev[t - 1]
        np.testing.assert_allclose(d[on], w[on], rtol=2e-7, atol=1e-15)
        np.testing.assert_array_equal(p[~on], p0[~on])  # inactive slots are never touched
        np.testing.assert_array_equal(m[~on], 0)
        if t == 2:
            assert np.all(d[7] != 0)


@pytest.mark.parametrize("C,group", [(1, 1), (12, 3)])
def test_vdb_adam_mode2_per_slot_rate(oracle_mod, C, group):
    p0, act, per, grads = _vdb_state(5, 150, C)
    a, b = p0.copy(), p0.copy()
    ma, va, mb, vb = (np.zeros_like(p0) for _ in range(4))
    for t, g in enumerate(grads, 1):
        pa, pb = a.copy(), b.copy()
        s = _stepsz(oracle_mod, 0.1, t)
        oracle_mod.vdb_adam(a, g.copy(), ma, va, act, group, 0, s, 1e-8, B0, B1, prec="f64")
        oracle_mod.vdb_adam(b, g.copy(), mb, vb, act, group, 2, s, 1e-8, B0, B1, per_lr=per,
                            prec="f64")
        np.testing.assert_array_equal(ma, mb)
        # one rate per slot, shared by all its channels
        np.testing.assert_allclose(b - pb, per[:, None] * (a - pa), rtol=1e-9, atol=1e-14)


def test_vdb_adam_mode1_group_skip_over_steps(oracle_mod):
    # density: a scalar is skipped when its gradient is 0; colour: a Vec3 group
    # only when all three components are 0 (a partly-zero group is updated)
    p0, act, _, grads = _vdb_state(6, 120, 12)
    act[:] = 1
    g2 = grads[1]
    g2[10, 0:3] = 0.0        # whole group 0 of slot 10: skipped at step 2
    g2[11, 3] = 0.0          # one component of group 1 of slot 11: updated
    p, m, v = p0.copy(), np.zeros_like(p0), np.zeros_like(p0)
    q, mq, vq = p0.copy(), np.zeros_like(p0), np.zeros_like(p0)
    snaps = []
    for t, g in enumerate(grads, 1):
        s = _stepsz(oracle_mod, 0.1, t)
        oracle_mod.vdb_adam(p, g.copy(), m, v, act, 3, 1, s, 1e-8, B0, B1, prec="f64")
        oracle_mod.vdb_adam(q, g.copy(), mq, vq, act, 3, 0, s, 1e-8, B0, B1, prec="f64")
        snaps.append((p.copy(), m.copy(), v.copy()))
    for k in range(3):
        np.testing.assert_array_equal(snaps[1][k][10, 0:3], snaps[0][k][10, 0:3])
    assert np.all(snaps[1][1][11, 3:6] != snaps[0][1][11, 3:6])
    np.testing.assert_array_equal(snaps[1][1][11, 3], F(B0) * snaps[0][1][11, 3])
    untouched = np.ones((120, 12), bool)
    untouched[10, 0:3] = False
    np.testing.assert_array_equal(p[untouched], q[untouched])
    # density (group 1): the scalar rule
    d0 = p0[:, :1].copy()
    gd = [g[:, :1].copy() for g in grads]
    gd[1][20, 0] = 0.0
    d, md, vd = d0.copy(), np.zeros_like(d0), np.zeros_like(d0)
    for t, g in enumerate(gd, 1):
        before = (d[20].copy(), md[20].copy())
        oracle_mod.vdb_adam(d, g, md, vd, act, 1, 1, _stepsz(oracle_mod, 0.1, t), 1e-8, B0, B1,
                            prec="f64")
        if t == 2:
            assert d[20] == before[0] and md[20] == before[1]


# ---------------------------------------------------------------------------
# Merged renderer per-sample rules (renderer.cu:222-367, Appendix A2-A5)
# ---------------------------------------------------------------------------
def _render_scene(seed=11, R=16):
    rng = np.random.default_rng(seed)
    act = rng.random((R, R, R)) < 0.7
    idx = np.zeros((R, R, R), np.int32)
    n = int(act.sum())
    idx[act] = np.arange(1, n + 1)
    den = rng.uniform(0.0, 7.0, n + 1).astype(np.float32)
    col = rng.normal(0, 1.0, (n + 1, 12)).astype(np.float32)
    # row 0 (read by inactive and out-of-grid corners) gets distinctive values;
    # the merged export stores zeros there, which would hide the rule
    den[0] = 7.5
    col[0] = np.linspace(-2, 2, 12)
    return idx, den, col


def _grid_sample(field, xyz, R, pad):
    """Trilinear interpolation of a voxel-centred field [R,R,R,C] at index-space
    points with F.grid_sample(align_corners=True) in FP64.  The field is padded
    by one voxel of `pad` on the high side, so the +1 corners of points within
    one voxel of R-1 read `pad` as the renderer's out-of-grid corners do."""
    C = field.shape[-1]
    f = np.empty((R + 1, R + 1, R + 1, C))
    f[:] = pad
    f[:R, :R, :R] = field
    g = torch.tensor(f).permute(3, 0, 1, 2)[None]
    ind = torch.tensor(np.asarray(xyz, np.float64) / R * 2 - 1).flip(-1).reshape(1, 1, 1, -1, 3)
    out = torch.nn.functional.grid_sample(g, ind, mode="bilinear", align_corners=True)
    return out.reshape(C, -1).T.numpy()


def test_render_samples_follow_the_corner_and_activity_rules(oracle_mod):
    from plenvdb_amd import scenes

    R, H, W = 16, 18, 21
    idx, den, col = _render_scene(R=R)
    mlp = scenes.mlp_params(scenes.CONFIGS["c2"])
    pose = scenes.orbit_poses(8, radius=4.0)[2]
    K = scenes.intrinsics(H, W, 16.0)
    mn, mx = np.full(3, -1.5, np.float32), np.full(3, 1.5, np.float32)
    shift, interval, stepdist, bg = -4.595, 0.5, 3.0 / R * 0.5, 0.3
    img, info = oracle_mod.reference_render(idx, den, col, mlp, pose, K, mn, mx, H, W, 0.5,
                                            stepdist, shift, interval, 0.0, bg, t_stop=1e-3,
                                            prec="f64")
    xyz = info["xyz"]
    assert len(xyz) > 500
    # every sample lies in [0, R-1]^3 and its rint-rounded voxel is active
    assert np.all((xyz >= 0) & (xyz <= R - 1))
    r = np.rint(xyz).astype(int)
    assert np.all(idx[r[:, 0], r[:, 1], r[:, 2]] != 0)
    # corners read the voxel's merged row, inactive voxels and out-of-grid
    # corners read row 0: the interpolant of that field by F.grid_sample
    act = idx != 0
    dfield = np.where(act, den[idx], den[0])[..., None].astype(np.float64)
    cfield = np.where(act[..., None], col[idx], col[0]).astype(np.float64)
    sig = _grid_sample(dfield, xyz, R, den[0])[:, 0]
    feat = _grid_sample(cfield, xyz, R, col[0])
    np.testing.assert_allclose(info["feat"], feat, rtol=0, atol=1e-9)
    e = np.exp(sig + F(shift))
    np.testing.assert_allclose(info["alpha"], 1 - (1 + e) ** (-interval), rtol=1e-9, atol=1e-12)
    # the rule matters here: many samples have inactive corners
    w_inact = 1.0 - _grid_sample(act[..., None].astype(np.float64), xyz, R, 0.0)[:, 0]
    assert (w_inact > 0.1).sum() > 100
    # with fast_color_thres = 0 every alpha-passing sample is kept, so the
    # weights are the alpha product down each ray, and the pixel is their
    # sum with the MLP colours plus T * bg
    pix = info["pix"]
    a = info["alpha"].astype(np.float64)
    wsum = np.zeros(H * W)
    T = np.ones(H * W)
    T_before_last = np.ones(H * W)
    for k in range(len(a)):
        p = pix[k]
        np.testing.assert_allclose(info["weight"][k], T[p] * a[k], rtol=1e-9)
        wsum[p] += info["weight"][k]
        T_before_last[p] = T[p]
        T[p] *= 1 - a[k]
    np.testing.assert_allclose(info["t_last"], T, rtol=1e-9)
    # the early stop: a ray ends at its first sample that leaves T < 1e-3
    assert np.all(T_before_last >= 1e-3) and (T < 1e-3).sum() > 20
    # consecutive samples of a ray sit a whole number of steps apart, on the
    # line through the camera centre; one step is stepdist*(R-1)/(max-min)
    # index units here (equal box sides and resolutions)
    L = stepdist * (R - 1) / 3.0
    o = (pose[:3, 3].astype(np.float64) - mn) / (mx - mn) * (R - 1)
    for p in np.unique(pix)[:80]:
        s = xyz[pix == p]
        d = np.linalg.norm(np.diff(s, axis=0), axis=1) / L
        np.testing.assert_allclose(d, np.rint(d), atol=1e-3)
        assert np.all(np.rint(d) >= 1)
        u = (s - o) / np.linalg.norm(s - o, axis=1, keepdims=True)
        np.testing.assert_allclose(u, np.broadcast_to(u[0], u.shape), atol=1e-5)


def test_render_inactive_corner_rule_changes_the_image(oracle_mod):
    """Row 0 is what inactive corners read: changing it changes the samples next
    to inactive voxels and nothing else."""
    from plenvdb_amd import scenes

    R, H, W = 16, 14, 15
    idx, den, col = _render_scene(seed=12, R=R)
    mlp = scenes.mlp_params(scenes.CONFIGS["c2"])
    pose = scenes.orbit_poses(8, radius=4.0)[5]
    K = scenes.intrinsics(H, W, 12.0)
    mn, mx = np.full(3, -1.5, np.float32), np.full(3, 1.5, np.float32)
    args = (mlp, pose, K, mn, mx, H, W, 0.5, 3.0 / R * 0.5, -4.595, 0.5, 0.0, 0.2)
    _, i1 = oracle_mod.reference_render(idx, den, col, *args, t_stop=1e-3, prec="f64")
    col2 = col.copy()
    col2[0] += 1.0
    _, i2 = oracle_mod.reference_render(idx, den, col2, *args, t_stop=1e-3, prec="f64")
    # same samples (density unchanged); feature change = weight of row-0 corners
    np.testing.assert_array_equal(i1["pix"], i2["pix"])
    xyz = i1["xyz"]
    act = (idx != 0).astype(np.float64)[..., None]
    w_row0 = 1.0 - _grid_sample(act, xyz, R, 0.0)[:, 0]
    delta = col2[0].astype(np.float64) - col[0]  # 1 up to the float rounding of row 0
    np.testing.assert_allclose(i2["feat"] - i1["feat"], w_row0[:, None] * delta[None, :],
                               atol=1e-9)
    assert (w_row0 > 0.05).sum() > 20 and (w_row0 == 0).sum() > 20
