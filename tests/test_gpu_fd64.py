"""Central finite differences in FP64 on the device for the hand-written
adjoint pairs of the plain path, through their FP64 twins (fd64.hip):
the trilinear gather / "leaf must exist" scatter (density and 12-channel
colour), raw2alpha and Alphas2Weights.  Each check perturbs inputs on the
device, evaluates the forward kernel twice per entry and compares
(L(x+h) - L(x-h)) / 2h with the adjoint's gradient of L = <forward(x), G>."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _dev(a, dtype=torch.float64):
    return torch.as_tensor(np.asarray(a)).to(dtype).cuda().contiguous()


@pytest.mark.parametrize("C", [1, 12])
def test_trilinear_adjoint_fd64(C):
    from plenvdb_amd import lib, scenes
    from plenvdb_amd.topology import Topology

    print("[case] trilinear fd64, channels", C, flush=True)
    sc = scenes.make_scene("c1", n_rays=8)
    topo = Topology.from_mask(sc["mask"])
    R = sc["reso"][0]
    rng = np.random.default_rng(C)
    vals = _dev(rng.normal(size=(topo.nslots, C)))
    n = 8192
    pts = rng.uniform(-1.0, R, size=(n, 3))  # also outside the grid and in missing leaves
    xs, ys, zs = (_dev(pts[:, a]) for a in range(3))
    G = _dev(rng.normal(size=(n, C)))
    out = torch.empty(n, C, dtype=torch.float64, device="cuda")

    def loss(v):
        lib.call("pvdb_trilinear_fwd_f64", topo.handle, lib.ptr(v), C, lib.ptr(xs), lib.ptr(ys),
                 lib.ptr(zs), n, lib.ptr(out), lib.stream())
        return float((out * G).sum())

    grad = torch.zeros_like(vals)
    lib.call("pvdb_trilinear_bwd_f64", topo.handle, lib.ptr(grad), C, lib.ptr(xs), lib.ptr(ys),
             lib.ptr(zs), lib.ptr(G), n, lib.stream())
    g = grad.cpu().numpy()
    nz = np.argwhere(g != 0)
    assert len(nz) > 100
    h = 1e-3
    worst = 0.0
    for s, c in nz[rng.choice(len(nz), 96, replace=False)]:
        vals[s, c] += h
        lp = loss(vals)
        vals[s, c] -= 2 * h
        lm = loss(vals)
        vals[s, c] += h
        fd = (lp - lm) / (2 * h)
        # L is linear in the values: no truncation error; the sums' rounding
        # (~1e5 terms, ~3e-14) over 2h gives ~2e-11
        assert abs(fd - g[s, c]) <= 1e-9 * abs(fd) + 1e-10, (s, c, fd, g[s, c])
        worst = max(worst, abs(fd - g[s, c]) / max(abs(fd), 1e-3))
    # a slot no point reaches has zero gradient and no effect on the loss
    z = np.argwhere(g == 0)[0]
    l0 = loss(vals)
    vals[z[0], z[1]] += 1.0
    assert loss(vals) == l0
    print(f"[result] trilinear C={C}: worst relative FD error {worst:.2e}", flush=True)


def test_raw2alpha_adjoint_fd64():
    from plenvdb_amd import lib

    print("[case] raw2alpha fd64", flush=True)
    rng = np.random.default_rng(5)
    n = 4096
    sigma = _dev(rng.normal(0, 3, n))
    G = _dev(rng.normal(size=n))
    shift, interval = -4.595119953155518, 0.5
    e = torch.empty_like(sigma)
    a = torch.empty_like(sigma)

    def alpha(s):
        lib.call("pvdb_raw2alpha_f64", lib.ptr(s), n, shift, interval, lib.ptr(e), lib.ptr(a),
                 lib.stream())
        return a.clone()

    alpha(sigma)
    gs = torch.empty_like(sigma)
    lib.call("pvdb_raw2alpha_bwd_f64", lib.ptr(e), lib.ptr(G), n, interval, lib.ptr(gs),
             lib.stream())
    h = 1e-5
    fd = (alpha(sigma + h) - alpha(sigma - h)) / (2 * h) * G  # elementwise: all entries at once
    # alpha = 1 - (1+e)^-interval carries ~1e-16 absolute rounding (cancellation
    # for small e), which the quotient turns into ~eps/h |G|: allow 1e-9 |G|
    tol = 1e-6 * gs.abs() + 1e-9 * G.abs()
    assert bool(((fd - gs).abs() <= tol).all())
    err = ((fd - gs).abs() / gs.abs().clamp_min(1e-30))[gs.abs() > 1e-3].max().item()
    print(f"[result] raw2alpha: worst relative FD error where |g| > 1e-3: {err:.2e}", flush=True)


def test_alpha2weight_adjoint_fd64():
    from plenvdb_amd import lib

    print("[case] alpha2weight fd64", flush=True)
    rng = np.random.default_rng(6)
    n_rays = 64
    counts = rng.integers(1, 60, n_rays)
    counts[:4] = [1, 2, 3, 0]  # short rays and an empty one
    incl = np.cumsum(counts)
    i_start = _dev(incl - counts, torch.int64)
    i_end0 = _dev(incl, torch.int64)
    M = int(incl[-1])
    alpha = _dev(rng.uniform(0.0, 0.25, M))
    gw = _dev(rng.normal(size=M))
    glast = _dev(rng.normal(size=n_rays))
    w = torch.zeros_like(alpha)
    T = torch.zeros_like(alpha)
    tl = torch.empty(n_rays, dtype=torch.float64, device="cuda")
    i_end = torch.empty_like(i_end0)

    def loss(al):
        w.zero_()
        T.zero_()
        i_end.copy_(i_end0)
        lib.call("pvdb_alpha2weight_f64", lib.ptr(al), n_rays, lib.ptr(w), lib.ptr(T), lib.ptr(tl),
                 lib.ptr(i_start), lib.ptr(i_end), lib.stream())
        return float((w * gw).sum() + (tl * glast).sum())

    loss(alpha)
    stops = (i_end != i_end0).sum().item()
    ga = torch.zeros_like(alpha)
    lib.call("pvdb_alpha2weight_bwd_f64", lib.ptr(alpha), lib.ptr(w), lib.ptr(T), lib.ptr(tl),
             lib.ptr(i_start), lib.ptr(i_end), n_rays, lib.ptr(gw), lib.ptr(glast), lib.ptr(ga),
             lib.stream())
    g = ga.cpu().numpy()
    h = 1e-5
    worst = 0.0
    for i in rng.choice(M, 128, replace=False):
        alpha[i] += h
        lp = loss(alpha)
        alpha[i] -= 2 * h
        lm = loss(alpha)
        alpha[i] += h
        fd = (lp - lm) / (2 * h)
        # truncation (h^2 times the third derivative / 6, ~2e-10) plus the sums'
        # rounding over 2h (~2.5e-10)
        assert abs(fd - g[i]) <= 1e-6 * abs(fd) + 2e-9, (i, fd, g[i])
        worst = max(worst, abs(fd - g[i]) / max(abs(fd), 1e-3))
    print(f"[result] alpha2weight: {stops} rays stop early; worst relative FD error {worst:.2e}",
          flush=True)
    assert stops > 0


def test_fp32_kernels_track_their_fp64_twins():
    """The FP32 kernels the plain path runs agree with the FD-checked FP64 twins
    on the same inputs (to FP32 rounding), forward and backward."""
    from plenvdb_amd import lib, scenes
    from plenvdb_amd.topology import Topology

    print("[case] fp32 kernels vs fp64 twins", flush=True)
    sc = scenes.make_scene("c1", n_rays=8)
    topo = Topology.from_mask(sc["mask"])
    R = sc["reso"][0]
    rng = np.random.default_rng(9)
    n = 8192
    pts = rng.uniform(-1.0, R, size=(n, 3)).astype(np.float32)
    for C, order in ((1, 0), (12, 1)):
        v32 = _dev(rng.normal(size=(topo.nslots, C)), torch.float32)
        G32 = _dev(rng.normal(size=(n, C)), torch.float32)
        x32 = [_dev(pts[:, a], torch.float32) for a in range(3)]
        x64 = [t.double() for t in x32]
        o32 = torch.empty(n, C, device="cuda")
        o64 = torch.empty(n, C, dtype=torch.float64, device="cuda")
        lib.call("pvdb_trilinear_fwd", topo.handle, lib.ptr(v32), C, order, *map(lib.ptr, x32), n,
                 lib.ptr(o32), lib.stream())
        lib.call("pvdb_trilinear_fwd_f64", topo.handle, lib.ptr(v32.double()), C,
                 *map(lib.ptr, x64), n, lib.ptr(o64), lib.stream())
        torch.testing.assert_close(o32.double(), o64, rtol=1e-5, atol=1e-5)
        g32 = torch.zeros_like(v32)
        g64 = torch.zeros(topo.nslots, C, dtype=torch.float64, device="cuda")
        lib.call("pvdb_trilinear_bwd", topo.handle, lib.ptr(g32), C, order, *map(lib.ptr, x32),
                 lib.ptr(G32), n, lib.stream())
        lib.call("pvdb_trilinear_bwd_f64", topo.handle, lib.ptr(g64), C, *map(lib.ptr, x64),
                 lib.ptr(G32.double()), n, lib.stream())
        torch.testing.assert_close(g32.double(), g64, rtol=1e-5, atol=1e-5)
    sigma = _dev(rng.normal(0, 3, n), torch.float32)
    e32, a32 = torch.empty_like(sigma), torch.empty_like(sigma)
    e64 = torch.empty(n, dtype=torch.float64, device="cuda")
    a64 = torch.empty_like(e64)
    shift, interval = -4.595119953155518, 0.5
    lib.call("pvdb_raw2alpha", lib.ptr(sigma), n, shift, interval, lib.ptr(e32), lib.ptr(a32),
             lib.stream())
    lib.call("pvdb_raw2alpha_f64", lib.ptr(sigma.double()), n, float(np.float32(shift)), interval,
             lib.ptr(e64), lib.ptr(a64), lib.stream())
    torch.testing.assert_close(a32.double(), a64, rtol=1e-5, atol=1e-7)
    ga = _dev(rng.normal(size=n), torch.float32)
    gs32 = torch.empty_like(sigma)
    gs64 = torch.empty_like(e64)
    lib.call("pvdb_raw2alpha_bwd", lib.ptr(e32), lib.ptr(ga), n, interval, lib.ptr(gs32),
             lib.stream())
    lib.call("pvdb_raw2alpha_bwd_f64", lib.ptr(e64), lib.ptr(ga.double()), n, interval,
             lib.ptr(gs64), lib.stream())
    torch.testing.assert_close(gs32.double(), gs64, rtol=1e-5, atol=1e-7)
