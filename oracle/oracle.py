"""CPU oracle of the PlenVDB hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline may import
this module, and only as the checker.  The product (plenvdb_amd/) never
imports it.

Pieces
  * NanoVDB leaf order and slot numbering restated in numpy
    (openvdb/nanovdb/nanovdb/util/GridBuilder.h:362-376, NanoVDB.h:932-935,
    3272-3283, 3893-3900).
  * ctypes wrappers of liboracle.so (pvdb_oracle_impl.h / pvdb_oracle_render.h:
    C restatements of the reference kernels in FP32 reference order and FP64).
  * `reference_train_step`: the reference training step composed from those
    pieces (plenvdb/lib/dvgo.py:296-388 forward, run.py:551-576 MSE loss and
    backward), with the MLP, composite and loss in torch CPU autograd (FP32 or
    FP64) and every custom CUDA op restated in C.
  * `reference_render`: the merged-VDB renderer (renderer.cu) with the MLP and
    composite in numpy.

Pinning (see tests/test_oracle_*.py): leaf order/slots/getValue/isActive/
trilinear against golden vectors from NanoVDB itself (tests/golden/), trilinear
against exact reproduction of linear fields and F.grid_sample, scatter by the
adjoint identity, alpha2weight by closed forms, Adam by a closed form.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_BIN = os.path.join(HERE, "_ref", "nanovdb_golden")

_lib = None


def build(with_ref: bool | None = None) -> None:
    """Compiles liboracle.so and, when the reference tree exists, oracle/_ref."""
    subprocess.run(["make", "-s", "-C", HERE, "-B", "all"], check=True)
    if with_ref is None:
        with_ref = os.path.isdir("/root/reference/openvdb/nanovdb")
    if with_ref:
        subprocess.run(["make", "-s", "-C", HERE, "-B", "ref"], check=True)


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build(with_ref=False)
        _lib = C.CDLL(LIB_PATH)
    return _lib


def set_threads(n: int) -> int:
    """Threads of the oracle's OpenMP loops (and of torch's CPU ops) from here
    on; returns the OpenMP maximum in effect."""
    import torch

    torch.set_num_threads(int(n))
    fn = load().pvdb_oracle_set_threads
    fn.argtypes = [C.c_int]
    fn.restype = C.c_int
    return int(fn(int(n)))


def _p(a):
    return C.c_void_p(None) if a is None else C.c_void_p(a.ctypes.data)


def _f3(v):
    a = np.ascontiguousarray(np.asarray(v, dtype=np.float32).reshape(3))
    return a


def _dt(prec):
    return np.float32 if prec == "f32" else np.float64


# ----------------------------------------------------------------------------
# Topology (NanoVDB order) -- numpy restatement
# ----------------------------------------------------------------------------
def nanovdb_leaf_order(origins: np.ndarray) -> np.ndarray:
    """Permutation sorting leaf origins into NanoVDB's depth-first order:
    upper nodes (4096^3) by lexicographic origin, lower nodes by
    ((x>>7)&31)<<10|((y>>7)&31)<<5|((z>>7)&31), leaves by ((x>>3)&15)<<8|..."""
    o = np.asarray(origins, dtype=np.int64).reshape(-1, 3)
    up = o & ~np.int64(4095)
    low = (((o[:, 0] & 4095) >> 7) << 10) | (((o[:, 1] & 4095) >> 7) << 5) | ((o[:, 2] & 4095) >> 7)
    leaf = (((o[:, 0] & 127) >> 3) << 8) | (((o[:, 1] & 127) >> 3) << 4) | ((o[:, 2] & 127) >> 3)
    return np.lexsort((leaf, low, up[:, 2], up[:, 1], up[:, 0]))


def leaves_from_mask(mask: np.ndarray):
    """Leaf origins (NanoVDB order) and [L,512] bool value masks of a dense mask."""
    m = np.asarray(mask).astype(bool)
    X, Y, Z = m.shape
    px, py, pz = -X % 8, -Y % 8, -Z % 8
    mp = np.pad(m, ((0, px), (0, py), (0, pz)))
    lx, ly, lz = mp.shape[0] // 8, mp.shape[1] // 8, mp.shape[2] // 8
    blocks = mp.reshape(lx, 8, ly, 8, lz, 8).transpose(0, 2, 4, 1, 3, 5).reshape(lx, ly, lz, 512)
    on = blocks.any(axis=-1)
    idx = np.argwhere(on)
    origins = (idx * 8).astype(np.int64)
    vm = blocks[on]
    order = nanovdb_leaf_order(origins)
    return origins[order], vm[order]


def slot_of(origins: np.ndarray, ijk: np.ndarray) -> np.ndarray:
    """leaf*512 + ((x&7)<<6|(y&7)<<3|(z&7)) or -1 (LeafNode::CoordToOffset)."""
    o = np.asarray(origins, dtype=np.int64)
    q = np.asarray(ijk, dtype=np.int64).reshape(-1, 3)
    key = {tuple(r): i for i, r in enumerate(o.tolist())}
    base = q & ~np.int64(7)
    off = ((q[:, 0] & 7) << 6) | ((q[:, 1] & 7) << 3) | (q[:, 2] & 7)
    leaf = np.array([key.get(tuple(b), -1) for b in base.tolist()], dtype=np.int64)
    return np.where(leaf >= 0, leaf * 512 + off, -1)


class SlotDomain:
    """Leaf set in a given order with slot-major payloads (slot = leaf*512 +
    in-leaf offset); leaves are found by binary search on packed origin keys."""

    def __init__(self, origins: np.ndarray, vmask: np.ndarray):
        self.origins = np.asarray(origins, dtype=np.int64).reshape(-1, 3)
        self.vmask = np.asarray(vmask, dtype=bool).reshape(-1, 512)
        lk = (self.origins >> 3) + (1 << 20)
        key = (lk[:, 0] << 42) | (lk[:, 1] << 21) | lk[:, 2]
        order = np.argsort(key, kind="stable")
        self.keys = np.ascontiguousarray(key[order])
        self.kleaf = np.ascontiguousarray(order.astype(np.int64))
        off = np.arange(512)
        self._local = np.stack([off >> 6, (off >> 3) & 7, off & 7], axis=1)

    @property
    def nleaf(self):
        return len(self.origins)

    @property
    def nslots(self):
        return len(self.origins) * 512

    def coords(self) -> np.ndarray:
        """[nslots,3] voxel coordinates of every slot."""
        return (self.origins[:, None, :] + self._local[None]).reshape(-1, 3)

    def active(self) -> np.ndarray:
        return self.vmask.reshape(-1)



# ----------------------------------------------------------------------------
# C kernels
# ----------------------------------------------------------------------------
def trilinear_fwd(dom: SlotDomain, slot_vals, xs, ys, zs, order: int, prec: str = "f32"):
    lib = load()
    dt = _dt(prec)
    sv = np.ascontiguousarray(np.asarray(slot_vals, dtype=dt).reshape(dom.nslots, -1))
    C_ = sv.shape[1]
    xs, ys, zs = (np.ascontiguousarray(np.asarray(a, dtype=np.float32)) for a in (xs, ys, zs))
    n = len(xs)
    out = np.zeros((n, C_), dtype=dt)
    fn = getattr(lib, f"pvdb_trilinear_fwd_{prec}")
    fn(_p(sv), _p(dom.keys), _p(dom.kleaf), C.c_long(dom.nleaf), C.c_int(C_), C.c_int(order),
       _p(xs), _p(ys), _p(zs), C.c_long(n), _p(out))
    return out


def trilinear_bwd(dom: SlotDomain, C_: int, xs, ys, zs, g, order: int, prec: str = "f32"):
    """Returns the slot gradient [nslots, C]."""
    lib = load()
    dt = _dt(prec)
    xs, ys, zs = (np.ascontiguousarray(np.asarray(a, dtype=np.float32)) for a in (xs, ys, zs))
    n = len(xs)
    gg = np.ascontiguousarray(np.asarray(g, dtype=dt).reshape(n, C_))
    grad = np.zeros((dom.nslots, C_), dtype=dt)
    fn = getattr(lib, f"pvdb_trilinear_bwd_{prec}")
    fn(_p(grad), _p(dom.keys), _p(dom.kleaf), C.c_long(dom.nleaf), C.c_int(C_), C.c_int(order),
       _p(xs), _p(ys), _p(zs), _p(gg), C.c_long(n))
    return grad


def rays_of_view(H, W, K, c2w, ndc=False, inverse_y=False, flip_x=False, flip_y=False,
                 mode="center", jitter=None):
    """get_rays_of_a_view (dvgo.py:522-527) with get_rays (:471-499) and ndc_rays
    (:502-519, near 1, focal K[0][0]) in float32, one numpy operation per torch
    operation: K's entries and the NDC constants are rounded to float32 where
    torch takes them as scalars.  Returns rays_o, rays_d, viewdirs [H, W, 3]."""
    f = np.float32
    K = np.asarray(K, np.float64)
    c2w = np.asarray(c2w, np.float32)
    i, j = np.meshgrid(np.arange(W, dtype=f), np.arange(H, dtype=f))  # [H, W]: i = column
    if mode == "center":
        i, j = i + f(0.5), j + f(0.5)
    elif mode == "random":  # i + rand_like(i), j + rand_like(j) (dvgo.py:482-484)
        jit = np.asarray(jitter, f).reshape(2, H, W)
        i, j = i + jit[0], j + jit[1]
    elif mode != "lefttop":
        raise NotImplementedError(mode)
    if flip_x:
        i = i[:, ::-1]
    if flip_y:
        j = j[::-1, :]
    d0 = (i - f(K[0][2])) / f(K[0][0])
    if inverse_y:
        d1 = (j - f(K[1][2])) / f(K[1][1])
        d2 = np.ones_like(i)
    else:
        d1 = -((j - f(K[1][2])) / f(K[1][1]))
        d2 = -np.ones_like(i)
    R = c2w[:3, :3]
    rd = np.stack([(d0 * R[k, 0] + d1 * R[k, 1]) + d2 * R[k, 2] for k in range(3)], -1).astype(f)
    nrm = np.sqrt((rd[..., 0] * rd[..., 0] + rd[..., 1] * rd[..., 1]) + rd[..., 2] * rd[..., 2])
    vd = rd / nrm[..., None]
    ro = np.broadcast_to(c2w[:3, 3], rd.shape).astype(f)
    if ndc:
        near = 1.0
        focal = float(K[0][0])
        t = -(f(near) + ro[..., 2]) / rd[..., 2]
        ro = ro + t[..., None] * rd
        ax, ay = f(-1. / (W / (2. * focal))), f(-1. / (H / (2. * focal)))
        o0 = (ax * ro[..., 0]) / ro[..., 2]
        o1 = (ay * ro[..., 1]) / ro[..., 2]
        o2 = f(1.) + f(2. * near) / ro[..., 2]
        e0 = ax * (rd[..., 0] / rd[..., 2] - ro[..., 0] / ro[..., 2])
        e1 = ay * (rd[..., 1] / rd[..., 2] - ro[..., 1] / ro[..., 2])
        e2 = f(-2. * near) / ro[..., 2]
        ro = np.stack([o0, o1, o2], -1)
        rd = np.stack([e0, e1, e2], -1)
    return ro.astype(f), rd.astype(f), vd.astype(f)


def ray_bounds(rays_o, rays_d, xyz_min, xyz_max, near, far, stepdist, far_per_ray=None,
               prec="f32"):
    lib = load()
    dt = _dt(prec)
    ro = np.ascontiguousarray(np.asarray(rays_o, np.float32).reshape(-1, 3))
    rd = np.ascontiguousarray(np.asarray(rays_d, np.float32).reshape(-1, 3))
    n = len(ro)
    tmin = np.zeros(n, dt)
    tmax = np.zeros(n, dt)
    ns = np.zeros(n, np.int64)
    fpr = None if far_per_ray is None else np.ascontiguousarray(np.asarray(far_per_ray, np.float32))
    mn, mx = _f3(xyz_min), _f3(xyz_max)
    getattr(lib, f"pvdb_ray_bounds_{prec}")(
        _p(ro), _p(rd), C.c_long(n), _p(mn), _p(mx), C.c_float(near),
        C.c_float(far), _p(fpr), C.c_float(stepdist), _p(tmin), _p(tmax), _p(ns))
    return tmin, tmax, ns


def sample_pts(rays_o, rays_d, t_min, n_steps, stepdist, xyz_min, xyz_max, prec="f32"):
    lib = load()
    dt = _dt(prec)
    ro = np.ascontiguousarray(np.asarray(rays_o, np.float32).reshape(-1, 3))
    rd = np.ascontiguousarray(np.asarray(rays_d, np.float32).reshape(-1, 3))
    n = len(ro)
    tm = np.ascontiguousarray(np.asarray(t_min, dt))
    ns = np.ascontiguousarray(np.asarray(n_steps, np.int64))
    offs = np.ascontiguousarray(np.concatenate([[0], np.cumsum(ns)[:-1]]).astype(np.int64))
    total = int(ns.sum())
    pts = np.zeros((total, 3), dt)
    rid = np.zeros(total, np.int64)
    sid = np.zeros(total, np.int64)
    inb = np.zeros(total, np.uint8)
    mn, mx = _f3(xyz_min), _f3(xyz_max)
    getattr(lib, f"pvdb_sample_pts_{prec}")(
        _p(ro), _p(rd), _p(tm), _p(ns), _p(offs), C.c_long(n), C.c_float(stepdist),
        _p(mn), _p(mx), _p(pts), _p(rid), _p(sid), _p(inb))
    return pts, rid, sid, inb.astype(bool)


def wld2idx(pts, xyz_min, xyz_max, reso, prec="f32"):
    lib = load()
    dt = _dt(prec)
    p = np.ascontiguousarray(np.asarray(pts, dt).reshape(-1, 3))
    m = len(p)
    xs, ys, zs = (np.zeros(m, np.float32) for _ in range(3))
    r = np.ascontiguousarray(np.asarray(reso, np.int32))
    mn, mx = _f3(xyz_min), _f3(xyz_max)
    getattr(lib, f"pvdb_wld2idx_{prec}")(_p(p), C.c_long(m), _p(mn), _p(mx), _p(r), _p(xs), _p(ys),
                                          _p(zs))
    return xs, ys, zs


def mask_lookup(mask, pts, scale, shift, prec="f32"):
    lib = load()
    dt = _dt(prec)
    m_ = np.ascontiguousarray(np.asarray(mask).astype(np.uint8))
    shape = np.ascontiguousarray(np.asarray(m_.shape, np.int32))
    p = np.ascontiguousarray(np.asarray(pts, dt).reshape(-1, 3))
    out = np.zeros(len(p), np.uint8)
    sc, sh = _f3(scale), _f3(shift)
    getattr(lib, f"pvdb_mask_lookup_{prec}")(_p(m_), _p(shape), _p(p), C.c_long(len(p)), _p(sc),
                                              _p(sh), _p(out))
    return out.astype(bool)


def raw2alpha(sigma, shift, interval, prec="f32"):
    lib = load()
    dt = _dt(prec)
    s = np.ascontiguousarray(np.asarray(sigma, dt).reshape(-1))
    e = np.zeros_like(s)
    a = np.zeros_like(s)
    getattr(lib, f"pvdb_raw2alpha_{prec}")(_p(s), C.c_long(len(s)), C.c_float(shift),
                                            C.c_float(interval), _p(e), _p(a))
    return e, a


def raw2alpha_bwd(e, g_a, interval, prec="f32"):
    lib = load()
    dt = _dt(prec)
    e = np.ascontiguousarray(np.asarray(e, dt).reshape(-1))
    g = np.ascontiguousarray(np.asarray(g_a, dt).reshape(-1))
    out = np.zeros_like(e)
    getattr(lib, f"pvdb_raw2alpha_bwd_{prec}")(_p(e), _p(g), C.c_long(len(e)), C.c_float(interval),
                                                _p(out))
    return out


def segment_bounds(ray_id, n_rays):
    rid = np.asarray(ray_id, np.int64)
    i_start = np.zeros(n_rays, np.int64)
    i_end = np.zeros(n_rays, np.int64)
    if len(rid):
        chg = np.nonzero(rid[1:] != rid[:-1])[0] + 1
        i_start[rid[chg]] = chg
        i_end[rid[chg - 1]] = chg
        i_end[rid[-1]] = len(rid)
    return i_start, i_end


def alpha2weight(alpha, ray_id, n_rays, prec="f32"):
    """Returns w, T, T_last, i_start, i_end (render_utils_kernel.cu:605-651)."""
    lib = load()
    dt = _dt(prec)
    a = np.ascontiguousarray(np.asarray(alpha, dt).reshape(-1))
    i_start, i_end = segment_bounds(ray_id, n_rays)
    w = np.zeros_like(a)
    T = np.ones_like(a)
    tl = np.ones(n_rays, dt)
    if len(a):
        getattr(lib, f"pvdb_alpha2weight_{prec}")(_p(a), C.c_long(n_rays), _p(i_start), _p(i_end),
                                                   _p(w), _p(T), _p(tl))
    return w, T, tl, i_start, i_end


def alpha2weight_bwd(alpha, w, T, t_last, i_start, i_end, g_w, g_last, prec="f32"):
    lib = load()
    dt = _dt(prec)
    arrs = [np.ascontiguousarray(np.asarray(x, dt).reshape(-1)) for x in (alpha, w, T, t_last, g_w, g_last)]
    a, w_, T_, tl, gw, gl = arrs
    out = np.zeros_like(a)
    n_rays = len(tl)
    getattr(lib, f"pvdb_alpha2weight_bwd_{prec}")(_p(a), _p(w_), _p(T_), _p(tl), _p(i_start),
                                                   _p(i_end), C.c_long(n_rays), _p(gw), _p(gl),
                                                   _p(out))
    return out


def vdb_adam(param, grad, m, v, active, group, mode, stepsz, eps, b0, b1, per_lr=None, prec="f32"):
    """In-place Adam on slot payloads; param/grad/m/v are [nslots, C]."""
    lib = load()
    dt = _dt(prec)
    for x in (param, grad, m, v):
        assert x.dtype == dt and x.flags.c_contiguous
    act = np.ascontiguousarray(np.asarray(active).astype(np.uint8))
    pl = None if per_lr is None else np.ascontiguousarray(np.asarray(per_lr, dt))
    nch = param.size // len(act)
    getattr(lib, f"pvdb_vdb_adam_{prec}")(_p(param), _p(grad), _p(m), _p(v), _p(pl), _p(act),
                                           C.c_long(len(act)), C.c_int(nch), C.c_int(group),
                                           C.c_int(mode),
                                           C.c_float(stepsz), C.c_float(eps), C.c_float(b0),
                                           C.c_float(b1))


def dense_adam(param, grad, m, v, mode, step_size, eps, b0, b1, per_lr=None, prec="f32"):
    lib = load()
    dt = _dt(prec)
    for x in (param, grad, m, v):
        assert x.dtype == dt and x.flags.c_contiguous
    pl = None if per_lr is None else np.ascontiguousarray(np.asarray(per_lr, dt))
    getattr(lib, f"pvdb_dense_adam_{prec}")(_p(param), _p(grad), _p(m), _p(v), _p(pl),
                                             C.c_long(param.size), C.c_int(mode),
                                             C.c_float(step_size), C.c_float(eps), C.c_float(b0),
                                             C.c_float(b1))


def adam_stepsz(lr, beta0, beta1, step):
    """lr*sqrt(1-beta1^t)/(1-beta0^t) in float (plenvdb.h:753; adam_upd_kernel.cu:72)."""
    f = np.float32
    t = f(step)
    return float(f(lr) * np.sqrt(f(1) - np.power(f(beta1), t)) / (f(1) - np.power(f(beta0), t)))


# ----------------------------------------------------------------------------
# Total-variation gradient
# ----------------------------------------------------------------------------
def tv_add_grad_dense(param, grad, wx, wy, wz, dense_mode):
    """Dense TV gradient (total_variation_kernel.cu:13-35, host scaling :45-47)
    in FP32: param/grad are [C, X, Y, Z]; returns the new grad.  Each element
    whose grad is nonzero (or every element in dense mode) gains
    sum over its six neighbours of w * clamp(p - p_nb, -1, 1), in the order
    z-1, z+1, y-1, y+1, x-1, x+1, with w = weight / 6 and no term past the
    border.  The reference weighs the x axis with wz, not wx (lines 31-32);
    that is kept, so wx has no effect."""
    f = np.float32
    p = np.asarray(param, dtype=f)
    g = np.array(grad, dtype=f)
    wy6, wz6 = f(wy) / f(6), f(wz) / f(6)
    gta = np.zeros_like(p)
    for w, axis in ((wz6, 3), (wy6, 2), (wz6, 1)):
        n = p.shape[axis]
        lo = [slice(None)] * 4
        hi = [slice(None)] * 4
        lo[axis], hi[axis] = slice(0, n - 1), slice(1, n)
        lo, hi = tuple(lo), tuple(hi)
        prev = np.zeros_like(p)  # p - p[index - 1] where that neighbour exists
        prev[hi] = p[hi] - p[lo]
        nxt = np.zeros_like(p)  # p - p[index + 1]
        nxt[lo] = p[lo] - p[hi]
        gta = gta + w * np.clip(prev, f(-1), f(1))
        gta = gta + w * np.clip(nxt, f(-1), f(1))
    upd = np.ones_like(g, dtype=bool) if dense_mode else g != 0
    return np.where(upd, g + gta, g).astype(f)


def tv_add_grad_vdb(dom: SlotDomain, vals, grad, reso, wx, wy, wz, dense_mode):
    """The dense rule above on a VDB payload: every slot whose voxel lies in
    [0, reso) takes the dense kernel's update on the grid's dense view
    (stored value in an existing leaf, 0 elsewhere, as get_dense_grid).
    vals/grad are [nslots, C]; returns the new grad."""
    f = np.float32
    vals = np.asarray(vals, dtype=f)
    grad = np.asarray(grad, dtype=f)
    C_ = vals.shape[1]
    R = [int(r) for r in reso]
    xyz = dom.coords()
    inside = np.all((xyz >= 0) & (xyz < np.array(R)), axis=1)
    x, y, z = xyz[inside].T
    dense = np.zeros((C_, *R), dtype=f)
    gd = np.zeros((C_, *R), dtype=f)
    dense[:, x, y, z] = vals[inside].T
    gd[:, x, y, z] = grad[inside].T
    new = tv_add_grad_dense(dense, gd, wx, wy, wz, dense_mode)
    out = grad.copy()
    out[inside] = new[:, x, y, z].T
    return out


# ----------------------------------------------------------------------------
# Composed reference training step
# ----------------------------------------------------------------------------
def positional_encoding(viewdirs, n_freq=4):
    """[v, sin(v*2^k), cos(v*2^k)] ordered [vx*1..8, vy.., vz..] (dvgo.py:354-357)."""
    v = np.asarray(viewdirs)
    freq = (2.0 ** np.arange(n_freq)).astype(v.dtype)
    emb = (v[:, :, None] * freq[None, None, :]).reshape(len(v), -1)
    return np.concatenate([v, np.sin(emb), np.cos(emb)], axis=1)


def reference_train_step(scene: dict, prec: str = "f32", with_mlp: bool = True):
    """One training step's forward and backward with the reference's semantics.

    scene keys: origins, vmask (topology), density [nslots], color [nslots,C],
    reso, xyz_min, xyz_max, rays_o, rays_d, viewdirs, target, near, stepsize,
    voxel_size, voxel_size_ratio, act_shift, fast_color_thres, bg, mlp (dict of
    w0,b0,w1,b1,w2,b2 with w as [in,out]), optional far_per_ray and optional
    loss weights weight_main (1), weight_entropy_last (0), weight_rgbper (0).

    The loss is run.py:551-574: weight_main * MSE, plus weight_entropy_last *
    the mean over rays of -(p log p + (1-p) log(1-p)) with p = alphainv_last
    clamped to [1e-6, 1-1e-6], plus weight_rgbper * sum over the weight-kept
    samples of weights.detach() * |raw_rgb - target[ray_id]|^2, divided by N.

    The occupancy mask is the topology's active-voxel set, looked up with the
    MaskGrid rule (grid.py:234-243).  Returns a dict with the loss, rendered
    rgb, the sample set, the slot gradients of both grids and the MLP grads.
    """
    import torch

    dt = _dt(prec)
    tdt = torch.float32 if prec == "f32" else torch.float64
    dom = SlotDomain(scene["origins"], scene["vmask"])
    reso = np.asarray(scene["reso"], np.int32)
    mn = np.asarray(scene["xyz_min"], np.float32)
    mx = np.asarray(scene["xyz_max"], np.float32)
    N = len(scene["rays_o"])
    stepdist = np.float32(scene["stepsize"]) * np.float32(scene["voxel_size"])
    # sample_ray (dvgo.py:272-294) with far forced to 1e9.  Sample positions and
    # the occupancy decision are the reference's FP32 integer work in both
    # precisions; FP64 applies to the arithmetic on those samples.
    tmin, tmax, ns = ray_bounds(scene["rays_o"], scene["rays_d"], mn, mx, scene["near"], 1e9,
                                stepdist, far_per_ray=scene.get("far_per_ray"), prec="f32")
    pts, rid, sid, inb = sample_pts(scene["rays_o"], scene["rays_d"], tmin, ns, stepdist, mn, mx,
                                    prec="f32")
    pts, rid, sid = pts[inb], rid[inb], sid[inb]
    # mask_cache (dvgo.py:311-316), mask = active voxels of the topology
    occ = dense_active_mask(dom, reso)
    scale = ((reso.astype(np.float32) - np.float32(1)) / (mx - mn)).astype(np.float32)
    shift = (-mn * scale).astype(np.float32)
    keep = mask_lookup(occ, pts, scale, shift, prec="f32")
    pts, rid, sid = pts[keep], rid[keep], sid[keep]
    n_mask = len(rid)
    # density + raw2alpha
    xs, ys, zs = wld2idx(pts, mn, mx, reso, prec="f32")
    sigma = trilinear_fwd(dom, scene["density"], xs, ys, zs, 0, prec)[:, 0]
    interval = np.float32(scene["stepsize"]) * np.float32(scene["voxel_size_ratio"])
    e, alpha = raw2alpha(sigma, scene["act_shift"], interval, prec)
    thres = float(scene["fast_color_thres"])
    idx_a = np.arange(n_mask)
    if thres > 0:
        ka = alpha > thres
        idx_a = idx_a[ka]
    a_c = alpha[idx_a]
    rid_a = rid[idx_a]
    w, T, tl, i_s, i_e = alpha2weight(a_c, rid_a, N, prec)
    idx_w = np.arange(len(idx_a))
    if thres > 0:
        idx_w = idx_w[w > thres]
    sel = idx_a[idx_w]  # indices into the mask-kept sample list
    # colour, MLP, composite, loss in torch CPU autograd
    cx, cy, cz = xs[sel], ys[sel], zs[sel]
    k0 = trilinear_fwd(dom, scene["color"], cx, cy, cz, 1, prec)
    k0_t = torch.tensor(k0, dtype=tdt, requires_grad=True)
    w_t = torch.tensor(w[idx_w], dtype=tdt, requires_grad=True)
    tl_t = torch.tensor(tl, dtype=tdt, requires_grad=True)
    rid_s = torch.tensor(rid[sel], dtype=torch.int64)
    params = {}
    if with_mlp:
        for k in ("w0", "b0", "w1", "b1", "w2", "b2"):
            params[k] = torch.tensor(np.asarray(scene["mlp"][k], dt), dtype=tdt, requires_grad=True)
        pe = torch.tensor(positional_encoding(np.asarray(scene["viewdirs"], dt)), dtype=tdt)
        feat = torch.cat([k0_t, pe[rid_s]], dim=1)
        h = torch.relu(feat @ params["w0"] + params["b0"])
        h = torch.relu(h @ params["w1"] + params["b1"])
        logit = h @ params["w2"] + params["b2"]
        rgb = torch.sigmoid(logit)
    else:
        rgb = torch.sigmoid(k0_t)
    nch = rgb.shape[1]
    acc = torch.zeros(N, nch, dtype=tdt).index_add_(0, rid_s, w_t[:, None] * rgb)
    rgb_marched = acc + tl_t[:, None] * float(scene["bg"])
    target = torch.tensor(np.asarray(scene["target"], dt), dtype=tdt)
    loss = float(scene.get("weight_main", 1.0)) * torch.mean((rgb_marched - target) ** 2)
    w_ent = float(scene.get("weight_entropy_last", 0.0))
    if w_ent > 0:  # run.py:553-556
        pout = tl_t.clamp(1e-6, 1 - 1e-6)
        loss = loss + w_ent * (-(pout * torch.log(pout) + (1 - pout) * torch.log(1 - pout))).mean()
    w_rp = float(scene.get("weight_rgbper", 0.0))
    if w_rp > 0:  # run.py:570-573
        rgbper = (rgb - target[rid_s]).pow(2).sum(-1)
        loss = loss + w_rp * (rgbper * w_t.detach()).sum() / N
    loss.backward()
    g_k0 = k0_t.grad.numpy()
    g_w = np.zeros(len(a_c), dt)
    g_w[idx_w] = w_t.grad.numpy()
    g_tl = tl_t.grad.numpy()
    # Alphas2Weights / Raw2Alpha backward, then the two scatters
    g_a = alpha2weight_bwd(a_c, w, T, tl, i_s, i_e, g_w, g_tl, prec)
    g_alpha_full = np.zeros(n_mask, dt)
    g_alpha_full[idx_a] = g_a
    g_sigma = raw2alpha_bwd(e, g_alpha_full, interval, prec)
    g_den = trilinear_bwd(dom, 1, xs, ys, zs, g_sigma, 0, prec)[:, 0]
    g_col = trilinear_bwd(dom, k0.shape[1], cx, cy, cz, g_k0, 1, prec)
    out = {
        "loss": float(loss.detach()),
        "rgb_marched": rgb_marched.detach().numpy(),
        "alphainv_last": tl,
        "n_mask": n_mask,
        "ray_id_mask": rid,
        "step_id_mask": sid,
        "sel": sel,
        "ray_id": rid[sel],
        "step_id": sid[sel],
        "weights": w[idx_w],
        "sigma": sigma,
        "grad_density": g_den,
        "grad_color": g_col,
        "n_samples_total": int(ns.sum()),
    }
    if with_mlp:
        out["grad_mlp"] = {k: v.grad.numpy() for k, v in params.items()}
    return out


def dense_active_mask(dom: SlotDomain, reso) -> np.ndarray:
    """Dense [X,Y,Z] bool mask of the active voxels inside [0, reso)."""
    reso = [int(r) for r in reso]
    m = np.zeros(reso, dtype=bool)
    c = dom.coords()[dom.active()]
    ok = np.all((c >= 0) & (c < np.asarray(reso)), axis=1)
    c = c[ok]
    m[c[:, 0], c[:, 1], c[:, 2]] = True
    return m


# ----------------------------------------------------------------------------
# Merged-VDB renderer
# ----------------------------------------------------------------------------
def reference_render(idx_grid, den, col, mlp, c2w, K, xyz_min, xyz_max, H, W, near, stepdist,
                     act_shift, interval, fast_color_thres, bg, inverse_y=False, t_stop=1e-3,
                     prec="f32"):
    """MGRenderer::render_an_image (renderer.cu:370-424) -> [H*W*3]."""
    lib = load()
    dt = _dt(prec)
    idx = np.ascontiguousarray(np.asarray(idx_grid, np.int32))
    reso = np.ascontiguousarray(np.asarray(idx.shape, np.int32))
    den = np.ascontiguousarray(np.asarray(den, np.float32).reshape(-1))
    col = np.ascontiguousarray(np.asarray(col, np.float32).reshape(-1))
    dcol = col.size // den.size
    c2w = np.ascontiguousarray(np.asarray(c2w, np.float32).reshape(16))
    K = np.ascontiguousarray(np.asarray(K, np.float32).reshape(9))
    mn, mx = _f3(xyz_min), _f3(xyz_max)
    HW = H * W
    tmin = np.zeros(HW, dt)
    tmax = np.zeros(HW, dt)
    step = np.zeros(HW, dt)
    pe = np.zeros((HW, 27), dt)
    cnt = np.zeros(HW, np.int64)
    getattr(lib, f"pvdb_render_count_{prec}")(
        _p(idx), _p(reso), _p(den), _p(c2w), _p(K), _p(mn), _p(mx), C.c_int(H), C.c_int(W),
        C.c_int(int(inverse_y)), C.c_float(near), C.c_float(stepdist), C.c_float(act_shift),
        C.c_float(interval), C.c_float(fast_color_thres), C.c_double(t_stop), _p(tmin), _p(tmax),
        _p(step), _p(pe), _p(cnt))
    offs = np.ascontiguousarray(np.concatenate([[0], np.cumsum(cnt)[:-1]]).astype(np.int64))
    M = int(cnt.sum())
    feat = np.zeros((M, dcol), dt)
    wts = np.zeros(M, dt)
    aux = np.zeros((M, 4), dt)
    pix = np.zeros(M, np.int64)
    written = np.zeros(HW, np.int64)
    tl = np.ones(HW, dt)
    getattr(lib, f"pvdb_render_fill_{prec}")(
        _p(idx), _p(reso), _p(den), _p(col), C.c_int(dcol), _p(c2w), _p(K), _p(mn), _p(mx),
        C.c_int(H), C.c_int(W), C.c_int(int(inverse_y)), C.c_float(near), C.c_float(stepdist),
        C.c_float(act_shift), C.c_float(interval), C.c_float(fast_color_thres), _p(tmin), _p(tmax),
        _p(cnt), _p(offs), _p(feat), _p(wts), _p(aux), _p(pix), _p(written), _p(tl))
    valid = np.zeros(M, bool)
    for n in np.nonzero(written)[0]:
        valid[offs[n]:offs[n] + written[n]] = True
    feat, wts, pix, aux = feat[valid], wts[valid], pix[valid], aux[valid]
    w0 = np.asarray(mlp["w0"], dt)
    h = feat @ w0[:dcol] + (pe @ w0[dcol:])[pix] + np.asarray(mlp["b0"], dt)
    h = np.maximum(h, 0)
    h = np.maximum(h @ np.asarray(mlp["w1"], dt) + np.asarray(mlp["b1"], dt), 0)
    raw = h @ np.asarray(mlp["w2"], dt) + np.asarray(mlp["b2"], dt)
    rgb = 1 / (1 + np.exp(-raw))
    img = np.repeat((tl * dt(bg))[:, None], 3, axis=1)
    np.add.at(img, pix, wts[:, None] * rgb)
    return img.reshape(-1), {"count": cnt, "written": written, "t_last": tl, "feat": feat,
                             "weight": wts, "pix": pix, "xyz": aux[:, :3], "alpha": aux[:, 3],
                             "tmin": tmin, "tmax": tmax, "step": step}
