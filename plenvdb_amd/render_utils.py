"""Device ops with the interface of the reference's `render_utils_cuda` and
`adam_upd_cuda` extensions (plenvdb/lib/cuda/render_utils.cpp:170-184,
adam_upd.cpp:79-86), backed by libplenvdb_hip.so.

Same names, argument order, return lists and dtypes (int64 ids, bool masks).
"""
from __future__ import annotations

import torch

from . import lib


def _f32c(t: torch.Tensor) -> torch.Tensor:
    lib.require_cuda(t)
    return t.contiguous() if t.dtype == torch.float32 else t.float().contiguous()


def _host3(t) -> list:
    if isinstance(t, torch.Tensor):
        return [float(v) for v in t.detach().reshape(-1).cpu().tolist()]
    return [float(v) for v in t]


def infer_t_minmax(rays_o, rays_d, xyz_min, xyz_max, near, far, far_per_ray=None):
    n = rays_o.shape[0]
    ro, rd = _f32c(rays_o), _f32c(rays_d)
    t_min = torch.empty(n, device=ro.device, dtype=torch.float32)
    t_max = torch.empty_like(t_min)
    n_steps = torch.empty(n, device=ro.device, dtype=torch.int64)
    mn, mx = lib.f3(_host3(xyz_min)), lib.f3(_host3(xyz_max))
    lib.call("pvdb_ray_bounds", lib.ptr(ro), lib.ptr(rd), n, mn, mx, float(near), float(far),
             lib.ptr(far_per_ray), 1.0, lib.ptr(t_min), lib.ptr(t_max), lib.ptr(n_steps), lib.stream())
    return [t_min, t_max]


def sample_pts_on_rays(rays_o, rays_d, xyz_min, xyz_max, near, far, stepdist, far_per_ray=None):
    """render_utils_kernel.cu:196-242 -> [rays_pts, mask_outbbox, ray_id, step_id,
    N_steps, t_min, t_max].  Reads the total sample count back to the host, as
    the reference does with .item()."""
    n = rays_o.shape[0]
    ro, rd = _f32c(rays_o), _f32c(rays_d)
    dev = ro.device
    t_min = torch.empty(n, device=dev, dtype=torch.float32)
    t_max = torch.empty_like(t_min)
    n_steps = torch.empty(n, device=dev, dtype=torch.int64)
    mn, mx = lib.f3(_host3(xyz_min)), lib.f3(_host3(xyz_max))
    fpr = None if far_per_ray is None else _f32c(far_per_ray)
    lib.call("pvdb_ray_bounds", lib.ptr(ro), lib.ptr(rd), n, mn, mx, float(near), float(far),
             lib.ptr(fpr), float(stepdist), lib.ptr(t_min), lib.ptr(t_max), lib.ptr(n_steps),
             lib.stream())
    incl = torch.cumsum(n_steps, 0)
    total = int(incl[-1].item()) if n else 0
    pts = torch.empty(total, 3, device=dev, dtype=torch.float32)
    ray_id = torch.empty(total, device=dev, dtype=torch.int64)
    step_id = torch.empty(total, device=dev, dtype=torch.int64)
    inb = torch.empty(total, device=dev, dtype=torch.uint8)
    lib.call("pvdb_sample_pts", lib.ptr(ro), lib.ptr(rd), lib.ptr(t_min), lib.ptr(incl), n, total,
             float(stepdist), mn, mx, lib.ptr(pts), lib.ptr(ray_id), lib.ptr(step_id), lib.ptr(inb),
             lib.stream())
    return [pts, inb == 0, ray_id, step_id, n_steps, t_min, t_max]


def maskcache_lookup(world, xyz, xyz2ijk_scale, xyz2ijk_shift):
    """render_utils_kernel.cu:392-424: dense bool occupancy lookup."""
    lib.require_cuda(world)
    w = world.to(torch.uint8).contiguous() if world.dtype != torch.uint8 else world.contiguous()
    p = _f32c(xyz.reshape(-1, 3))
    out = torch.zeros(p.shape[0], device=p.device, dtype=torch.uint8)
    lib.call("pvdb_mask_lookup", lib.ptr(w), world.shape[0], world.shape[1], world.shape[2],
             lib.ptr(p), p.shape[0], lib.f3(_host3(xyz2ijk_scale)), lib.f3(_host3(xyz2ijk_shift)),
             lib.ptr(out), lib.stream())
    return out.bool()


def active_lookup(topology, xyz, xyz2ijk_scale, xyz2ijk_shift):
    """The MaskGrid rule on a topology's active voxels (no dense bool grid)."""
    p = _f32c(xyz.reshape(-1, 3))
    out = torch.zeros(p.shape[0], device=p.device, dtype=torch.uint8)
    lib.call("pvdb_active_lookup", topology.handle, lib.ptr(p), p.shape[0],
             lib.f3(_host3(xyz2ijk_scale)), lib.f3(_host3(xyz2ijk_shift)), lib.ptr(out), lib.stream())
    return out.bool()


def raw2alpha(density, shift, interval):
    """render_utils_kernel.cu:464-489 -> [exp, alpha]."""
    d = _f32c(density)
    e = torch.empty_like(d)
    a = torch.empty_like(d)
    lib.call("pvdb_raw2alpha", lib.ptr(d), d.numel(), float(shift), float(interval), lib.ptr(e),
             lib.ptr(a), lib.stream())
    return [e, a]


def raw2alpha_backward(exp, grad_back, interval):
    e, g = _f32c(exp), _f32c(grad_back)
    out = torch.empty_like(e)
    lib.call("pvdb_raw2alpha_bwd", lib.ptr(e), lib.ptr(g), e.numel(), float(interval), lib.ptr(out),
             lib.stream())
    return out


def alpha2weight(alpha, ray_id, n_rays):
    """render_utils_kernel.cu:605-651 -> [weight, T, alphainv_last, i_start, i_end]."""
    a = _f32c(alpha)
    dev = a.device
    n = a.numel()
    weight = torch.zeros_like(a)
    T = torch.ones_like(a)
    last = torch.ones(n_rays, device=dev, dtype=torch.float32)
    i_start = torch.zeros(n_rays, device=dev, dtype=torch.int64)
    i_end = torch.zeros(n_rays, device=dev, dtype=torch.int64)
    if n == 0:
        return [weight, T, last, i_start, i_end]
    rid = ray_id.contiguous()
    lib.call("pvdb_segment_bounds", lib.ptr(rid), n, lib.ptr(i_start), lib.ptr(i_end), lib.stream())
    lib.call("pvdb_alpha2weight", lib.ptr(a), n_rays, lib.ptr(weight), lib.ptr(T), lib.ptr(last),
             lib.ptr(i_start), lib.ptr(i_end), lib.stream())
    return [weight, T, last, i_start, i_end]


def alpha2weight_backward(alpha, weight, T, alphainv_last, i_start, i_end, n_rays, grad_weights,
                          grad_last):
    a = _f32c(alpha)
    grad = torch.zeros_like(a)
    if n_rays == 0:
        return grad
    gw, gl = _f32c(grad_weights), _f32c(grad_last)
    lib.call("pvdb_alpha2weight_bwd", lib.ptr(a), lib.ptr(_f32c(weight)), lib.ptr(_f32c(T)),
             lib.ptr(_f32c(alphainv_last)), lib.ptr(i_start.contiguous()), lib.ptr(i_end.contiguous()),
             n_rays, lib.ptr(gw), lib.ptr(gl), lib.ptr(grad), lib.stream())
    return grad


def adam_step_size(lr, beta1, beta2, step):
    """lr*sqrt(1-beta2^t)/(1-beta1^t) in float (adam_upd_kernel.cu:72)."""
    import numpy as np

    f = np.float32
    t = f(step)
    return float(f(lr) * np.sqrt(f(1) - np.power(f(beta2), t)) / (f(1) - np.power(f(beta1), t)))


def _dense_adam(mode, param, grad, exp_avg, exp_avg_sq, perlr, step, beta1, beta2, lr, eps):
    for t in (param, grad, exp_avg, exp_avg_sq):
        if t.dtype != torch.float32 or not t.is_contiguous():
            raise ValueError("adam tensors must be contiguous float32")
    lib.call("pvdb_dense_adam", lib.ptr(param), lib.ptr(grad), lib.ptr(exp_avg), lib.ptr(exp_avg_sq),
             lib.ptr(perlr), param.numel(), mode, adam_step_size(lr, beta1, beta2, step), float(eps),
             float(beta1), float(beta2), lib.stream())


def adam_upd(param, grad, exp_avg, exp_avg_sq, step, beta1, beta2, lr, eps):
    """adam_upd_kernel.cu:60-85."""
    _dense_adam(0, param, grad, exp_avg, exp_avg_sq, None, step, beta1, beta2, lr, eps)


def masked_adam_upd(param, grad, exp_avg, exp_avg_sq, step, beta1, beta2, lr, eps):
    """adam_upd_kernel.cu:87-108 (skip elements with zero grad)."""
    _dense_adam(1, param, grad, exp_avg, exp_avg_sq, None, step, beta1, beta2, lr, eps)


def adam_upd_with_perlr(param, grad, exp_avg, exp_avg_sq, perlr, step, beta1, beta2, lr, eps):
    """adam_upd_kernel.cu:110-132."""
    _dense_adam(2, param, grad, exp_avg, exp_avg_sq, perlr.contiguous(), step, beta1, beta2, lr, eps)
