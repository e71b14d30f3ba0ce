"""Camera rays and the training-ray builders of plenvdb/lib/dvgo.py:471-627.

`get_rays_of_a_view` runs `pvdb_rays_of_view` (one HIP thread per pixel, the
reference's float32 operation order); the builders assemble the per-view rays
of a training set the way the reference does, on the current device:
`get_training_rays` keeps the [N, H, W, 3] layout, `get_training_rays_flatten`
concatenates all pixels, and `get_training_rays_in_maskcache_sampling` keeps
only the rays that hit the occupancy cache (`DirectVoxGO.hit_coarse_geo`), the
fine stage's `ray_sampler='in_maskcache'` (configs/default.py:72, run.py:446).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import lib

_MODES = {"center": 0, "lefttop": 1, "random": 2}


def random_jitter(H: int, W: int, device) -> torch.Tensor:
    """The 'random' mode's pixel offsets [2, H, W], drawn from torch's generator
    with the reference's calls (dvgo.py:472-484: rand_like of the transposed
    meshgrid for i, then for j), so a seeded run draws the same numbers."""
    i, j = torch.meshgrid(torch.linspace(0, W - 1, W, device=device),
                          torch.linspace(0, H - 1, H, device=device), indexing="ij")
    i = i.t().float()
    j = j.t().float()
    ri = torch.rand_like(i)
    rj = torch.rand_like(j)
    return torch.stack([ri, rj]).contiguous()


def get_rays_of_a_view(H, W, K, c2w, ndc, inverse_y, flip_x, flip_y, mode="center"):
    """dvgo.py:522-527: rays_o, rays_d, viewdirs [H, W, 3] (float32, device)."""
    if mode not in _MODES:
        raise NotImplementedError(f"get_rays mode {mode!r}")
    H, W = int(H), int(W)
    k = np.ascontiguousarray(np.asarray(K, np.float64).reshape(3, 3))
    c = c2w.detach().cpu().numpy() if isinstance(c2w, torch.Tensor) else np.asarray(c2w)
    c = np.ascontiguousarray(np.asarray(c, np.float32)[:3, :4])
    dev = torch.device("cuda")
    jit = random_jitter(H, W, dev) if mode == "random" else None
    ro = torch.empty(H, W, 3, dtype=torch.float32, device=dev)
    rd = torch.empty_like(ro)
    vd = torch.empty_like(ro)
    lib.call("pvdb_rays_of_view", H, W, k.ctypes.data_as(C.c_void_p),
             c.ctypes.data_as(C.c_void_p), int(bool(ndc)), int(bool(inverse_y)),
             int(bool(flip_x)), int(bool(flip_y)), _MODES[mode], lib.ptr(ro), lib.ptr(rd),
             lib.ptr(vd), lib.ptr(jit), lib.stream())
    return ro, rd, vd


@torch.no_grad()
def get_training_rays(rgb_tr, train_poses, HW, Ks, ndc, inverse_y, flip_x, flip_y):
    """dvgo.py:531-553 (one resolution and intrinsics for every view)."""
    assert len(np.unique(np.asarray(HW), axis=0)) == 1
    assert len(np.unique(np.asarray(Ks).reshape(len(Ks), -1), axis=0)) == 1
    assert len(rgb_tr) == len(train_poses) == len(Ks) == len(HW)
    H, W = (int(v) for v in HW[0])
    K = Ks[0]
    dev = rgb_tr.device
    rays_o_tr = torch.zeros([len(rgb_tr), H, W, 3], device=dev)
    rays_d_tr = torch.zeros_like(rays_o_tr)
    viewdirs_tr = torch.zeros_like(rays_o_tr)
    for i, c2w in enumerate(train_poses):
        ro, rd, vd = get_rays_of_a_view(H, W, K, c2w, ndc, inverse_y, flip_x, flip_y)
        rays_o_tr[i].copy_(ro)
        rays_d_tr[i].copy_(rd)
        viewdirs_tr[i].copy_(vd)
    return rgb_tr, rays_o_tr, rays_d_tr, viewdirs_tr, [1] * len(rgb_tr)


@torch.no_grad()
def get_training_rays_flatten(rgb_tr_ori, train_poses, HW, Ks, ndc, inverse_y, flip_x, flip_y):
    """dvgo.py:556-585: every pixel of every view, concatenated."""
    assert len(rgb_tr_ori) == len(train_poses) == len(Ks) == len(HW)
    dev = rgb_tr_ori[0].device
    N = sum(im.shape[0] * im.shape[1] for im in rgb_tr_ori)
    rgb_tr = torch.zeros([N, 3], device=dev)
    rays_o_tr = torch.zeros_like(rgb_tr)
    rays_d_tr = torch.zeros_like(rgb_tr)
    viewdirs_tr = torch.zeros_like(rgb_tr)
    imsz, top = [], 0
    for c2w, img, (H, W), K in zip(train_poses, rgb_tr_ori, HW, Ks):
        assert img.shape[:2] == (H, W)
        ro, rd, vd = get_rays_of_a_view(H, W, K, c2w, ndc, inverse_y, flip_x, flip_y)
        n = H * W
        rgb_tr[top:top + n].copy_(img.flatten(0, 1))
        rays_o_tr[top:top + n].copy_(ro.flatten(0, 1))
        rays_d_tr[top:top + n].copy_(rd.flatten(0, 1))
        viewdirs_tr[top:top + n].copy_(vd.flatten(0, 1))
        imsz.append(n)
        top += n
    assert top == N
    return rgb_tr, rays_o_tr, rays_d_tr, viewdirs_tr, imsz


@torch.no_grad()
def get_training_rays_in_maskcache_sampling(rgb_tr_ori, train_poses, HW, Ks, ndc, inverse_y,
                                            flip_x, flip_y, model, render_kwargs):
    """dvgo.py:588-627: the flattened rays that hit the model's occupancy cache
    (rows of 64 image lines at a time, as the reference)."""
    assert len(rgb_tr_ori) == len(train_poses) == len(Ks) == len(HW)
    CHUNK = 64
    dev = rgb_tr_ori[0].device
    N = sum(im.shape[0] * im.shape[1] for im in rgb_tr_ori)
    rgb_tr = torch.zeros([N, 3], device=dev)
    rays_o_tr = torch.zeros_like(rgb_tr)
    rays_d_tr = torch.zeros_like(rgb_tr)
    viewdirs_tr = torch.zeros_like(rgb_tr)
    imsz, top = [], 0
    for c2w, img, (H, W), K in zip(train_poses, rgb_tr_ori, HW, Ks):
        assert img.shape[:2] == (H, W)
        ro, rd, vd = get_rays_of_a_view(H, W, K, c2w, ndc, inverse_y, flip_x, flip_y)
        mask = torch.empty(img.shape[:2], device=dev, dtype=torch.bool)
        for i in range(0, img.shape[0], CHUNK):
            mask[i:i + CHUNK] = model.hit_coarse_geo(rays_o=ro[i:i + CHUNK],
                                                     rays_d=rd[i:i + CHUNK], **render_kwargs)
        n = int(mask.sum())
        rgb_tr[top:top + n].copy_(img[mask])
        rays_o_tr[top:top + n].copy_(ro[mask])
        rays_d_tr[top:top + n].copy_(rd[mask])
        viewdirs_tr[top:top + n].copy_(vd[mask])
        imsz.append(n)
        top += n
    return rgb_tr[:top], rays_o_tr[:top], rays_d_tr[:top], viewdirs_tr[:top], imsz
