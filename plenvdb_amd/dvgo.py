"""Reference-structured model forward on the device (plain FP32 path).

`DirectVoxGO.forward` follows plenvdb/lib/dvgo.py:296-388 stage by stage: a
materialised sample list, the occupancy skip, the density gather, raw2alpha,
the alpha cull, Alphas2Weights, the weight cull, the colour gather, the
view-dependent MLP (torch FP32) and the segment-sum composite.  Every custom
CUDA op of the reference is a libplenvdb_hip.so kernel here.  This path is
both the drop-in for callers that use the reference model API and the
stage-per-kernel FP32 baseline the fused path (fused.py) is measured against.
"""
from __future__ import annotations

import numpy as np
import torch
import torch.nn as nn
import torch.nn.functional as F

from . import render_utils
from .grid import MaskGrid, VDBGrid
from .topology import Topology


class Raw2Alpha(torch.autograd.Function):
    """dvgo.py:408-432."""

    @staticmethod
    def forward(ctx, density, shift, interval):
        exp, alpha = render_utils.raw2alpha(density, shift, interval)
        if density.requires_grad:
            ctx.save_for_backward(exp)
            ctx.interval = interval
        return alpha

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, grad_back):
        (exp,) = ctx.saved_tensors
        return render_utils.raw2alpha_backward(exp, grad_back.contiguous(), ctx.interval), None, None


class Alphas2Weights(torch.autograd.Function):
    """dvgo.py:450-466."""

    @staticmethod
    def forward(ctx, alpha, ray_id, N):
        weights, T, alphainv_last, i_start, i_end = render_utils.alpha2weight(alpha, ray_id, N)
        if alpha.requires_grad:
            ctx.save_for_backward(alpha, weights, T, alphainv_last, i_start, i_end)
            ctx.n_rays = N
        return weights, alphainv_last

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, grad_weights, grad_last):
        alpha, weights, T, alphainv_last, i_start, i_end = ctx.saved_tensors
        grad = render_utils.alpha2weight_backward(alpha, weights, T, alphainv_last, i_start, i_end,
                                                  ctx.n_rays, grad_weights.contiguous(),
                                                  grad_last.contiguous())
        return grad, None, None


def positional_encoding(viewdirs: torch.Tensor, viewfreq: torch.Tensor) -> torch.Tensor:
    """[v, sin(v*2^k), cos(v*2^k)] (dvgo.py:354-357)."""
    emb = (viewdirs.unsqueeze(-1) * viewfreq).flatten(-2)
    return torch.cat([viewdirs, emb.sin(), emb.cos()], -1)


def build_rgbnet(dim0: int, width: int = 128, depth: int = 3) -> nn.Module:
    """The rgbnet of dvgo.py:91-109: Linear-ReLU, (depth-2) x [Linear-ReLU], Linear
    to 3, last bias zero."""
    layers = [nn.Linear(dim0, width), nn.ReLU(inplace=True)]
    for _ in range(depth - 2):
        layers += [nn.Sequential(nn.Linear(width, width), nn.ReLU(inplace=True))]
    layers += [nn.Linear(width, 3)]
    net = nn.Sequential(*layers)
    nn.init.constant_(net[-1].bias, 0)
    return net


class DirectVoxGO(nn.Module):
    """Fine-stage DVGO on VDB grids (dvgo.py:24-388), VDBGrid density + k0.

    Construction mirrors the reference's arguments that matter on the hot path;
    `topology` (optional) gives both grids and the occupancy a shared sparse
    topology, otherwise the reference's dense topology is used.
    """

    def __init__(self, xyz_min, xyz_max, world_size, alpha_init=1e-2, fast_color_thres=0.0,
                 rgbnet_dim=12, rgbnet_width=128, rgbnet_depth=3, viewbase_pe=4,
                 voxel_size=None, voxel_size_ratio=1.0, topology: Topology | None = None,
                 mlp_params: dict | None = None, mask=None, k0_dim: int | None = None):
        super().__init__()
        self.register_buffer("xyz_min", torch.as_tensor(xyz_min, dtype=torch.float32).clone())
        self.register_buffer("xyz_max", torch.as_tensor(xyz_max, dtype=torch.float32).clone())
        self.world_size = torch.as_tensor(world_size).long()
        self.fast_color_thres = float(fast_color_thres)
        self.alpha_init = alpha_init
        self.act_shift = float(np.float32(np.log(1 / (1 - alpha_init) - 1)))
        if voxel_size is None:
            vol = float(torch.prod(self.xyz_max - self.xyz_min))
            voxel_size = (vol / float(torch.prod(self.world_size.float()))) ** (1 / 3)
        self.voxel_size = float(voxel_size)
        self.voxel_size_ratio = float(voxel_size_ratio)
        if topology is None and mask is not None:
            topology = Topology.from_mask(mask)
        self.topology = topology
        self.density = VDBGrid(1, self.world_size, self.xyz_min, self.xyz_max, topology=topology)
        # rgbnet-less models composite sigmoid(k0); k0_dim overrides the coarse default of 3
        self.k0_dim = rgbnet_dim if rgbnet_dim > 0 else (k0_dim or 3)
        self.k0 = VDBGrid(self.k0_dim, self.world_size, self.xyz_min, self.xyz_max,
                          topology=topology)
        if rgbnet_dim > 0:
            self.register_buffer("viewfreq", torch.tensor([2.0 ** i for i in range(viewbase_pe)]))
            dim0 = 3 + 3 * viewbase_pe * 2 + self.k0_dim
            self.rgbnet = build_rgbnet(dim0, rgbnet_width, rgbnet_depth)
            if mlp_params is not None:
                self.load_mlp(mlp_params)
        else:
            self.rgbnet = None
        if topology is not None:
            self.mask_cache = MaskGrid(xyz_min=self.xyz_min, xyz_max=self.xyz_max, topology=topology)
        else:
            self.mask_cache = MaskGrid(mask=torch.ones(self.world_size.tolist(), dtype=torch.bool,
                                                       device=self.xyz_min.device),
                                       xyz_min=self.xyz_min, xyz_max=self.xyz_max)

    # -- grid files (dvgo.py:390-396): density.vdb and color.vdb under one prefix --
    def save_to(self, savedir: str) -> None:
        for g in (self.density, self.k0):
            g.save_to(savedir)

    def load_from(self, loaddir: str) -> None:
        """Each grid takes the file's (pruned) leaf set; the occupancy cache
        keeps its topology, whose active voxels pruning does not change."""
        for g in (self.density, self.k0):
            g.load_from(loaddir)

    @torch.no_grad()
    def load_mlp(self, p: dict):
        """Weights given as [in, out] (Linear.weight.T, run.py:98-104)."""
        lin = [m for m in self.rgbnet.modules() if isinstance(m, nn.Linear)]
        for i, l in enumerate(lin):
            l.weight.copy_(torch.as_tensor(p[f"w{i}"]).t())
            l.bias.copy_(torch.as_tensor(p[f"b{i}"]))

    def sample_ray(self, rays_o, rays_d, near, far, stepsize, far_per_ray=None, **kw):
        """dvgo.py:272-294 (far forced to 1e9 unless per-ray far bounds are given)."""
        stepdist = stepsize * self.voxel_size
        ray_pts, mask_outbbox, ray_id, step_id, *_ = render_utils.sample_pts_on_rays(
            rays_o.contiguous(), rays_d.contiguous(), self.xyz_min, self.xyz_max, near, 1e9, stepdist,
            far_per_ray=far_per_ray)
        inb = ~mask_outbbox
        return ray_pts[inb], ray_id[inb], step_id[inb]

    def _grid_xyz(self, shape):
        """World positions of a grid of `shape` points spanning the bbox
        (torch.meshgrid of three linspaces, dvgo.py:121-125)."""
        ls = [torch.linspace(float(self.xyz_min[a]), float(self.xyz_max[a]), int(shape[a]),
                             device=self.xyz_min.device) for a in range(3)]
        return torch.stack(torch.meshgrid(*ls, indexing="ij"), -1)

    def _dense_mask(self) -> torch.Tensor:
        """The occupancy as a dense bool grid (a topology MaskGrid's active voxels)."""
        mc = self.mask_cache
        if mc.topology is None:
            return mc.mask
        return torch.from_numpy(mc.topology.dense_active(self.world_size.tolist())).to(
            self.xyz_min.device)

    @torch.no_grad()
    def hit_coarse_geo(self, rays_o, rays_d, near, far, stepsize, **render_kwargs):
        """Whether each ray has a sample inside the bbox that the occupancy keeps
        (dvgo.py:258-270; far is forced to 1e9 as there)."""
        shape = rays_o.shape[:-1]
        rays_o = rays_o.reshape(-1, 3).contiguous()
        rays_d = rays_d.reshape(-1, 3).contiguous()
        stepdist = stepsize * self.voxel_size
        ray_pts, mask_outbbox, ray_id, *_ = render_utils.sample_pts_on_rays(
            rays_o, rays_d, self.xyz_min, self.xyz_max, near, 1e9, stepdist)
        inb = ~mask_outbbox
        hit = torch.zeros(len(rays_o), dtype=torch.bool, device=rays_o.device)
        hit[ray_id[inb][self.mask_cache(ray_pts[inb])]] = True
        return hit.reshape(shape)

    def _set_grid_resolution(self, num_voxels, num_voxels_base=None):
        """dvgo.py:133-142: voxel size from the voxel budget, world size and the
        step-interval ratio against the base voxel size."""
        self.num_voxels = int(num_voxels)
        vol = (self.xyz_max - self.xyz_min).prod()
        self.voxel_size = float((vol / num_voxels).pow(1 / 3))
        self.world_size = ((self.xyz_max - self.xyz_min) / self.voxel_size).long().cpu()
        if num_voxels_base is not None:
            self.voxel_size_base = float((vol / num_voxels_base).pow(1 / 3))
        base = getattr(self, "voxel_size_base", self.voxel_size)
        self.voxel_size_ratio = self.voxel_size / base

    @torch.no_grad()
    def maskout_near_cam_vox(self, cam_o, near_clip):
        """Density of the grid points within near_clip of any camera centre <- -100
        (dvgo.py:164-175)."""
        xyz = self._grid_xyz(self.world_size.tolist())
        cam_o = torch.as_tensor(cam_o, dtype=torch.float32, device=xyz.device)
        nearest = torch.stack([
            (xyz.unsqueeze(-2) - co).pow(2).sum(-1).sqrt().amin(-1) for co in cam_o.split(100)
        ]).amin(0)
        self.density.setValuesOn_bymask(nearest <= near_clip, -100)

    @torch.no_grad()
    def update_occupancy_cache(self):
        """mask &= maxpool3(alpha(density at the cache grid points)) > fast_color_thres
        (dvgo.py:201-210).  A topology-backed cache becomes the dense mask of its
        active voxels first, as the reference keeps it."""
        mask = self._dense_mask()
        xyz = self._grid_xyz(mask.shape)
        alpha = self.activate_density(self.density(xyz))[None, None]
        alpha = F.max_pool3d(alpha, kernel_size=3, padding=1, stride=1)[0, 0]
        self.mask_cache = MaskGrid(mask=mask & (alpha > self.fast_color_thres),
                                   xyz_min=self.xyz_min, xyz_max=self.xyz_max)

    @torch.no_grad()
    def scale_volume_grid(self, num_voxels):
        """dvgo.py:178-198: new resolution, both grids resampled (VDBGrid
        scale_volume_grid), and for grids up to 256^3 a dense occupancy cache
        from the old cache and the max-pooled alpha."""
        self._set_grid_resolution(num_voxels)
        old_mask = self.mask_cache
        self.density.scale_volume_grid(self.world_size)
        self.k0.scale_volume_grid(self.world_size)
        if int(torch.prod(self.world_size)) <= 256 ** 3:
            xyz = self._grid_xyz(self.world_size.tolist())
            alpha = F.max_pool3d(self.activate_density(self.density.get_dense_grid()),
                                 kernel_size=3, padding=1, stride=1)[0, 0]
            self.mask_cache = MaskGrid(mask=old_mask(xyz) & (alpha > self.fast_color_thres),
                                       xyz_min=self.xyz_min, xyz_max=self.xyz_max)
        self.topology = None

    def activate_density(self, density, interval=None):
        interval = interval if interval is not None else self.voxel_size_ratio
        shape = density.shape
        return Raw2Alpha.apply(density.flatten(), self.act_shift, interval).reshape(shape)

    def forward(self, rays_o, rays_d, viewdirs, global_step=None, **render_kwargs):
        """dvgo.py:296-388."""
        N = len(rays_o)
        ray_pts, ray_id, step_id = self.sample_ray(rays_o=rays_o, rays_d=rays_d, **render_kwargs)
        interval = render_kwargs["stepsize"] * self.voxel_size_ratio
        mask = self.mask_cache(ray_pts)
        ray_pts, ray_id, step_id = ray_pts[mask], ray_id[mask], step_id[mask]
        n_samples = ray_pts.shape[0]
        density = self.density(ray_pts)
        alpha = self.activate_density(density, interval)
        if self.fast_color_thres > 0:
            m = alpha > self.fast_color_thres
            ray_pts, ray_id, step_id, density, alpha = ray_pts[m], ray_id[m], step_id[m], density[m], alpha[m]
        weights, alphainv_last = Alphas2Weights.apply(alpha, ray_id, N)
        if self.fast_color_thres > 0:
            m = weights > self.fast_color_thres
            weights, alpha, ray_pts, ray_id, step_id = weights[m], alpha[m], ray_pts[m], ray_id[m], step_id[m]
        k0 = self.k0(ray_pts)
        if self.rgbnet is None:
            rgb = torch.sigmoid(k0)
        else:
            emb = positional_encoding(viewdirs, self.viewfreq)
            emb = emb.flatten(0, -2)[ray_id]
            rgb = torch.sigmoid(self.rgbnet(torch.cat([k0, emb], -1)))
        rgb_marched = torch.zeros(N, rgb.shape[-1], device=rgb.device, dtype=rgb.dtype)
        rgb_marched = rgb_marched.index_add(0, ray_id, weights.unsqueeze(-1) * rgb)
        rgb_marched = rgb_marched + alphainv_last.unsqueeze(-1) * render_kwargs["bg"]
        return {
            "alphainv_last": alphainv_last,
            "weights": weights,
            "rgb_marched": rgb_marched,
            "raw_alpha": alpha,
            "raw_rgb": rgb,
            "ray_id": ray_id,
            "n_samples": n_samples,
        }
