"""VDB grid modules with a device-resident autograd boundary.

`VDBGrid` keeps the interface of plenvdb/lib/grid.py:65-137 but its
`QueryVerticalInVDB` runs the gather and the scatter on the device: the
reference copies the points to the host and back on every forward and
backward (grid.py:40-60, plenvdb.h:441-485).
"""
from __future__ import annotations

import numpy as np
import torch
import torch.nn as nn
import torch.nn.functional as F

from . import lib
from .plenvdb import ColorVDB, DensityVDB
from .topology import Topology


class QueryVerticalInVDB(torch.autograd.Function):
    """grid.py:40-60 on the device: forward gathers [N, ndim], backward
    scatters grad_output into the grid's gradient payload (the points get no
    gradient, as in the reference)."""

    @staticmethod
    def forward(ctx, pts, grid):  # pts: (3, N) index space
        N = pts.shape[-1]
        ctx.grid = grid
        p = pts.detach().contiguous()
        ctx.save_for_backward(p)
        if N == 0:
            return torch.zeros(0, grid.getndim(), device=pts.device)
        return grid.forward(p[0], p[1], p[2])

    @staticmethod
    def backward(ctx, grad_output):
        if ctx.needs_input_grad[0]:
            (p,) = ctx.saved_tensors
            if p.shape[-1] > 0:
                ctx.grid.backward(p[0], p[1], p[2], grad_output.contiguous())
        return None, None


class VDBGrid(nn.Module):
    """grid.py:65-137."""

    def __init__(self, channels, world_size, xyz_min, xyz_max, topology: Topology | None = None,
                 mask=None, **kwargs):
        super().__init__()
        self.channels = int(channels)
        self.world_size = torch.as_tensor(world_size).long()
        self.register_buffer("xyz_min", torch.as_tensor(xyz_min, dtype=torch.float32).clone())
        self.register_buffer("xyz_max", torch.as_tensor(xyz_max, dtype=torch.float32).clone())
        ws = self.world_size.tolist()
        self.register_buffer("ws_minus_1", (self.world_size - 1).float())
        if self.channels == 1:
            self.grid = DensityVDB(ws, 1, topology=topology, mask=mask)
        else:
            self.grid = ColorVDB(ws, self.channels, topology=topology, mask=mask)

    def wld2idx(self, pts):
        """(pts-min)/(max-min)*(world_size-1), grid.py:77-78."""
        return (pts - self.xyz_min) / (self.xyz_max - self.xyz_min) * self.ws_minus_1

    def forward(self, xyz):
        shape = xyz.shape[:-1]
        pts = self.wld2idx(xyz).reshape(-1, 3).t().contiguous().float().requires_grad_()
        out = QueryVerticalInVDB.apply(pts, self.grid).reshape([*shape, self.channels])
        if self.channels == 1:
            out = out.squeeze(-1)
        return out

    def scale_volume_grid(self, new_world_size):
        """grid.py:91-102: dense get -> trilinear resample -> new dense grid."""
        data = self.get_dense_grid()
        new_ws = torch.as_tensor(new_world_size).long()
        new_data = F.interpolate(data, size=tuple(new_ws.tolist()), mode="trilinear",
                                 align_corners=True).permute(0, 2, 3, 4, 1).squeeze(0).contiguous()
        ws = new_ws.tolist()
        if self.channels == 1:
            self.grid = DensityVDB(ws, 1)
        else:
            self.grid = ColorVDB(ws, self.channels)
        self.grid.copyFromDense(new_data.reshape(-1))
        self.world_size = new_ws
        self.ws_minus_1 = (new_ws - 1).float().to(self.xyz_min.device)

    def total_variation_add_grad(self, wx, wy, wz, dense_mode):
        """Stubbed in the reference (grid.py:104-107)."""
        return

    def setValuesOn_bymask(self, mask, val):
        assert self.channels == 1
        self.grid.setValuesOn_bymask(mask.reshape(-1), val)

    def get_dense_grid(self):
        d = self.grid.get_dense_grid_device()  # [X,Y,Z,C]
        return d.permute(3, 0, 1, 2).unsqueeze(0).contiguous()

    def load_from(self, loadpath):
        self.grid.load_from(loadpath + ("density.vdb" if self.channels == 1 else "color.vdb"))

    def save_to(self, savepath):
        self.grid.save_to(savepath + ("density.vdb" if self.channels == 1 else "color.vdb"))

    def resetTimer(self):
        self.grid.resetTimer()

    def getTimer(self):
        pass

    def extra_repr(self):
        return f"channels={self.channels}, world_size={self.world_size.tolist()}"


class MaskGrid(nn.Module):
    """Occupancy for empty-space skipping (grid.py:207-246).  Either a dense bool
    mask (reference form, looked up with maskcache_lookup) or a topology whose
    active voxels are the occupied set (looked up on the VDB value masks)."""

    def __init__(self, mask=None, xyz_min=None, xyz_max=None, topology: Topology | None = None):
        super().__init__()
        xyz_min = torch.as_tensor(xyz_min, dtype=torch.float32).cpu()
        xyz_max = torch.as_tensor(xyz_max, dtype=torch.float32).cpu()
        self.topology = topology
        if topology is not None:
            shape = torch.tensor(topology.reso, dtype=torch.float32)
            self.register_buffer("mask", torch.zeros(0, dtype=torch.bool))
        else:
            mask = torch.as_tensor(mask).bool()
            shape = torch.tensor(list(mask.shape), dtype=torch.float32)
            self.register_buffer("mask", mask)
        xyz_len = xyz_max - xyz_min
        scale = (shape - 1) / xyz_len
        self.register_buffer("xyz2ijk_scale", scale)
        self.register_buffer("xyz2ijk_shift", -xyz_min * scale)
        self._scale = scale.tolist()
        self._shift = (-xyz_min * scale).tolist()

    @torch.no_grad()
    def forward(self, xyz):
        from . import render_utils

        shape = xyz.shape[:-1]
        xyz = xyz.reshape(-1, 3)
        if self.topology is not None:
            m = render_utils.active_lookup(self.topology, xyz, self._scale, self._shift)
        else:
            m = render_utils.maskcache_lookup(self.mask, xyz, self._scale, self._shift)
        return m.reshape(shape)
