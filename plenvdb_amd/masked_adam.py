"""Optimizers of the hot path behind the reference's names: `VDBAdam` drives the
two VDB grids' device optimizers and `MaskedAdam` the MLP's dense Adam
(interface of plenvdb/lib/masked_adam.py:18-158; the `vdbopt.tar` keys of
:70-100 are kept so checkpoints interchange).

Design: each VDB grid is one `_GridGroup` (its optimizer handle plus the rule
that picks the kernel's stepmode), so the density and colour sides share one
code path; the optimizer's scalar state is described once by `_SCALARS`, which
both the checkpoint writer and reader walk.  `MaskedAdam` resolves one kernel
mode per parameter and hands it to the dense Adam kernel directly.
"""
from __future__ import annotations

from dataclasses import dataclass

import torch

from . import lib, render_utils
from .plenvdb import ColorOpt, DensityOpt

# stepmodes of the VDB and dense Adam kernels (densityvdb.cu:185-368,
# adam_upd_kernel.cu:8-132)
PLAIN, SKIP_ZERO_GRAD, PER_VOXEL_LR = 0, 1, 2


def _validate(lr: float, betas, eps: float) -> None:
    problems = []
    if not lr >= 0.0:
        problems.append(f"Invalid learning rate: {lr}")
    if not eps >= 0.0:
        problems.append(f"Invalid epsilon value: {eps}")
    problems += [f"Invalid beta parameter at index {i}: {b}"
                 for i, b in enumerate(betas[:2]) if not 0.0 <= b < 1.0]
    if problems:
        raise ValueError("; ".join(problems))


@dataclass
class _GridGroup:
    """One VDB grid under optimisation: its device optimizer, whether zero
    gradients are skipped and whether a per-voxel rate was installed."""

    grid: object           # the VDBGrid module the group optimises
    opt: object            # DensityOpt / ColorOpt handle
    skip_zero_grad: bool
    per_voxel_lr: bool = False

    def mode(self) -> int:
        # a per-voxel rate wins over skipping (masked_adam.py:56-68)
        if self.per_voxel_lr:
            return PER_VOXEL_LR
        return SKIP_ZERO_GRAD if self.skip_zero_grad else PLAIN


# the optimizer scalars saved in vdbopt.tar: (key suffix, getter, setter)
_SCALARS = (("lr", "getLr", "setLr"), ("eps", "getEps", "setEps"),
            ("beta0", "getBeta0", "setBeta0"), ("beta1", "getBeta1", "setBeta1"),
            ("step", "getStep", "setStep"))
# key prefix and file prefix of each grid in a checkpoint
_ROLES = {"density": ("d", "density_"), "color": ("c", "color_")}


class VDBAdam:
    """Adam over the density and colour VDB grids (masked_adam.py:18-100).

    params: two groups {'params': VDBGrid, 'lr': float, 'skip_zero_grad': bool},
    one of one channel (density) and one of more (colour).
    """

    def __init__(self, params, lr=1e-3, betas=(0.9, 0.99), eps=1e-8):
        _validate(lr, betas, eps)
        if len(params) != 2:
            raise AssertionError("VDBAdam takes exactly two parameter groups")
        self.groups: dict[str, _GridGroup] = {}
        for spec in params:
            grid = spec["params"]
            role, cls = ("density", DensityOpt) if grid.channels == 1 else ("color", ColorOpt)
            self.groups[role] = _GridGroup(grid, cls(grid.grid, spec["lr"], eps, betas[0], betas[1]),
                                           bool(spec["skip_zero_grad"]))
        if set(self.groups) != set(_ROLES):
            raise AssertionError("VDBAdam needs one density and one colour grid")

    # the reference exposes the two handles under these names
    @property
    def densityOpt(self):
        return self.groups["density"].opt

    @property
    def colorOpt(self):
        return self.groups["color"].opt

    def set_pervoxel_lr(self, count):
        """Per-voxel rate count / max(count) on the density grid (stepmode 2)."""
        g = self.groups["density"]
        g.opt.set_pervoxel_lr((count.float() / count.max()).reshape(-1))
        g.per_voxel_lr = True

    def update_lr(self, factor):
        for g in self.groups.values():
            g.opt.update_lr(factor)

    def zero_grad(self):
        for g in self.groups.values():
            g.opt.zero_grad()

    def step(self):
        for g in self.groups.values():
            g.opt.step(g.mode())

    def save_to(self, savepath):
        state = {}
        for role, (key, _) in _ROLES.items():
            opt = self.groups[role].opt
            state.update({key + name: getattr(opt, get)() for name, get, _ in _SCALARS})
        torch.save(state, savepath + "vdbopt.tar")
        for role, (_, fprefix) in _ROLES.items():
            self.groups[role].opt.save_to(savepath + fprefix)

    def load_from(self, loadpath):
        state = torch.load(loadpath + "vdbopt.tar", weights_only=True)
        for role, (key, fprefix) in _ROLES.items():
            opt = self.groups[role].opt
            for name, _, put in _SCALARS:
                getattr(opt, put)(state[key + name])
            opt.load_from(loadpath + fprefix)


class MaskedAdam(torch.optim.Optimizer):
    """Dense Adam of the MLP with the reference's zero-grad skip and per-voxel
    rate (masked_adam.py:105-158) on the library's dense Adam kernel.  Each
    group may carry 'skip_zero_grad'; a per-voxel rate set by
    `set_pervoxel_lr` applies to parameters of its shape."""

    def __init__(self, params, lr=1e-3, betas=(0.9, 0.99), eps=1e-8):
        _validate(lr, betas, eps)
        self.per_lr = None
        super().__init__(params, dict(lr=lr, betas=betas, eps=eps, skip_zero_grad=False))

    def set_pervoxel_lr(self, count):
        assert self.param_groups[0]["params"][0].shape == count.shape
        self.per_lr = count.float() / count.max()

    def _mode(self, param, group) -> int:
        if self.per_lr is not None and param.shape == self.per_lr.shape:
            return PER_VOXEL_LR
        return SKIP_ZERO_GRAD if group["skip_zero_grad"] else PLAIN

    def _moments(self, param):
        st = self.state[param]
        if not st:  # created at a parameter's first update
            st["step"] = 0
            st["exp_avg"] = torch.zeros_like(param, memory_format=torch.preserve_format)
            st["exp_avg_sq"] = torch.zeros_like(param, memory_format=torch.preserve_format)
        st["step"] += 1
        return st

    @torch.no_grad()
    def step(self):
        for group in self.param_groups:
            b1, b2 = group["betas"]
            for param in (p for p in group["params"] if p.grad is not None):
                st = self._moments(param)
                mode = self._mode(param, group)
                grad = param.grad.contiguous()
                for t in (param, grad, st["exp_avg"], st["exp_avg_sq"]):
                    if t.dtype != torch.float32 or not t.is_contiguous():
                        raise ValueError("adam tensors must be contiguous float32")
                rate = self.per_lr.contiguous() if mode == PER_VOXEL_LR else None
                lib.call("pvdb_dense_adam", lib.ptr(param), lib.ptr(grad), lib.ptr(st["exp_avg"]),
                         lib.ptr(st["exp_avg_sq"]), lib.ptr(rate), param.numel(), mode,
                         render_utils.adam_step_size(group["lr"], b1, b2, st["step"]),
                         float(group["eps"]), float(b1), float(b2), lib.stream())
