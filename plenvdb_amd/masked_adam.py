"""Optimizers of the hot path: `VDBAdam` for the two VDB grids and
`MaskedAdam` for the MLP (plenvdb/lib/masked_adam.py:18-158)."""
from __future__ import annotations

import torch

from . import render_utils
from .plenvdb import ColorOpt, DensityOpt


class VDBAdam:
    """masked_adam.py:18-100.  params: two groups {'params': VDBGrid, 'lr',
    'skip_zero_grad'}; stepmode 2 with per-voxel lr, 1 when skipping zero
    gradients, else 0 (masked_adam.py:56-68)."""

    def __init__(self, params, lr=1e-3, betas=(0.9, 0.99), eps=1e-8):
        if not 0.0 <= lr:
            raise ValueError(f"Invalid learning rate: {lr}")
        if not 0.0 <= eps:
            raise ValueError(f"Invalid epsilon value: {eps}")
        for i in (0, 1):
            if not 0.0 <= betas[i] < 1.0:
                raise ValueError(f"Invalid beta parameter at index {i}: {betas[i]}")
        assert len(params) == 2
        self.has_per_lr = False
        self.skip_density, self.skip_color = False, False
        for group in params:
            param = group["params"]
            glr = group["lr"]
            if param.channels == 1:
                self.densityparam = param
                self.densityOpt = DensityOpt(param.grid, glr, eps, betas[0], betas[1])
                self.skip_density = group["skip_zero_grad"]
            else:
                self.colorparam = param
                self.colorOpt = ColorOpt(param.grid, glr, eps, betas[0], betas[1])
                self.skip_color = group["skip_zero_grad"]

    def set_pervoxel_lr(self, count):
        self.has_per_lr = True
        self.densityOpt.set_pervoxel_lr((count.float() / count.max()).reshape(-1))

    def update_lr(self, factor):
        self.densityOpt.update_lr(factor)
        self.colorOpt.update_lr(factor)

    def zero_grad(self):
        self.densityOpt.zero_grad()
        self.colorOpt.zero_grad()

    def step(self):
        if self.has_per_lr:
            self.densityOpt.step(2)
        elif self.skip_density:
            self.densityOpt.step(1)
        else:
            self.densityOpt.step(0)
        self.colorOpt.step(1 if self.skip_color else 0)

    def save_to(self, savepath):
        torch.save({
            "dlr": self.densityOpt.getLr(), "clr": self.colorOpt.getLr(),
            "deps": self.densityOpt.getEps(), "ceps": self.colorOpt.getEps(),
            "dbeta0": self.densityOpt.getBeta0(), "cbeta0": self.colorOpt.getBeta0(),
            "dbeta1": self.densityOpt.getBeta1(), "cbeta1": self.colorOpt.getBeta1(),
            "dstep": self.densityOpt.getStep(), "cstep": self.colorOpt.getStep(),
        }, savepath + "vdbopt.tar")
        self.densityOpt.save_to(savepath + "density_")
        self.colorOpt.save_to(savepath + "color_")

    def load_from(self, loadpath):
        ck = torch.load(loadpath + "vdbopt.tar", weights_only=True)
        d, c = self.densityOpt, self.colorOpt
        d.setStep(ck["dstep"]); d.setLr(ck["dlr"]); d.setBeta0(ck["dbeta0"])
        d.setBeta1(ck["dbeta1"]); d.setEps(ck["deps"])
        c.setStep(ck["cstep"]); c.setLr(ck["clr"]); c.setBeta0(ck["cbeta0"])
        c.setBeta1(ck["cbeta1"]); c.setEps(ck["ceps"])
        d.load_from(loadpath + "density_")
        c.load_from(loadpath + "color_")


class MaskedAdam(torch.optim.Optimizer):
    """masked_adam.py:105-158 on libplenvdb_hip.so's dense Adam kernel."""

    def __init__(self, params, lr=1e-3, betas=(0.9, 0.99), eps=1e-8):
        if not 0.0 <= lr:
            raise ValueError(f"Invalid learning rate: {lr}")
        if not 0.0 <= eps:
            raise ValueError(f"Invalid epsilon value: {eps}")
        for i in (0, 1):
            if not 0.0 <= betas[i] < 1.0:
                raise ValueError(f"Invalid beta parameter at index {i}: {betas[i]}")
        defaults = dict(lr=lr, betas=betas, eps=eps, skip_zero_grad=False)
        self.per_lr = None
        super().__init__(params, defaults)

    def set_pervoxel_lr(self, count):
        assert self.param_groups[0]["params"][0].shape == count.shape
        self.per_lr = count.float() / count.max()

    @torch.no_grad()
    def step(self):
        for group in self.param_groups:
            lr = group["lr"]
            beta1, beta2 = group["betas"]
            eps = group["eps"]
            skip = group["skip_zero_grad"]
            for param in group["params"]:
                if param.grad is None:
                    continue
                state = self.state[param]
                if len(state) == 0:
                    state["step"] = 0
                    state["exp_avg"] = torch.zeros_like(param, memory_format=torch.preserve_format)
                    state["exp_avg_sq"] = torch.zeros_like(param, memory_format=torch.preserve_format)
                state["step"] += 1
                g = param.grad.contiguous()
                if self.per_lr is not None and param.shape == self.per_lr.shape:
                    render_utils.adam_upd_with_perlr(param, g, state["exp_avg"], state["exp_avg_sq"],
                                                     self.per_lr, state["step"], beta1, beta2, lr, eps)
                elif skip:
                    render_utils.masked_adam_upd(param, g, state["exp_avg"], state["exp_avg_sq"],
                                                 state["step"], beta1, beta2, lr, eps)
                else:
                    render_utils.adam_upd(param, g, state["exp_avg"], state["exp_avg_sq"],
                                          state["step"], beta1, beta2, lr, eps)
