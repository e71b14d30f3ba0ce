"""Training-step drivers over a synthetic scene (see scenes.py).

`PlainTrainer` runs the reference's iteration (plenvdb/run.py:541-598) on the
stage-per-kernel model of dvgo.py: forward, zero_grad, the loss (`step_loss`), backward, the
MLP's MaskedAdam step, the VDB Adam step (skip-zero-grad for both grids, as
the fine stage configures, configs/default.py:74) and the lr decay.
"""
from __future__ import annotations

import numpy as np
import torch

from . import scenes
from .dvgo import DirectVoxGO
from .masked_adam import MaskedAdam, VDBAdam
from .topology import Topology


def build_topology(scene: dict) -> Topology:
    return Topology.from_mask(scene["mask"])


def init_values(scene: dict, topo: Topology):
    """Seeded per-slot density / colour values for a topology."""
    den, col = scenes.grid_values(scene["cfg"], topo.nslots)
    return den, col


def step_loss(out: dict, target: torch.Tensor, weight_main: float = 1.0,
              weight_entropy_last: float = 0.0, weight_rgbper: float = 0.0) -> torch.Tensor:
    """The training loss of run.py:551-574 from a forward's outputs: weighted
    MSE, background entropy of alphainv_last and the per-sample rgb loss with
    the weights detached (nearclip and distortion are off in every configuration)."""
    loss = weight_main * torch.nn.functional.mse_loss(out["rgb_marched"], target)
    if weight_entropy_last > 0:
        pout = out["alphainv_last"].clamp(1e-6, 1 - 1e-6)
        ent = -(pout * torch.log(pout) + (1 - pout) * torch.log(1 - pout)).mean()
        loss = loss + weight_entropy_last * ent
    if weight_rgbper > 0:
        rgbper = (out["raw_rgb"] - target[out["ray_id"]]).pow(2).sum(-1)
        loss = loss + weight_rgbper * (rgbper * out["weights"].detach()).sum() / len(target)
    return loss


class PlainTrainer:
    def __init__(self, scene: dict, topo: Topology | None = None, lr_grid=0.1, lr_mlp=1e-3,
                 lrate_decay=20, weight_main: float = 1.0, weight_entropy_last: float = 0.0,
                 weight_rgbper: float = 0.0, device="cuda"):
        self.scene = scene
        self.weights = (float(weight_main), float(weight_entropy_last), float(weight_rgbper))
        cfg = scene["cfg"]
        self.topo = topo if topo is not None else build_topology(scene)
        self.model = DirectVoxGO(scene["xyz_min"], scene["xyz_max"], scene["reso"],
                                 alpha_init=cfg.alpha_init, fast_color_thres=cfg.fast_color_thres,
                                 rgbnet_dim=cfg.channels if cfg.use_mlp else 0,
                                 voxel_size=scene["voxel_size"], topology=self.topo,
                                 mlp_params=scene.get("mlp"), k0_dim=cfg.channels).to(device)
        den, col = init_values(scene, self.topo)
        self.model.density.grid.grid.copy_(torch.from_numpy(den).reshape(-1, 1))
        self.model.k0.grid.grid.copy_(torch.from_numpy(col))
        self.vdbopt = VDBAdam([
            {"params": self.model.density, "lr": lr_grid, "skip_zero_grad": True},
            {"params": self.model.k0, "lr": lr_grid, "skip_zero_grad": True},
        ])
        self.opt = None
        if self.model.rgbnet is not None:
            self.opt = MaskedAdam([{"params": self.model.rgbnet.parameters(), "lr": lr_mlp,
                                    "skip_zero_grad": False}])
        self.decay = 0.1 ** (1 / (lrate_decay * 1000))
        dev = torch.device(device)
        self.rays_o = torch.from_numpy(scene["rays_o"]).to(dev)
        self.rays_d = torch.from_numpy(scene["rays_d"]).to(dev)
        self.viewdirs = torch.from_numpy(scene["viewdirs"]).to(dev)
        self.target = torch.from_numpy(scene["target"]).to(dev)
        fpr = scene.get("far_per_ray")
        self.far_per_ray = None if fpr is None else torch.from_numpy(fpr).to(dev)
        self.render_kwargs = dict(near=scene["near"], far=1e9, stepsize=scene["stepsize"],
                                  bg=scene["bg"], far_per_ray=self.far_per_ray)
        self.last = None

    def forward_backward(self):
        out = self.model(self.rays_o, self.rays_d, self.viewdirs, **self.render_kwargs)
        if self.opt is not None:
            self.opt.zero_grad(set_to_none=True)
        self.vdbopt.zero_grad()
        loss = step_loss(out, self.target, *self.weights)
        loss.backward()
        self.last = out
        return loss, out

    # -- checkpoints (run.py:404-424 resume, :611-626 save at the end of a stage) --
    def save_checkpoint(self, prefix: str, global_step: int) -> None:
        """The reference's end-of-stage save: model.save_to and vdbopt.save_to
        under `prefix` (density.vdb, color.vdb, the moments and vdbopt.tar), and
        `prefix + 'last.tar'` with the step, the MLP state and its optimizer
        state (run.py:611-626; fine_last.tar there)."""
        self.model.save_to(prefix)
        self.vdbopt.save_to(prefix)
        mlp = None if self.model.rgbnet is None else self.model.rgbnet.state_dict()
        opt = None if self.opt is None else self.opt.state_dict()
        torch.save({"global_step": int(global_step), "rgbnet_state_dict": mlp,
                    "optimizer_state_dict": opt}, prefix + "last.tar")

    def load_checkpoint(self, prefix: str) -> int:
        """Resume (run.py:404-424 with load_checkpoint): the grids and their
        optimizer from `prefix`, the MLP and its optimizer from last.tar.
        Returns the saved global step."""
        self.model.load_from(prefix)
        # the grids' optimizers follow the (possibly pruned) loaded topologies
        lr_d, lr_c = self.vdbopt.densityOpt.getLr(), self.vdbopt.colorOpt.getLr()
        self.vdbopt = VDBAdam([
            {"params": self.model.density, "lr": lr_d, "skip_zero_grad": True},
            {"params": self.model.k0, "lr": lr_c, "skip_zero_grad": True},
        ])
        self.vdbopt.load_from(prefix)
        ck = torch.load(prefix + "last.tar", weights_only=True)
        if self.model.rgbnet is not None:
            self.model.rgbnet.load_state_dict(ck["rgbnet_state_dict"])
            self.opt.load_state_dict(ck["optimizer_state_dict"])
        return int(ck["global_step"])

    def step(self):
        loss, out = self.forward_backward()
        if self.opt is not None:
            self.opt.step()
        self.vdbopt.step()
        if self.opt is not None:
            for g in self.opt.param_groups:
                g["lr"] = g["lr"] * self.decay
        self.vdbopt.update_lr(self.decay)
        return loss, out
