"""Fused (batched) training step on the device: `FusedTrainer`.

Holds the grids (density fp32, colour fp32 or bf16), their gradients and Adam
moments, the MLP master parameters (fp32) and a workspace, and runs one
iteration of the reference's training loop (plenvdb/run.py:541-598) per call
through `pvdb_fused_step` -- a fixed sequence of gfx950 kernels with no host
synchronisation, optionally replayed from a captured hipGraph.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import lib, scenes
from .topology import Topology

MLP_PARAMS = 39 * 128 + 128 + 128 * 128 + 128 + 128 * 3 + 3


class StepDesc(C.Structure):
    """Mirror of pvdb_step_desc (include/plenvdb_hip.h)."""

    _fields_ = [
        ("rays_o", C.c_void_p), ("rays_d", C.c_void_p), ("viewdirs", C.c_void_p),
        ("target", C.c_void_p), ("far_per_ray", C.c_void_p),
        ("n_rays", C.c_int32), ("color_bf16", C.c_int32),
        ("xyz_min", C.c_float * 3), ("xyz_max", C.c_float * 3), ("reso", C.c_int32 * 3),
        ("near_", C.c_float), ("stepdist", C.c_float), ("act_shift", C.c_float),
        ("interval", C.c_float), ("fast_color_thres", C.c_float), ("bg", C.c_float),
        ("density", C.c_void_p), ("grad_density", C.c_void_p), ("density_m", C.c_void_p),
        ("density_v", C.c_void_p),
        ("color", C.c_void_p), ("grad_color", C.c_void_p), ("color_m", C.c_void_p),
        ("color_v", C.c_void_p),
        ("mlp", C.c_void_p), ("mlp_grad", C.c_void_p), ("mlp_m", C.c_void_p), ("mlp_v", C.c_void_p),
        ("stepsz_density", C.c_float), ("stepsz_color", C.c_float), ("step_size_mlp", C.c_float),
        ("eps", C.c_float), ("beta0", C.c_float), ("beta1", C.c_float),
        ("do_adam", C.c_int32), ("k0_dim", C.c_int32),
        ("workspace", C.c_void_p), ("workspace_bytes", C.c_int64), ("capacity", C.c_int64),
        ("ray_out", C.c_void_p), ("loss", C.c_void_p), ("adam_scalars", C.c_void_p),
        ("launch", C.c_int32 * 4),
        ("weight_main", C.c_float), ("weight_entropy_last", C.c_float),
        ("weight_rgbper", C.c_float), ("grad_stride", C.c_int32),
        ("color_master", C.c_void_p),
        ("density_lr", C.c_void_p), ("stepmode_density", C.c_int32), ("stepmode_color", C.c_int32),
    ]


lib.EXTRA_SIGS.update({
    "pvdb_fused_timing": [C.c_int32, C.c_void_p, C.c_void_p, C.c_int32],
    "pvdb_fused_init": [C.c_void_p, C.POINTER(StepDesc), C.c_void_p],
    "pvdb_fused_step": [C.c_void_p, C.POINTER(StepDesc), C.c_void_p],
    "pvdb_fused_launch_defaults": [C.c_void_p],
    "pvdb_fused_records": [C.c_void_p, C.POINTER(StepDesc)] + [C.c_void_p] * 7,
})


def _workspace_bytes(n_rays: int, capacity: int, nleaf: int) -> int:
    L = lib.load()
    fn = L.pvdb_fused_workspace_bytes
    fn.restype = C.c_int64
    fn.argtypes = [C.c_int64, C.c_int64, C.c_int64]
    b = int(fn(n_rays, capacity, nleaf))
    if b < 0:
        raise RuntimeError("bad workspace sizes")
    return b


def pack_mlp(p: dict) -> np.ndarray:
    """W0[39][128] b0 W1[128][128] b1 W2[128][3] b2 (W as Linear.weight.T)."""
    return np.concatenate([np.asarray(p[k], np.float32).ravel()
                           for k in ("w0", "b0", "w1", "b1", "w2", "b2")])


def unpack_mlp(v: np.ndarray) -> dict:
    v = np.asarray(v)
    out, o = {}, 0
    for k, shp in (("w0", (39, 128)), ("b0", (128,)), ("w1", (128, 128)), ("b1", (128,)),
                   ("w2", (128, 3)), ("b2", (3,))):
        n = int(np.prod(shp))
        out[k] = v[o:o + n].reshape(shp)
        o += n
    return out


def _f32(x) -> float:
    return float(np.float32(x))


class FusedTrainer:
    """One training iteration per `step()` on the fused kernels.

    scene: scenes.make_scene(...) output.  A configuration with the MLP runs
    the fine-stage model (12 colour features, rgbnet); one without it runs
    rgb = sigmoid(k0) over its colour grid's channels (3, the coarse stage's
    k0, or 12; dvgo.py:72-75,344-346).  color_bf16 defaults to the
    configuration's choice.  weight_main, weight_entropy_last and
    weight_rgbper are the loss weights of run.py:551-574 (the fine stage of
    configs/default.py:73-74 uses 1, 0.001 and 0.01, the coarse stage
    configs/default.py:52-55 uses 1, 0.01 and 0.1; the defaults here are MSE
    only).  stepmode_density / stepmode_color are VDBAdam's stepmodes
    (masked_adam.py:56-68): 1 is skip_zero_grad (the fine stage, the
    default), 0 updates every active voxel, density 2 applies the per-voxel
    rate of `set_pervoxel_lr` (the coarse stage: density 2, colour 0).
    """

    def __init__(self, scene: dict, topo: Topology | None = None, color_bf16: bool | None = None,
                 lr_grid=0.1, lr_mlp=1e-3, lrate_decay=20, capacity: int | None = None,
                 weight_main: float = 1.0, weight_entropy_last: float = 0.0,
                 weight_rgbper: float = 0.0, color_master: bool = True,
                 stepmode_density: int = 1, stepmode_color: int = 1, device="cuda"):
        cfg = scene["cfg"]
        self.k0_dim = 0 if cfg.use_mlp else int(cfg.channels)
        if self.k0_dim not in (0, 3, 12):
            raise ValueError("a model without the MLP takes 3 or 12 colour channels")
        if stepmode_density not in (0, 1, 2) or stepmode_color not in (0, 1):
            raise ValueError("stepmode_density must be 0, 1 or 2 and stepmode_color 0 or 1")
        self.stepmodes = (int(stepmode_density), int(stepmode_color))
        self.scene = scene
        dev = torch.device(device)
        self.topo = topo if topo is not None else Topology.from_mask(scene["mask"])
        self.color_bf16 = cfg.color_bf16 if color_bf16 is None else bool(color_bf16)
        if self.k0_dim and self.color_bf16:
            raise ValueError("a model without the MLP takes an fp32 colour grid")
        nc = self.k0_dim or 12  # colour channels per slot
        self.n_out = self.k0_dim or 3  # rendered channels
        ns = self.topo.nslots
        den, col = scenes.grid_values(cfg, ns)
        self.density = torch.from_numpy(den).to(dev)
        colt = torch.from_numpy(col).to(dev)
        self.color = colt.to(torch.bfloat16).contiguous() if self.color_bf16 else colt
        # the bf16 payload's fp32 master, which the colour Adam updates
        self.color_master = colt if (self.color_bf16 and color_master) else None
        # the gradients as one row per slot (colour, then density, padded to 16
        # bytes): a voxel's gradients share a cache line for the step's atomics
        self.grad_row = 16 if nc == 12 else 4
        self.grad_rows = torch.zeros(ns, self.grad_row, device=dev)
        self.grad_color = self.grad_rows[:, :nc]
        self.grad_density = self.grad_rows[:, nc]
        self.m_density = torch.zeros(ns, device=dev)
        self.v_density = torch.zeros(ns, device=dev)
        self.m_color = torch.zeros(ns, nc, device=dev)
        self.v_color = torch.zeros(ns, nc, device=dev)
        self.density_lr = None  # per-voxel rate of density stepmode 2
        if self.k0_dim:
            self.mlp = self.mlp_grad = self.mlp_m = self.mlp_v = None
        else:
            self.mlp = torch.from_numpy(pack_mlp(scene["mlp"])).to(dev)
            self.mlp_grad = torch.zeros_like(self.mlp)
            self.mlp_m = torch.zeros_like(self.mlp)
            self.mlp_v = torch.zeros_like(self.mlp)
        self.rays_o = torch.from_numpy(scene["rays_o"]).to(dev).contiguous()
        self.rays_d = torch.from_numpy(scene["rays_d"]).to(dev).contiguous()
        self.viewdirs = torch.from_numpy(scene["viewdirs"]).to(dev).contiguous()
        self.target = torch.from_numpy(scene["target"]).to(dev).contiguous()
        fpr = scene.get("far_per_ray")
        self.far_per_ray = None if fpr is None else torch.from_numpy(fpr).to(dev).contiguous()
        self.n_rays = self.rays_o.shape[0]
        self.stepdist = _f32(np.float32(scene["stepsize"]) * np.float32(scene["voxel_size"]))
        if capacity is None:
            capacity = self._capacity_bound()
        self.capacity = int(capacity)
        self.ws_bytes = _workspace_bytes(self.n_rays, self.capacity, self.topo.nleaf)
        self.workspace = torch.empty(self.ws_bytes, dtype=torch.uint8, device=dev)
        self.ray_out = torch.zeros(self.n_rays, self.n_out + 1, device=dev)
        self.loss = torch.zeros(1, device=dev)
        self.lr_grid = float(lr_grid)
        self.lr_mlp = float(lr_mlp)
        self.decay = 0.1 ** (1 / (lrate_decay * 1000))
        self.step_count = 0
        # per-step Adam step sizes, read on the device (a captured step replays them)
        self.adam_scalars = torch.zeros(4, device=dev)
        self.weights = (float(weight_main), float(weight_entropy_last), float(weight_rgbper))
        self.desc = self._make_desc()
        lib.call("pvdb_fused_init", self.topo.handle, C.byref(self.desc), lib.stream())
        self.graph = None

    def _capacity_bound(self) -> int:
        """Upper bound of the records of a step: the total number of ray steps."""
        mn, mx = self.scene["xyz_min"], self.scene["xyz_max"]
        n = self.rays_o.shape[0]
        t_min = torch.empty(n, device=self.rays_o.device)
        t_max = torch.empty_like(t_min)
        n_steps = torch.empty(n, dtype=torch.int64, device=self.rays_o.device)
        lib.call("pvdb_ray_bounds", lib.ptr(self.rays_o), lib.ptr(self.rays_d), n, lib.f3(mn),
                 lib.f3(mx), float(self.scene["near"]), 1e9, lib.ptr(self.far_per_ray),
                 self.stepdist, lib.ptr(t_min), lib.ptr(t_max), lib.ptr(n_steps), lib.stream())
        # record offsets are int32 on the device; the step takes at most 2^31 - 2
        return min(int(n_steps.sum().item()) + 64, 2 ** 31 - 2)

    def _make_desc(self) -> StepDesc:
        sc = self.scene
        d = StepDesc()
        d.rays_o, d.rays_d = self.rays_o.data_ptr(), self.rays_d.data_ptr()
        d.viewdirs, d.target = self.viewdirs.data_ptr(), self.target.data_ptr()
        d.far_per_ray = self.far_per_ray.data_ptr() if self.far_per_ray is not None else None
        d.n_rays = self.n_rays
        d.color_bf16 = int(self.color_bf16)
        for a in range(3):
            d.xyz_min[a] = float(sc["xyz_min"][a])
            d.xyz_max[a] = float(sc["xyz_max"][a])
            d.reso[a] = int(sc["reso"][a])
        d.near_ = float(sc["near"])
        d.stepdist = self.stepdist
        d.act_shift = float(sc["act_shift"])
        d.interval = _f32(np.float32(sc["stepsize"]) * np.float32(sc["voxel_size_ratio"]))
        d.fast_color_thres = float(sc["fast_color_thres"])
        d.bg = float(sc["bg"])
        d.density, d.grad_density = self.density.data_ptr(), self.grad_density.data_ptr()
        d.grad_stride = self.grad_row
        d.density_m, d.density_v = self.m_density.data_ptr(), self.v_density.data_ptr()
        d.color, d.grad_color = self.color.data_ptr(), self.grad_color.data_ptr()
        d.color_m, d.color_v = self.m_color.data_ptr(), self.v_color.data_ptr()
        if self.mlp is not None:
            d.mlp, d.mlp_grad = self.mlp.data_ptr(), self.mlp_grad.data_ptr()
            d.mlp_m, d.mlp_v = self.mlp_m.data_ptr(), self.mlp_v.data_ptr()
        d.k0_dim = self.k0_dim
        d.stepmode_density, d.stepmode_color = self.stepmodes
        d.density_lr = None if self.density_lr is None else self.density_lr.data_ptr()
        d.eps, d.beta0, d.beta1 = 1e-8, 0.9, 0.99
        d.workspace = self.workspace.data_ptr()
        d.workspace_bytes = self.ws_bytes
        d.capacity = self.capacity
        d.ray_out = self.ray_out.data_ptr()
        d.loss = self.loss.data_ptr()
        d.adam_scalars = self.adam_scalars.data_ptr()
        d.weight_main, d.weight_entropy_last, d.weight_rgbper = self.weights
        d.color_master = None if self.color_master is None else self.color_master.data_ptr()
        return d

    @staticmethod
    def _stepsz(lr, t):
        f = np.float32
        return float(f(lr) * np.sqrt(f(1) - np.power(f(0.99), f(t))) / (f(1) - np.power(f(0.9), f(t))))

    def forward_backward(self):
        """Forward + backward only: accumulates gradients, no optimizer step."""
        self.desc.do_adam = 0
        lib.call("pvdb_fused_step", self.topo.handle, C.byref(self.desc), lib.stream())
        return self.loss

    def _next_step_sizes(self) -> list:
        """Advances the step counter; the Adam step sizes of this step (grid,
        grid, MLP), also written to the device buffer a captured step reads."""
        self.step_count += 1
        t = self.step_count
        sz = [self._stepsz(self.lr_grid, t), self._stepsz(self.lr_grid, t),
              self._stepsz(self.lr_mlp, t), 0.0]
        self.adam_scalars.copy_(torch.tensor(sz, dtype=torch.float32))
        self.desc.stepsz_density, self.desc.stepsz_color, self.desc.step_size_mlp = sz[:3]
        return sz

    def _decay_lr(self):
        # lr decay of run.py:592-598 (the VDB optimizers multiply in float)
        self.lr_grid = _f32(np.float32(self.lr_grid) * np.float32(self.decay))
        self.lr_mlp = self.lr_mlp * self.decay

    def step(self):
        """One full iteration: forward, backward, Adam (grids stepmode 1, MLP
        stepmode 0), gradients re-zeroed, lr decay (run.py:592-598).  Replays
        the captured graph when `capture_graph` was called."""
        self._next_step_sizes()
        if self.graph is not None:
            self.graph.replay()
        else:
            self.desc.do_adam = 1
            lib.call("pvdb_fused_step", self.topo.handle, C.byref(self.desc), lib.stream())
        self._decay_lr()
        return self.loss

    def adam_only(self):
        """The optimizer step alone on the gradients accumulated by earlier
        `forward_backward` calls (VDBAdam.step + MaskedAdam.step after a
        separate backward): grid Adam over the touched leaves with their
        gradients re-zeroed, MLP Adam on `mlp_grad`, then the lr decay."""
        self._next_step_sizes()
        self.desc.do_adam = 2
        lib.call("pvdb_fused_step", self.topo.handle, C.byref(self.desc), lib.stream())
        self._decay_lr()

    def capture_graph(self):
        """Captures one full step (do_adam) into a hipGraph; later `step()` calls
        replay it.  Run at least one eager step first (library set-up calls must
        not happen under capture)."""
        torch.cuda.synchronize()
        self.desc.do_adam = 1
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.graph(g, stream=s):
            lib.call("pvdb_fused_step", self.topo.handle, C.byref(self.desc), lib.stream())
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        self.graph = g

    def autotune(self, reps: int = 3, force: bool = False) -> dict:
        """Chooses the step's launch shapes for this device and workload (see
        plenvdb_amd.autotune); a captured graph is captured again with them."""
        from . import autotune

        out = autotune.tune(self, reps=reps, force=force)
        if self.graph is not None:
            self.capture_graph()
        return out

    def set_pervoxel_lr(self, count) -> None:
        """Per-voxel rate of density stepmode 2 from a dense [X, Y, Z] view-count
        grid: count / count.max() (masked_adam.py:43-46), copied onto the slots
        (plenvdb.h:713-719); the step then runs density stepmode 2."""
        c = torch.as_tensor(count, device=self.density.device).float()
        reso = [int(r) for r in self.scene["reso"]]
        if c.numel() != reso[0] * reso[1] * reso[2]:
            raise ValueError("set_pervoxel_lr: size does not match the resolution")
        c = (c / c.max()).reshape(-1).contiguous()
        self.density_lr = torch.zeros(self.topo.nslots, device=self.density.device)
        lib.call("pvdb_copy_from_dense", self.topo.handle, lib.ptr(self.density_lr), 1, lib.ptr(c),
                 *reso, lib.stream())
        torch.cuda.current_stream().synchronize()
        self.stepmodes = (2, self.stepmodes[1])
        self.desc.density_lr = self.density_lr.data_ptr()
        self.desc.stepmode_density = 2
        if self.graph is not None:
            self.capture_graph()

    # -- checkpoints: the plain path's file set (PlainTrainer.save_checkpoint) --
    def _rgbnet(self):
        """The MLP as DirectVoxGO's rgbnet module (its state_dict layout)."""
        from .dvgo import build_rgbnet

        net = build_rgbnet(39, 128, 3)
        p = unpack_mlp(self.mlp.detach().cpu().numpy())
        lin = [m for m in net.modules() if isinstance(m, torch.nn.Linear)]
        with torch.no_grad():
            for i, l in enumerate(lin):
                l.weight.copy_(torch.from_numpy(np.ascontiguousarray(p[f"w{i}"].T)))
                l.bias.copy_(torch.from_numpy(p[f"b{i}"]))
        return net, lin

    def save_to(self, prefix: str, global_step: int | None = None) -> None:
        """Writes density.vdb, color.vdb (the fp32 master of a bf16 grid), the
        grid moments and vdbopt.tar (masked_adam.py:83-100 keys) and last.tar
        with the MLP as rgbnet_state_dict plus its Adam state in torch's
        optimizer format, so PlainTrainer.load_checkpoint reads the same files."""
        from . import vdbio

        t, reso = self.topo, self.scene["reso"]
        col = self.color_master if self.color_master is not None else self.color.float()
        for name, val in (("density.vdb", self.density), ("color.vdb", col),
                          ("density_exp_avg.vdb", self.m_density),
                          ("density_exp_avg_sq.vdb", self.v_density),
                          ("color_exp_avg.vdb", self.m_color), ("color_exp_avg_sq.vdb", self.v_color)):
            vdbio.save_grid(prefix + name, t, val, reso)
        sc = {"lr": self.lr_grid, "eps": 1e-8, "beta0": 0.9, "beta1": 0.99, "step": self.step_count}
        torch.save({k + n: v for k in ("d", "c") for n, v in sc.items()}, prefix + "vdbopt.tar")
        if self.mlp is None:
            torch.save({"global_step": int(self.step_count if global_step is None else global_step),
                        "rgbnet_state_dict": None, "optimizer_state_dict": None},
                       prefix + "last.tar")
            return
        net, lin = self._rgbnet()
        pm = unpack_mlp(self.mlp_m.detach().cpu().numpy())
        pv = unpack_mlp(self.mlp_v.detach().cpu().numpy())
        state = {}
        for i in range(3):
            for j, (kw, tr) in enumerate(((f"w{i}", True), (f"b{i}", False))):
                conv = (lambda a: np.ascontiguousarray(a.T)) if tr else np.ascontiguousarray
                state[2 * i + j] = {"step": self.step_count,
                                    "exp_avg": torch.from_numpy(conv(pm[kw])),
                                    "exp_avg_sq": torch.from_numpy(conv(pv[kw]))}
        opt = {"state": state, "param_groups": [{"lr": self.lr_mlp, "betas": (0.9, 0.99),
                                                 "eps": 1e-8, "skip_zero_grad": False,
                                                 "params": list(range(6))}]}
        torch.save({"global_step": int(self.step_count if global_step is None else global_step),
                    "rgbnet_state_dict": net.state_dict(), "optimizer_state_dict": opt},
                   prefix + "last.tar")

    def load_from(self, prefix: str) -> int:
        """Reads the files save_to (or PlainTrainer.save_checkpoint) writes; the
        grids are pruned on load like the reference's, the moments follow them
        by leaf origin.  Returns the saved global step."""
        from . import vdbio

        nc = self.k0_dim or 12
        topo, den, _ = vdbio.load_grid(prefix + "density.vdb", 1)
        ctopo, col, _ = vdbio.load_grid(prefix + "color.vdb", nc)
        if ctopo.nleaf != topo.nleaf:  # a colour leaf pruned alone: keep the density's set
            col = vdbio.load_onto(prefix + "color.vdb", topo, nc)
        if topo.nleaf != self.topo.nleaf:
            raise ValueError("the checkpoint's leaf set differs from this trainer's topology")
        self.density.copy_(den)
        if self.color_master is not None:
            self.color_master.copy_(col)
        self.color.copy_(col.to(self.color.dtype))
        for name, dst in (("density_exp_avg", self.m_density), ("density_exp_avg_sq", self.v_density),
                          ("color_exp_avg", self.m_color), ("color_exp_avg_sq", self.v_color)):
            dst.copy_(vdbio.load_onto(prefix + name + ".vdb", self.topo, dst.numel() // self.topo.nslots))
        st = torch.load(prefix + "vdbopt.tar", weights_only=True)
        self.lr_grid = _f32(st["dlr"])
        self.step_count = int(st["dstep"])
        ck = torch.load(prefix + "last.tar", weights_only=True)
        if self.mlp is None:
            self.grad_density.zero_()
            self.grad_color.zero_()
            lib.call("pvdb_fused_init", self.topo.handle, C.byref(self.desc), lib.stream())
            return int(ck["global_step"])
        net, lin = self._rgbnet()
        net.load_state_dict(ck["rgbnet_state_dict"])
        vec = lambda ts: np.concatenate([  # noqa: E731
            np.ascontiguousarray(t.detach().cpu().numpy().T if t.dim() == 2 else t.detach().cpu().numpy()
                                 ).ravel() for t in ts]).astype(np.float32)
        self.mlp.copy_(torch.from_numpy(vec([x for l in lin for x in (l.weight, l.bias)])))
        opt = ck["optimizer_state_dict"]
        if opt is not None:
            order = sorted(opt["state"])
            self.mlp_m.copy_(torch.from_numpy(vec([opt["state"][k]["exp_avg"] for k in order])))
            self.mlp_v.copy_(torch.from_numpy(vec([opt["state"][k]["exp_avg_sq"] for k in order])))
            self.lr_mlp = float(opt["param_groups"][0]["lr"])
        self.grad_density.zero_()
        self.grad_color.zero_()
        lib.call("pvdb_fused_init", self.topo.handle, C.byref(self.desc), lib.stream())
        return int(ck["global_step"])

    def zero_grad(self):
        self.grad_rows.zero_()
        if self.mlp_grad is not None:
            self.mlp_grad.zero_()

    def overflowed(self) -> bool:
        """Whether the last step had more records than the capacity (the step
        then covered the first `capacity` of them; reads the count back)."""
        return self.total_records() > self.capacity

    def total_records(self) -> int:
        ptrs = [C.c_void_p() for _ in range(7)]
        lib.call("pvdb_fused_records", self.topo.handle, C.byref(self.desc), *[C.byref(p) for p in ptrs])
        return int(_read_i32(ptrs[6].value, 1)[0])

    def records(self) -> dict:
        """Host copies of the per-sample records and per-ray counts of the last step."""
        ptrs = [C.c_void_p() for _ in range(7)]
        lib.call("pvdb_fused_records", self.topo.handle, C.byref(self.desc), *[C.byref(p) for p in ptrs])
        torch.cuda.synchronize()
        n = self.n_rays
        total = _read_i32(ptrs[6].value, 1)[0]
        total = min(total, self.capacity)
        out = {"total": total, "count": _read_i32(ptrs[4].value, n), "offset": _read_i32(ptrs[5].value, n)}
        for k, p, w in zip(("xyze", "ab", "rgb", "adj"), ptrs[:4], (4, 2, 4, 4)):
            out[k] = _read_f32(p.value, w * total).reshape(total, w)
        return out


N_STAGES = 9


def stage_timing(enable: int) -> dict:
    """Collects the fused step's per-stage device times (ms totals, calls);
    enable 1 starts, 0 stops, -1 only reads (see pvdb_fused_timing)."""
    L = lib.load()
    L.pvdb_fused_stage_name.restype = C.c_char_p
    L.pvdb_fused_stage_name.argtypes = [C.c_int32]
    tot = (C.c_double * N_STAGES)()
    calls = (C.c_int32 * N_STAGES)()
    lib.call("pvdb_fused_timing", int(enable), tot, calls, N_STAGES)
    return {L.pvdb_fused_stage_name(i).decode(): (tot[i], calls[i]) for i in range(N_STAGES)}


def _wrap(ptr: int, n: int, typestr: str) -> torch.Tensor:
    """A CUDA tensor view of n elements at a raw device pointer (no ownership)."""

    class _Holder:
        __cuda_array_interface__ = {"shape": (n,), "typestr": typestr, "data": (ptr, False),
                                    "version": 3}

    return torch.as_tensor(_Holder(), device="cuda")


def _read_f32(ptr: int, n: int) -> np.ndarray:
    if n == 0:
        return np.zeros(0, np.float32)
    torch.cuda.current_stream().synchronize()
    return _wrap(ptr, n, "<f4").cpu().numpy()


def _read_i32(ptr: int, n: int) -> np.ndarray:
    torch.cuda.current_stream().synchronize()
    return _wrap(ptr, n, "<i4").cpu().numpy()
