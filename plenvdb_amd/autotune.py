"""Launch-shape autotuner of the fused training step.

Four launch choices of `pvdb_fused_step` are free (`pvdb_step_desc.launch`):
the forward MLP's workgroup count, the scatter's workgroup count, the
scatter's records per chunk (32 or 64) and the rays (wavefronts) per workgroup
of the wave-per-ray kernels (the two marches and the composite).  None changes
the computed values beyond the order of the scatter's float atomics.  `tune`
picks them per device and workload by coordinate descent: each candidate is
timed on the stages it affects (the step's own HIP-event stage timers, several
forward+backward passes per candidate) with the other choices held at their
current best.  Choices are cached per (device, rays, capacity, leaves, colour
type, model) for the process and, when PVDB_TUNE_CACHE names a JSON file,
across processes.
"""
from __future__ import annotations

import ctypes as C
import json
import os

import torch

from . import lib

CANDIDATES = {
    0: (256, 512, 768, 1024),         # forward-MLP workgroups
    2: (32, 64),                      # scatter records per chunk
    1: (512, 768, 1024, 1536, 2048),  # scatter workgroups
    3: (1, 2, 4, 8),                  # rays per workgroup of the wave-per-ray kernels
}
STAGE = {0: ("mlp_fwd",), 1: ("scatter",), 2: ("scatter",),
         3: ("ray_count", "ray_march", "composite")}
NAMES = {0: "mlp_fwd_workgroups", 1: "scatter_workgroups", 2: "scatter_chunk_records",
         3: "ray_waves"}

_cache: dict[str, list[int]] = {}


def defaults() -> list[int]:
    """The built-in launch shapes on the current device."""
    arr = (C.c_int32 * 4)()
    lib.call("pvdb_fused_launch_defaults", arr)
    return [int(v) for v in arr]


def _key(tr) -> str:
    dev = torch.cuda.get_device_properties(torch.cuda.current_device())
    return (f"{dev.name}|{getattr(dev, 'gcnArchName', '')}|rays={tr.n_rays}|cap={tr.capacity}"
            f"|leaves={tr.topo.nleaf}|bf16={int(tr.color_bf16)}|k0={tr.k0_dim}")


def _load_file() -> dict:
    path = os.environ.get("PVDB_TUNE_CACHE")
    if not path or not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def _save_file(entries: dict) -> None:
    path = os.environ.get("PVDB_TUNE_CACHE")
    if not path:
        return
    data = _load_file()
    data.update(entries)
    with open(path, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)


def _stage_ms(tr, stages: tuple, reps: int) -> float:
    """Mean time per step of the named stages (summed)."""
    from .fused import stage_timing

    tr.forward_backward()  # warm-up with this choice
    torch.cuda.synchronize()
    stage_timing(1)
    for _ in range(reps):
        tr.forward_backward()
    torch.cuda.synchronize()
    st = stage_timing(0)
    return sum(st[s][0] / max(st[s][1], 1) for s in stages)


def apply(tr, launch: list[int]) -> None:
    for i in range(4):
        tr.desc.launch[i] = int(launch[i])


def tune(tr, reps: int = 3, force: bool = False) -> dict:
    """Chooses and applies the launch shapes for trainer `tr`; returns them by
    name with the per-candidate stage times (ms) of this tuning run (empty when
    the choice came from a cache).  Gradients accumulated while timing are
    cleared."""
    key = _key(tr)
    if not force:
        if key not in _cache:
            hit = _load_file().get(key)
            if hit is not None:
                _cache[key] = [int(v) for v in hit]
        if key in _cache:
            apply(tr, _cache[key])
            return {"launch": {NAMES[i]: _cache[key][i] for i in NAMES}, "source": "cache"}
    best = defaults()
    apply(tr, best)
    timings = {}
    # the chunk length before the scatter's workgroup count; a model without the
    # MLP has no forward-MLP launch to shape
    for idx in ((2, 1, 3) if tr.k0_dim else (0, 2, 1, 3)):
        t = {}
        for v in CANDIDATES[idx]:
            tr.desc.launch[idx] = v
            t[v] = _stage_ms(tr, STAGE[idx], reps)
        best[idx] = min(t, key=t.get)
        tr.desc.launch[idx] = best[idx]
        timings[NAMES[idx]] = {str(k): round(v, 5) for k, v in t.items()}
    tr.zero_grad()
    _cache[key] = list(best)
    _save_file({key: list(best)})
    return {"launch": {NAMES[i]: best[i] for i in NAMES}, "source": "tuned",
            "stage_ms": timings}
