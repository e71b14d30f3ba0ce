"""ctypes binding of libplenvdb_hip.so (declarations in include/plenvdb_hip.h).

The library is loaded from plenvdb_amd/_lib.  There is no fallback: when the
library is missing or a call fails, a RuntimeError is raised.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libplenvdb_hip.so")

_lock = threading.Lock()
_lib = None

P = C.c_void_p
I32 = C.c_int32
I64 = C.c_int64
F32 = C.c_float
F64 = C.c_double

# name -> argtypes (every function returns int32 status)
_SIGS = {
    "pvdb_abi_version": [],
    "pvdb_topo_create_from_mask": [P, I32, I32, I32, C.POINTER(P)],
    "pvdb_topo_create_from_leaves": [P, P, I64, I32, C.POINTER(P)],
    "pvdb_topo_destroy": [P],
    "pvdb_topo_info": [P, C.POINTER(I64), C.POINTER(I64), P, P, P],
    "pvdb_topo_leaves": [P, P, P],
    "pvdb_topo_slots": [P, P, I64, P, P],
    "pvdb_topo_device_arrays": [P, P, P, P],
    "pvdb_trilinear_fwd": [P, P, I32, I32, P, P, P, I64, P, P],
    "pvdb_trilinear_bwd": [P, P, I32, I32, P, P, P, P, I64, P],
    "pvdb_copy_from_dense": [P, P, I32, P, I32, I32, I32, P],
    "pvdb_copy_to_dense": [P, P, I32, P, I32, I32, I32, F32, P],
    "pvdb_set_values_by_mask": [P, P, P, F32, I32, I32, I32, P],
    "pvdb_zero_active": [P, P, I32, P],
    "pvdb_tv_add_grad": [P, P, P, I32, I32, I32, I32, F32, F32, F32, I32, P],
    "pvdb_vdb_adam": [P, P, P, P, P, P, I32, I32, I32, F32, F32, F32, F32, P],
    "pvdb_dense_adam": [P, P, P, P, P, I64, I32, F32, F32, F32, F32, P],
    "pvdb_rays_of_view": [I32, I32, P, P, I32, I32, I32, I32, I32, P, P, P, P, P],
    "pvdb_ray_bounds": [P, P, I64, P, P, F32, F32, P, F32, P, P, P, P],
    "pvdb_sample_pts": [P, P, P, P, I64, I64, F32, P, P, P, P, P, P, P],
    "pvdb_mask_lookup": [P, I32, I32, I32, P, I64, P, P, P, P],
    "pvdb_active_lookup": [P, P, I64, P, P, P, P],
    "pvdb_wld2idx": [P, I64, P, P, P, P, P, P, P],
    "pvdb_raw2alpha": [P, I64, F32, F32, P, P, P],
    "pvdb_raw2alpha_bwd": [P, P, I64, F32, P, P],
    "pvdb_segment_bounds": [P, I64, P, P, P],
    "pvdb_alpha2weight": [P, I64, P, P, P, P, P, P],
    "pvdb_alpha2weight_bwd": [P, P, P, P, P, P, I64, P, P, P, P],
    "pvdb_mfma_selftest": [P, P, P, P, P, P, P],
    "pvdb_trilinear_fwd_f64": [P, P, I32, P, P, P, I64, P, P],
    "pvdb_trilinear_bwd_f64": [P, P, I32, P, P, P, P, I64, P],
    "pvdb_raw2alpha_f64": [P, I64, F64, F64, P, P, P],
    "pvdb_raw2alpha_bwd_f64": [P, P, I64, F64, P, P],
    "pvdb_alpha2weight_f64": [P, I64, P, P, P, P, P, P],
    "pvdb_alpha2weight_bwd_f64": [P, P, P, P, P, P, I64, P, P, P, P],
}

# functions with extra declarations added by later modules register here
EXTRA_SIGS: dict[str, list] = {}


def declared_symbols() -> list[str]:
    return sorted(set(_SIGS) | set(EXTRA_SIGS) |
                  {"pvdb_last_error", "pvdb_fused_workspace_bytes", "pvdb_fused_stage_name"})


def load() -> C.CDLL:
    """Loads the library once (torch is imported first so both share one HIP runtime)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
                " (or python plenvdb_amd/build.py)")
        lib = C.CDLL(LIB_PATH)
        lib.pvdb_last_error.restype = C.c_char_p
        lib.pvdb_last_error.argtypes = []
        for name, args in {**_SIGS, **EXTRA_SIGS}.items():
            fn = getattr(lib, name)
            fn.restype = C.c_int32
            fn.argtypes = args
        _lib = lib
        return lib


class CallTimer:
    """Brackets every library call made while active with HIP events on torch's
    current stream (the stream the calls launch on); `totals()` gives the
    device time per entry point in ms after a synchronize."""

    active: "CallTimer | None" = None

    def __init__(self):
        self.records: list[tuple[str, torch.cuda.Event, torch.cuda.Event]] = []

    def __enter__(self):
        CallTimer.active = self
        return self

    def __exit__(self, *exc):
        CallTimer.active = None

    def totals(self) -> dict[str, tuple[float, int]]:
        out: dict[str, list] = {}
        for name, a, b in self.records:
            t = out.setdefault(name, [0.0, 0])
            t[0] += a.elapsed_time(b)
            t[1] += 1
        return {k: (v[0], v[1]) for k, v in out.items()}


def function(name: str):
    """The library function `name` with its declared argument types (modules that
    register signatures after the library was loaded get them applied here)."""
    lib = load()
    fn = getattr(lib, name)
    if fn.argtypes is None:
        sig = _SIGS.get(name, EXTRA_SIGS.get(name))
        if sig is None:
            raise RuntimeError(f"{name}: no declared signature")
        fn.restype = C.c_int32
        fn.argtypes = sig
    return fn


def call(name: str, *args) -> None:
    """Calls a library function and raises RuntimeError on a non-zero status."""
    lib = load()
    fn = function(name)
    timer = CallTimer.active
    if timer is not None:
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        rc = fn(*args)
        e1.record()
        timer.records.append((name, e0, e1))
    else:
        rc = fn(*args)
    if rc != 0:
        msg = lib.pvdb_last_error().decode(errors="replace")
        raise RuntimeError(f"{name} failed ({rc}): {msg}")


def ptr(t: torch.Tensor | None) -> C.c_void_p:
    if t is None:
        return C.c_void_p(None)
    return C.c_void_p(t.data_ptr())


def hptr(a) -> C.c_void_p:
    """Pointer of a C-contiguous numpy array (host)."""
    if a is None:
        return C.c_void_p(None)
    return C.c_void_p(a.ctypes.data)


def stream(device: torch.device | None = None) -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def f3(vals) -> C.Array:
    a = (F32 * 3)()
    for i, v in enumerate(vals):
        a[i] = float(v)
    return a


def i3(vals) -> C.Array:
    a = (I32 * 3)()
    for i, v in enumerate(vals):
        a[i] = int(v)
    return a


def require_cuda(*tensors: torch.Tensor) -> None:
    for t in tensors:
        if t is not None and not t.is_cuda:
            raise RuntimeError("plenvdb_amd kernels need device tensors (got a CPU tensor)")
