"""MGRenderer: the merged-VDB renderer of the reference's `plenvdb` module
(plenvdb/lib/vdb/plenvdb.cpp:135-171, plenvdb.h:933-1068) on
`pvdb_render_image`.

Same constructor, setters and outputs: load_params takes the three Linear
layers as [in, out] row-major arrays (Linear.weight.T, run.py:98-104),
load_data the density [N] and colour [N*dcol] data of the merged export plus
the path of its index grid (run.py:86-95), setScene / setKwargs / input_a_c2w
the camera, and render_an_image renders only once all five have been called
(plenvdb.h:1029).  output_an_image returns the [H*W*3] float32 image.  Inputs
may be numpy arrays (host, copied like the reference does) or CUDA tensors.

Differences: the index grid is read from this package's grid container
(vdbio) rather than an OpenVDB file; the early-stop threshold is a parameter
(`set_t_stop`, default the reference's 1e-3); the MLP runs in fp16 with fp32
accumulation; the timer counts wall time around the render and its
synchronisation (the reference sums CPU clock() time, plenvdb.h:1030-1035).
The kept samples of a frame live in a device workspace sized in records; a
synchronous render that finds it too small grows it and renders again.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np
import torch

from . import lib, vdbio


class RenderDesc(C.Structure):
    """Mirror of pvdb_render_desc (include/plenvdb_hip.h)."""

    _fields_ = [
        ("c2w", C.c_float * 16), ("K", C.c_float * 9), ("xyz_min", C.c_float * 3),
        ("xyz_max", C.c_float * 3), ("reso", C.c_int32 * 3),
        ("H", C.c_int32), ("W", C.c_int32), ("inverse_y", C.c_int32),
        ("near_", C.c_float), ("stepdist", C.c_float), ("act_shift", C.c_float),
        ("interval", C.c_float), ("fast_color_thres", C.c_float), ("bg", C.c_float),
        ("t_stop", C.c_float), ("payload_f32", C.c_int32),
        ("payload", C.c_void_p), ("weights", C.c_void_p), ("out", C.c_void_p),
        ("counters", C.c_void_p), ("workspace", C.c_void_p), ("workspace_bytes", C.c_int64),
        ("capacity", C.c_int64),
    ]


lib.EXTRA_SIGS.update({
    "pvdb_render_pack_weights": [C.c_void_p] * 6 + [C.c_int32] * 4 + [C.c_void_p, C.c_void_p],
    "pvdb_render_build_payload": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                  C.c_int32, C.c_void_p, C.c_void_p],
    "pvdb_render_image": [C.c_void_p, C.POINTER(RenderDesc), C.c_void_p],
    "pvdb_render_image_noalloc": [C.c_void_p, C.POINTER(RenderDesc), C.c_void_p],
    "pvdb_render_frame_info": [C.POINTER(RenderDesc), C.c_void_p, C.c_void_p, C.c_void_p,
                               C.c_void_p],
    "pvdb_render_timing": [C.c_int32, C.c_void_p],
    "pvdb_mfma_selftest_f16": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
})


def _i64_fn(name, argtypes):
    fn = getattr(lib.load(), name)
    fn.restype = C.c_int64
    fn.argtypes = argtypes
    return fn


def _dev_f32(a, n: int | None = None, what: str = "array") -> torch.Tensor:
    if isinstance(a, torch.Tensor):
        t = a.detach().to(device="cuda", dtype=torch.float32).reshape(-1).contiguous()
    else:
        t = torch.from_numpy(np.ascontiguousarray(np.asarray(a, np.float32).reshape(-1))).cuda()
    if n is not None and t.numel() != n:
        raise ValueError(f"{what}: expected {n} values, got {t.numel()}")
    return t


def _fp16_exact(t: torch.Tensor) -> bool:
    h = t.to(torch.float16)
    return bool(torch.isfinite(h).all()) and bool((h.float() == t).all())


class MGRenderer:
    """plenvdb.MGRenderer(dcol, dpe, dhid, dout)."""

    def __init__(self, dcol: int, dpe: int, dhid: int, dout: int):
        if (dcol, dpe, dhid, dout) != (12, 27, 128, 3):
            raise ValueError("MGRenderer supports dcol 12, dpe 27, dhid 128, dout 3")
        self.Dcol, self.Dpe, self.Dhid, self.Dout = dcol, dpe, dhid, dout
        self.flags = [False] * 5  # load_data, load_params, setScene, setKwargs, input_a_c2w
        self.timer = 0.0
        self.t_stop = 1e-3  # renderer.cu:265
        self.desc = RenderDesc()
        self.weights = None
        self.payload = None
        self.topo = None
        self.image = None
        self.counters = torch.zeros(3, dtype=torch.int64, device="cuda")
        self.workspace = None
        self.capacity = 0
        self._ws_hw = None

    # -- setters (plenvdb.h:959-1027) ------------------------------------------------
    def load_params(self, w0, b0, w1, b1, w2, b2) -> None:
        din = self.Dcol + self.Dpe
        t = [_dev_f32(w0, din * self.Dhid, "w0"), _dev_f32(b0, self.Dhid, "b0"),
             _dev_f32(w1, self.Dhid * self.Dhid, "w1"), _dev_f32(b1, self.Dhid, "b1"),
             _dev_f32(w2, self.Dhid * self.Dout, "w2"), _dev_f32(b2, self.Dout, "b2")]
        nbytes = int(_i64_fn("pvdb_render_weights_bytes", [])())
        self.weights = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        lib.call("pvdb_render_pack_weights", *[lib.ptr(x) for x in t], self.Dcol, self.Dpe,
                 self.Dhid, self.Dout, lib.ptr(self.weights), lib.stream())
        torch.cuda.current_stream().synchronize()  # the fp32 inputs may be freed after
        self.flags[1] = True

    def load_data(self, ddata, cdata, vdbdir: str, N: int) -> None:
        topo, slot_idx, _ = vdbio.load_index_grid(vdbdir)
        self.load_data_grid(ddata, cdata, topo, slot_idx, N)

    def load_data_grid(self, ddata, cdata, topo, slot_idx: torch.Tensor, N: int) -> None:
        """load_data with the index grid given as a topology and its per-slot
        indices (vdbio.load_index_grid / index_grid_arrays) instead of a path."""
        den = _dev_f32(ddata, int(N), "density data")
        col = _dev_f32(cdata, int(N) * self.Dcol, "colour data")
        if not isinstance(slot_idx, torch.Tensor):
            slot_idx = torch.from_numpy(np.ascontiguousarray(slot_idx, np.int32))
        slot_idx = slot_idx.to(device="cuda", dtype=torch.int32).contiguous()
        if slot_idx.numel() != topo.nslots:
            raise ValueError("slot_idx must hold one index per slot of the topology")
        if slot_idx.numel() and (int(slot_idx.min()) < 0 or int(slot_idx.max()) >= int(N)):
            raise ValueError("index grid refers past the data arrays")
        self.payload_f32 = not (_fp16_exact(den) and _fp16_exact(col))
        nbytes = int(_i64_fn("pvdb_render_payload_bytes", [C.c_void_p, C.c_int32])(
            topo.handle, int(self.payload_f32)))
        self.payload = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        lib.call("pvdb_render_build_payload", topo.handle, lib.ptr(slot_idx), lib.ptr(den),
                 lib.ptr(col), int(N), int(self.payload_f32), lib.ptr(self.payload), lib.stream())
        torch.cuda.current_stream().synchronize()
        self.topo = topo
        self.flags[0] = True

    def setScene(self, reso, K, xyz_min, xyz_max) -> None:
        d = self.desc
        for i, v in enumerate(np.asarray(reso).reshape(-1)[:3]):
            d.reso[i] = int(v)
        for i, v in enumerate(np.asarray(K, np.float32).reshape(-1)[:9]):
            d.K[i] = float(v)
        for i in range(3):
            d.xyz_min[i] = float(np.asarray(xyz_min, np.float32).reshape(-1)[i])
            d.xyz_max[i] = float(np.asarray(xyz_max, np.float32).reshape(-1)[i])
        self.flags[2] = True

    def setKwargs(self, near, far, stepdist, act_shift, interval, fast_color_thres, bg,
                  inverse_y, H, W) -> None:
        d = self.desc
        d.near_ = float(near)  # far is forced to 1e9 (plenvdb.h:1008)
        d.stepdist, d.act_shift, d.interval = float(stepdist), float(act_shift), float(interval)
        d.fast_color_thres, d.bg = float(fast_color_thres), float(bg)
        d.inverse_y = int(bool(inverse_y))
        d.H, d.W = int(H), int(W)
        self.image = torch.zeros(int(H) * int(W) * 3, dtype=torch.float32, device="cuda")
        # records per frame: start at 16 per pixel; a frame that needs more is
        # rendered again with room for all of them (render_an_image)
        self._reserve(16 * int(H) * int(W))
        self.flags[3] = True

    def input_a_c2w(self, c2w) -> None:
        v = np.asarray(c2w.detach().cpu() if isinstance(c2w, torch.Tensor) else c2w,
                       np.float32).reshape(-1)
        if v.size != 16:
            raise ValueError("c2w must hold 16 values")
        for i in range(16):
            self.desc.c2w[i] = float(v[i])
        self.flags[4] = True

    def set_t_stop(self, t: float) -> None:
        """Early-stop transmittance (the reference hard-codes 1e-3)."""
        self.t_stop = float(t)

    def _reserve(self, capacity: int) -> None:
        """Workspace for frames of up to `capacity` records (kept samples)."""
        capacity = min(int(capacity), 2 ** 31 - 2)
        d = self.desc
        if self.workspace is not None and capacity <= self.capacity and \
                self._ws_hw == (d.H, d.W):
            return
        capacity = max(capacity, self.capacity if self._ws_hw == (d.H, d.W) else 0)
        nbytes = int(_i64_fn("pvdb_render_workspace_bytes", [C.c_int32, C.c_int32, C.c_int64])(
            d.H, d.W, capacity))
        if nbytes < 0:
            raise ValueError("bad image size or record capacity")
        self.workspace = None
        self.workspace = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        self.capacity = capacity
        self._ws_hw = (d.H, d.W)

    def frame_total(self) -> int:
        """Records (kept samples) the last frame needed (synchronises)."""
        v = C.c_uint64()
        lib.call("pvdb_render_frame_info", C.byref(self.desc), C.byref(v), None, None, None)
        return int(v.value)

    def pixel_counts(self) -> tuple[np.ndarray, np.ndarray]:
        """Per pixel: the first march's kept count and the samples the second
        march wrote (host copies of the last frame's)."""
        pc, pw = C.c_void_p(), C.c_void_p()
        lib.call("pvdb_render_frame_info", C.byref(self.desc), None, C.byref(pc), C.byref(pw),
                 None)
        n = self.desc.H * self.desc.W
        torch.cuda.current_stream().synchronize()
        from .fused import _read_i32

        return _read_i32(pc.value, n), _read_i32(pw.value, n)

    # -- rendering (plenvdb.h:1029-1049) ----------------------------------------------
    def _launch(self) -> None:
        d = self.desc
        d.t_stop = self.t_stop
        d.payload_f32 = int(self.payload_f32)
        d.payload = self.payload.data_ptr()
        d.weights = self.weights.data_ptr()
        d.out = self.image.data_ptr()
        d.counters = self.counters.data_ptr()
        d.workspace = self.workspace.data_ptr()
        d.workspace_bytes = self.workspace.numel()
        d.capacity = self.capacity
        lib.call("pvdb_render_image", self.topo.handle, C.byref(d), lib.stream())

    def render_an_image(self) -> None:
        if not all(self.flags):
            return
        t0 = time.perf_counter()
        self._launch()
        need = self.frame_total()  # synchronises
        if need > self.capacity:  # the frame kept more samples than there was room for
            self._reserve(need + need // 4)
            self._launch()
            torch.cuda.current_stream().synchronize()
        self.timer += time.perf_counter() - t0

    def render_async(self) -> torch.Tensor:
        """Launches a render on the current stream without waiting; returns the
        device image [H*W*3] (overwritten by the next render).  A frame with more
        kept samples than the record capacity renders only those that fit:
        render the poses once with render_an_image (which grows the capacity)
        or call reserve() first."""
        if not all(self.flags):
            raise RuntimeError("MGRenderer: load_data, load_params, setScene, setKwargs and "
                               "input_a_c2w must all be called before rendering")
        self._launch()
        return self.image

    def output_an_image(self):
        if not all(self.flags):
            return None
        return self.image.cpu().numpy()

    @staticmethod
    def stage_timing(enable: bool) -> np.ndarray:
        """Starts (True) or stops (False) the device timing of the march, MLP
        and composite kernels; returns the totals (ms) collected until now."""
        out = (C.c_double * 3)()
        lib.call("pvdb_render_timing", 1 if enable else 0, out)
        return np.array(out[:])

    def reserve(self, capacity: int) -> None:
        """Room for frames of up to `capacity` kept samples."""
        self._reserve(capacity)

    def sample_counts(self) -> tuple[int, int]:
        """(first-march density gathers, written samples) of the last render."""
        c = self.counters.cpu()
        return int(c[0]), int(c[1])

    def resetTimer(self) -> None:
        self.timer = 0.0

    def getTimer(self) -> float:
        return self.timer
