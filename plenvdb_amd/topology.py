"""Python handle of a flattened VDB topology (leaf table + value masks in HBM).

A topology is shared by every payload defined on it (data, gradient, Adam
moments, per-voxel lr), which keeps their slot numbering identical.  In the
reference each of those is a separate NanoVDB grid built from the same dense
topology (plenvdb/lib/vdb/plenvdb.h:117-125, 434-439, 706-708).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import lib


class Topology:
    """Owner of a pvdb_topology handle on the current HIP device."""

    def __init__(self, handle: C.c_void_p, reso):
        self._h = handle
        self.reso = [int(r) for r in reso]
        nleaf, nact = C.c_int64(), C.c_int64()
        lo, dim = lib.i3([0, 0, 0]), lib.i3([0, 0, 0])
        two = (C.c_int32 * 1)()
        lib.call("pvdb_topo_info", self._h, C.byref(nleaf), C.byref(nact), lo, dim, two)
        self.two_level = bool(two[0])
        self.nleaf = int(nleaf.value)
        self.nactive = int(nact.value)
        self.leaf_lo = [lo[i] for i in range(3)]
        self.leaf_dim = [dim[i] for i in range(3)]

    # -- constructors ---------------------------------------------------------
    @classmethod
    def from_mask(cls, mask) -> "Topology":
        """Leaves for every 8^3 block holding an active voxel of a dense [X,Y,Z] mask."""
        if isinstance(mask, torch.Tensor):
            mask = mask.detach().cpu().numpy()
        m = np.ascontiguousarray(np.asarray(mask).astype(np.uint8))
        if m.ndim != 3:
            raise ValueError("mask must be 3-D")
        lib.load()
        h = C.c_void_p()
        lib.call("pvdb_topo_create_from_mask", lib.hptr(m), m.shape[0], m.shape[1], m.shape[2],
                 C.byref(h))
        return cls(h, m.shape)

    @classmethod
    def dense(cls, reso) -> "Topology":
        """denseFill topology of [0,reso-1]^3 (plenvdb/lib/vdb/plenvdb.h:117-125)."""
        return cls.from_mask(np.ones([int(r) for r in reso], dtype=np.uint8))

    @classmethod
    def from_leaves(cls, origins, masks, reso, layout: int = 0) -> "Topology":
        """layout 0 automatic, 1 one dense leaf table, 2 two-level tables."""
        o = np.ascontiguousarray(np.asarray(origins, dtype=np.int32).reshape(-1, 3))
        m = np.ascontiguousarray(np.asarray(masks, dtype=np.uint64).reshape(-1, 8))
        if len(o) != len(m):
            raise ValueError("origins and masks disagree in length")
        lib.load()
        h = C.c_void_p()
        lib.call("pvdb_topo_create_from_leaves", lib.hptr(o), lib.hptr(m), len(o), int(layout),
                 C.byref(h))
        return cls(h, reso)

    # -- queries ----------------------------------------------------------------
    @property
    def handle(self) -> C.c_void_p:
        return self._h

    @property
    def nslots(self) -> int:
        return self.nleaf * 512

    def leaves(self):
        """Leaf origins [L,3] int32 and value masks [L,8] uint64 in NanoVDB order."""
        o = np.zeros((self.nleaf, 3), dtype=np.int32)
        m = np.zeros((self.nleaf, 8), dtype=np.uint64)
        lib.call("pvdb_topo_leaves", self._h, lib.hptr(o), lib.hptr(m))
        return o, m

    def empty_run(self) -> np.ndarray | None:
        """The marchers' empty-run map over the leaf grid box (uint8
        [leaf_dim], 0 where a leaf exists, else the radius r of the empty cube
        of cells around the cell, capped) or None for the two-level layout."""
        fn = lib.load().pvdb_topo_empty_run
        fn.restype = C.c_int64
        fn.argtypes = [C.c_void_p, C.c_void_p]
        n = int(fn(self._h, None))
        if n <= 0:
            return None
        out = np.zeros(self.leaf_dim, dtype=np.uint8)
        assert out.size == n
        fn(self._h, lib.hptr(out))
        return out

    def slots(self, ijk):
        """Slot index (-1 when the leaf does not exist) and active flag of voxels."""
        q = np.ascontiguousarray(np.asarray(ijk, dtype=np.int32).reshape(-1, 3))
        s = np.zeros(len(q), dtype=np.int64)
        a = np.zeros(len(q), dtype=np.uint8)
        lib.call("pvdb_topo_slots", self._h, lib.hptr(q), len(q), lib.hptr(s), lib.hptr(a))
        return s, a.astype(bool)

    def active_slot_mask(self) -> np.ndarray:
        """Boolean [nslots] array, True for active slots."""
        _, m = self.leaves()
        bits = np.unpackbits(m.view(np.uint8).reshape(-1, 64), axis=1, bitorder="little")
        return bits.reshape(-1).astype(bool)

    def dense_active(self, reso=None) -> np.ndarray:
        """Dense bool [X, Y, Z] grid of the active voxels inside [0, reso)."""
        reso = [int(r) for r in (reso if reso is not None else self.reso)]
        o, _ = self.leaves()
        act = self.active_slot_mask().reshape(-1, 512)
        v = np.arange(512)
        local = np.stack([v >> 6, (v >> 3) & 7, v & 7], 1)
        out = np.zeros(reso, dtype=bool)
        li, vi = np.nonzero(act)
        c = o[li] + local[vi]
        ok = np.all((c >= 0) & (c < np.asarray(reso)), axis=1)
        c = c[ok]
        out[c[:, 0], c[:, 1], c[:, 2]] = True
        return out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib.load().pvdb_topo_destroy(h)
            except Exception:
                pass
            self._h = None
