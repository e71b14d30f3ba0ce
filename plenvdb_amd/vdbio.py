"""Grid files: the container that `save_to` / `load_from` and the merged index
grid use in place of OpenVDB `.vdb` files.

The reference writes and reads OpenVDB files through libopenvdb
(plenvdb/lib/vdb/plenvdb.h:126-148, 211-276, 720-736; vdb_compression.py:52-53).
This package keeps the same calls and file names but stores a grid as a NumPy
`.npz` archive (written to exactly the path given, loaded with
allow_pickle=False):

    format        "plenvdb-amd-grid/1"
    kind          "float" (payload [slots, ndim] float32) or "index" (int32 [slots])
    reso          int32 [3]
    leaf_origins  int32 [L, 3]   leaf origins in NanoVDB depth-first order
    leaf_masks    uint64 [L, 8]  512-bit value masks (active voxels)
    values        payload in slot order (slot = leaf*512 + (x&7)<<6 | (y&7)<<3 | z&7)

An OpenVDB file cannot be read or written here (the OpenVDB library is not
available), so files are not interchangeable with the reference's.
"""
from __future__ import annotations

import numpy as np
import torch

from .topology import Topology

FORMAT = "plenvdb-amd-grid/1"


def _write(path: str, **arrays) -> None:
    with open(path, "wb") as f:  # a file object keeps np.savez from renaming the path
        np.savez(f, format=np.array(FORMAT), **arrays)


def _read(path: str) -> dict:
    with np.load(path, allow_pickle=False) as z:
        out = {k: z[k] for k in z.files}
    if str(out.get("format", "")) != FORMAT:
        raise ValueError(f"{path}: not a {FORMAT} file")
    return out


def save_grid(path: str, topo: Topology, values: torch.Tensor, reso) -> None:
    """Writes a float payload (density [slots] or colour [slots, C])."""
    o, m = topo.leaves()
    v = values.detach().float().cpu().numpy().reshape(topo.nslots, -1)
    _write(path, kind=np.array("float"), reso=np.asarray(reso, np.int32), leaf_origins=o,
           leaf_masks=m, values=v)


def active_voxel_dim(origins: np.ndarray, masks: np.ndarray) -> list[int]:
    """evalActiveVoxelDim: extent (max - min + 1) of the active voxels' bounding box."""
    if len(origins) == 0:
        return [0, 0, 0]
    bits = np.unpackbits(masks.view(np.uint8).reshape(-1, 64), axis=1, bitorder="little")
    leaf, off = np.nonzero(bits.astype(bool))
    if len(leaf) == 0:
        return [0, 0, 0]
    xyz = origins[leaf].astype(np.int64) + np.stack([off >> 6, (off >> 3) & 7, off & 7], 1)
    return [int(d) for d in (xyz.max(0) - xyz.min(0) + 1)]


def prune_leaves(origins: np.ndarray, masks: np.ndarray, values: np.ndarray):
    """The leaves pruneGrid keeps (plenvdb.h:135, 227).  OpenVDB's prune
    replaces a node whose values are all equal and whose active states are all
    equal by a tile; a leaf with no active voxel whose values are all the
    background 0 becomes background, which is no leaf at all, so it is dropped
    here.  A uniform leaf with active voxels (or with another value) would
    become a tile that this layout does not hold; it stays a leaf with the same
    values and activity.  Returns (origins, masks, values) of the kept leaves
    in their original (NanoVDB) order."""
    nl = len(origins)
    if nl == 0:
        return origins, masks, values
    per_leaf = values.reshape(nl, 512, -1)
    empty = ~masks.reshape(nl, 8).any(1) & ~per_leaf.reshape(nl, -1).any(1)
    keep = ~empty
    return origins[keep], masks[keep], per_leaf[keep].reshape(-1, values.shape[1])


def load_grid(path: str, ndim: int, device="cuda", prune: bool = True):
    """Returns (topology, values [slots] or [slots, ndim] on the device, reso) where
    reso is the active voxels' extent (plenvdb.h:126-142 sets reso from
    evalActiveVoxelDim).  With prune (as the reference's load_from) the leaves
    pruneGrid turns into background are dropped (prune_leaves)."""
    z = _read(path)
    if str(z["kind"]) != "float":
        raise ValueError(f"{path}: expected a float grid")
    vals = z["values"]
    if vals.shape[1] != ndim:
        raise ValueError(f"{path}: grid has {vals.shape[1]} channels, expected {ndim}")
    if prune:
        z["leaf_origins"], z["leaf_masks"], vals = prune_leaves(z["leaf_origins"], z["leaf_masks"],
                                                               vals)
    topo = Topology.from_leaves(z["leaf_origins"], z["leaf_masks"], z["reso"])
    v = torch.from_numpy(np.ascontiguousarray(vals, np.float32)).to(device)
    if ndim == 1:
        v = v.reshape(-1)
    return topo, v.contiguous(), active_voxel_dim(z["leaf_origins"], z["leaf_masks"])


def load_onto(path: str, topo: Topology, ndim: int, device="cuda") -> torch.Tensor:
    """Values of a saved grid placed on another topology by leaf origin: slots
    of leaves the file does not hold read 0.  Adam moments saved with one leaf
    set thus land on the right voxels of a parameter whose leaf set was pruned
    on load (the reference loads both with their own leaf counts and pairs
    them by index, which no longer matches then)."""
    z = _read(path)
    if str(z["kind"]) != "float":
        raise ValueError(f"{path}: expected a float grid")
    vals = z["values"]
    if vals.shape[1] != ndim:
        raise ValueError(f"{path}: grid has {vals.shape[1]} channels, expected {ndim}")
    src = {tuple(o): i for i, o in enumerate(np.asarray(z["leaf_origins"]).tolist())}
    dst, _ = topo.leaves()
    out = np.zeros((topo.nslots, ndim), np.float32)
    per_leaf = vals.reshape(-1, 512, ndim)
    for j, o in enumerate(np.asarray(dst).tolist()):
        i = src.get(tuple(o))
        if i is not None:
            out[512 * j:512 * (j + 1)] = per_leaf[i]
    v = torch.from_numpy(out).to(device)
    return (v.reshape(-1) if ndim == 1 else v).contiguous()


def index_grid_arrays(idx_grid: np.ndarray):
    """Topology of a dense merged index grid ([X,Y,Z] int, 0 = inactive; its
    leaves are the 8^3 blocks holding a non-zero index) and the index of every
    slot (host int32 [slots], 0 for inactive slots)."""
    idx = np.ascontiguousarray(np.asarray(idx_grid))
    if idx.ndim != 3:
        raise ValueError("index grid must be 3-D")
    on = idx != 0
    topo = Topology.from_mask(on)
    slots, _ = topo.slots(np.argwhere(on))
    per_slot = np.zeros(topo.nslots, np.int32)
    per_slot[slots] = idx[on].astype(np.int32)
    return topo, per_slot


def save_index_grid(path: str, idx_grid: np.ndarray) -> None:
    """Writes a merged index grid given densely ([X,Y,Z] int, 0 = inactive)."""
    topo, per_slot = index_grid_arrays(idx_grid)
    o, m = topo.leaves()
    idx = np.asarray(idx_grid)
    _write(path, kind=np.array("index"), reso=np.asarray(idx.shape, np.int32), leaf_origins=o,
           leaf_masks=m, values=per_slot)


def load_index_grid(path: str, device="cuda"):
    """Returns (topology, slot_idx int32 [slots] on the device, reso)."""
    z = _read(path)
    if str(z["kind"]) != "index":
        raise ValueError(f"{path}: expected an index grid")
    topo = Topology.from_leaves(z["leaf_origins"], z["leaf_masks"], z["reso"])
    v = torch.from_numpy(np.ascontiguousarray(z["values"], np.int32)).to(device)
    return topo, v, [int(r) for r in z["reso"]]
