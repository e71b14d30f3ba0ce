"""Merged-VDB export: the render-only data of a trained model
(plenvdb/vdb_compression.py:19-59).

Every active voxel of the occupancy mask gets index 1 + its C-order rank
(x slowest); index 0 is the value of every other voxel.  `den` [N+1] and
`col` [N+1, C] hold the density and colour of the indexed voxels with a zero
row 0, stored as float16 (vdb_compression.py:56-59) and read back as float32
by the renderer's caller (run.py:86-95).  The data file keeps the reference's
name and layout (`mergeddata.npz` with arrays `den` and `col`); the index grid
`mergedidxs.vdb` uses this package's grid container (vdbio).
"""
from __future__ import annotations

import os

import numpy as np


def export_merged(mask, den_dense, col_dense):
    """mask [X,Y,Z] bool, den_dense [X,Y,Z], col_dense [X,Y,Z,C] ->
    (idx [X,Y,Z] int32, den float16 [N+1], col float16 [N+1, C])."""
    m = np.asarray(mask).astype(bool)
    n = int(m.sum())
    idx = np.zeros(m.shape, np.int32)
    idx[m] = np.arange(1, n + 1, dtype=np.int32)  # boolean indexing walks C order
    den = np.concatenate([[0.0], np.asarray(den_dense)[m].reshape(-1)]).astype(np.float16)
    cd = np.asarray(col_dense)
    c = cd.shape[-1]
    col = np.concatenate([np.zeros((1, c)), cd[m].reshape(-1, c)]).astype(np.float16)
    return idx, den, col


def export_from_grids(mask, density_vdb, color_vdb):
    """export_merged from a trained DensityVDB / ColorVDB (their dense grids, as
    copyToArray gives them in vdb_compression.py:35-47)."""
    reso = list(np.asarray(mask).shape)
    den = np.asarray(density_vdb.get_dense_grid()).reshape(reso)
    col = np.asarray(color_vdb.get_dense_grid()).reshape(reso + [color_vdb.getndim()])
    return export_merged(mask, den, col)


def save_merged(ckptdir: str, idx, den, col) -> None:
    """Writes mergedidxs.vdb (index grid) and mergeddata.npz (den, col)."""
    from . import vdbio

    os.makedirs(ckptdir, exist_ok=True)
    vdbio.save_index_grid(os.path.join(ckptdir, "mergedidxs.vdb"), idx)
    np.savez_compressed(os.path.join(ckptdir, "mergeddata"), den=den, col=col)


def load_merged_data(ckptdir: str):
    """(den float32 [N+1], col float32 [(N+1)*C], N) as run.py:86-95 reads them."""
    with np.load(os.path.join(ckptdir, "mergeddata.npz"), allow_pickle=False) as z:
        den = z["den"].astype(np.float32)
        col = z["col"].astype(np.float32)
    return den.reshape(-1), col.reshape(-1), den.shape[0]
