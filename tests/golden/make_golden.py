"""Writes the NanoVDB golden fixtures of tests/golden/ (run in the build
container, where the reference tree exists).

It feeds active-voxel sets to oracle/_ref/nanovdb_golden -- a driver compiled
from the reference's own header-only NanoVDB (openvdb/nanovdb/nanovdb) by
oracle/Makefile -- and stores what NanoVDB reports: leaf order, value masks,
getValue / isActive / leaf-cached / slot of integer queries, and the density
trilinear rule of densityvdb.cu:101-125 evaluated through NanoVDB's accessor
(plus NanoVDB's own TrilinearSampler).

Fixtures
  nanovdb_c1.npz      topology of the 128^3 benchmark scene (337 leaves)
  nanovdb_sparse.npz  250 random partial leaves spread over [-6000, 6000]^3
                      (several upper nodes, negative coordinates)
Usage: python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
GEN = os.path.join(ROOT, "oracle", "_ref", "nanovdb_golden")
OUT = os.path.dirname(os.path.abspath(__file__))


def run_generator(vox: np.ndarray, vals: np.ndarray, queries: np.ndarray, points: np.ndarray):
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("<i", len(vox)))
            rec = np.zeros(len(vox), dtype=[("c", "<i4", 3), ("v", "<f4")])
            rec["c"] = vox
            rec["v"] = vals
            f.write(rec.tobytes())
            f.write(struct.pack("<i", len(queries)))
            f.write(np.ascontiguousarray(queries, dtype="<i4").tobytes())
            f.write(struct.pack("<i", len(points)))
            f.write(np.ascontiguousarray(points, dtype="<f4").tobytes())
        subprocess.run([GEN, fin, fout], check=True)
        raw = open(fout, "rb").read()
    nleaf = struct.unpack_from("<i", raw, 0)[0]
    off = 4
    lrec = np.frombuffer(raw, dtype=[("o", "<i4", 3), ("m", "<u8", 8)], count=nleaf, offset=off)
    off += lrec.nbytes
    qrec = np.frombuffer(raw, dtype=[("v", "<f4"), ("a", "<i4"), ("c", "<i4"), ("s", "<i8")],
                         count=len(queries), offset=off)
    off += qrec.nbytes
    prec = np.frombuffer(raw, dtype=[("r", "<f4"), ("s", "<f4")], count=len(points), offset=off)
    return {
        "leaf_origins": lrec["o"].copy(),
        "leaf_masks": lrec["m"].copy(),
        "q_value": qrec["v"].copy(),
        "q_active": qrec["a"].astype(np.uint8),
        "q_cached": qrec["c"].astype(np.uint8),
        "q_slot": qrec["s"].copy(),
        "p_tri_ref": prec["r"].copy(),
        "p_tri_sampler": prec["s"].copy(),
    }


def c1_case():
    from plenvdb_amd import scenes

    cfg = scenes.CONFIGS["c1"]
    mask = scenes.occupancy_mask(cfg)
    vox = np.argwhere(mask).astype(np.int32)
    rng = np.random.default_rng(11)
    vals = rng.normal(0, 2, size=len(vox)).astype(np.float32)
    R = cfg.reso
    q = np.concatenate([
        rng.integers(-4, R + 4, size=(1500, 3)),
        vox[rng.integers(0, len(vox), 400)],
        np.array([[0, 0, 0], [R - 1, R - 1, R - 1], [R, R, R], [-1, -1, -1], [-8, 0, 0]]),
    ]).astype(np.int32)
    # points inside the grid, near active voxels and across leaf faces
    pts = np.concatenate([
        rng.uniform(0, R - 1, size=(1500, 3)),
        vox[rng.integers(0, len(vox), 500)] + rng.uniform(-1, 1, size=(500, 3)),
        np.array([[7.5, 7.5, 7.5], [8.0, 8.0, 8.0], [R - 1, R - 1, R - 1], [0, 0, 0]]),
    ]).astype(np.float32)
    return vox, vals, q, pts


def sparse_case():
    rng = np.random.default_rng(12)
    L = 250
    org = (rng.integers(-750, 750, size=(L, 3)) * 8).astype(np.int32)
    org = np.unique(org, axis=0)
    # a cluster of neighbouring leaves so trilinear points straddle leaves
    base = np.array([[4088, -16, 96]], dtype=np.int32)
    cl = base + 8 * np.stack(np.meshgrid(np.arange(3), np.arange(3), np.arange(3), indexing="ij"),
                             -1).reshape(-1, 3).astype(np.int32)
    org = np.unique(np.concatenate([org, cl]), axis=0)
    vox = []
    for o in org:
        keep = rng.random(512) < 0.6
        keep[rng.integers(0, 512)] = True
        off = np.nonzero(keep)[0]
        vox.append(o[None] + np.stack([off >> 6, (off >> 3) & 7, off & 7], 1))
    vox = np.concatenate(vox).astype(np.int32)
    vals = rng.normal(0, 1, size=len(vox)).astype(np.float32)
    q = np.concatenate([
        vox[rng.integers(0, len(vox), 1200)],
        org[rng.integers(0, len(org), 600)] + rng.integers(0, 8, size=(600, 3)),
        rng.integers(-6000, 6000, size=(300, 3)),
    ]).astype(np.int32)
    lo = base[0].astype(np.float32)
    pts = np.concatenate([
        lo + rng.uniform(0, 23, size=(1500, 3)),
        vox[rng.integers(0, len(vox), 500)] + rng.uniform(-1, 1, size=(500, 3)),
    ]).astype(np.float32)
    return vox, vals, q, pts


def main():
    if not os.path.exists(GEN):
        raise SystemExit(f"{GEN} missing: run `make -C oracle ref` first")
    for name, case in (("nanovdb_c1", c1_case), ("nanovdb_sparse", sparse_case)):
        vox, vals, q, pts = case()
        # lexicographic voxel order, so the voxel list is recoverable from the
        # leaf masks and only the values need storing
        order = np.lexsort((vox[:, 2], vox[:, 1], vox[:, 0]))
        vox, vals = vox[order], vals[order]
        res = run_generator(vox, vals, q, pts)
        np.savez_compressed(os.path.join(OUT, name + ".npz"), vals=vals, queries=q, points=pts,
                            **res)
        print(name, "leaves", len(res["leaf_origins"]), "voxels", len(vox))


if __name__ == "__main__":
    main()
