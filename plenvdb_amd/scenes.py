"""Synthetic workloads of the benchmark configurations.

Seeds (numpy default_rng): 0 spheres, 1 grid values, 2 rays, 3 targets,
4 MLP weights.  A leaf (8^3 voxels) is active when its world-space AABB meets
a sphere (or the shell); every voxel of an active leaf inside [0,R)^3 is
active.  Index space follows VDBGrid.wld2idx (plenvdb/lib/grid.py:77-78):
voxel i sits at xyz_min + i*(xyz_max-xyz_min)/(R-1).

Configurations
  c1  128^3, 16 spheres r~U(0.1,0.4) (8.2% of voxels active, 337 leaves),
      4096 rays, no MLP (12-ch sigmoid composite), fast_color_thres 0.
  c2  160^3, 64 spheres r~U(0.04,0.16) (7.5% active, 597 leaves), 8192 rays,
      12-ch features + MLP 39-128-128-3, fast_color_thres 1e-4.  (The C1
      radius range would give 37% at 64 spheres; the radii are scaled to the
      stated ~8% occupancy.)
  c3  160^3 as c2, 65536 rays from the upper hemisphere, far ~U(2.5, 6).
  c4  320^3, 256 spheres r~U(0.015,0.035) plus a one-leaf shell of radius
      0.6 (3.1% active, 1.0M voxels, 1959 leaves), 131072 rays, BF16 colour
      features.
  c5  c4's topology as a merged index grid; 800x800 pinhole camera, focal
      1111, 100 poses on a radius-4 orbit at -30 degrees elevation.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field, replace

import numpy as np

XYZ_MIN = np.array([-1.5, -1.5, -1.5], dtype=np.float32)
XYZ_MAX = np.array([1.5, 1.5, 1.5], dtype=np.float32)


@dataclass
class SceneConfig:
    name: str
    reso: int
    n_spheres: int
    r_lo: float
    r_hi: float
    n_rays: int
    use_mlp: bool
    fast_color_thres: float
    shell: tuple | None = None  # (radius, half thickness)
    hemisphere: bool = False
    far_lo: float | None = None
    far_hi: float | None = None
    color_bf16: bool = False
    channels: int = 12
    stepsize: float = 0.5
    near: float = 2.0
    far: float = 6.0
    alpha_init: float = 1e-2
    bg: float = 1.0
    extra: dict = field(default_factory=dict)


CONFIGS = {
    "c1": SceneConfig("c1", 128, 16, 0.1, 0.4, 4096, False, 0.0),
    "c2": SceneConfig("c2", 160, 64, 0.04, 0.16, 8192, True, 1e-4),
    "c3": SceneConfig("c3", 160, 64, 0.04, 0.16, 65536, True, 1e-4, hemisphere=True, far_lo=2.5,
                      far_hi=6.0),
    "c4": SceneConfig("c4", 320, 256, 0.015, 0.035, 131072, True, 1e-4, shell=(0.6, 0.0),
                      color_bf16=True),
}


def leaf_aabbs(reso: int):
    """World AABB (lo, hi) [L,3] of every leaf of the dense leaf grid."""
    nl = (reso + 7) // 8
    h = (XYZ_MAX - XYZ_MIN).astype(np.float64) / (reso - 1)
    g = np.stack(np.meshgrid(np.arange(nl), np.arange(nl), np.arange(nl), indexing="ij"), -1)
    g = g.reshape(-1, 3).astype(np.float64)
    lo = XYZ_MIN + g * 8 * h
    hi = XYZ_MIN + np.minimum(g * 8 + 7, reso - 1) * h
    return g.astype(np.int64), lo, hi


def occupancy_mask(cfg: SceneConfig) -> np.ndarray:
    """Dense [R,R,R] bool mask of active voxels (leaf-granular)."""
    rng = np.random.default_rng(0)
    centers = rng.uniform(XYZ_MIN, XYZ_MAX, size=(cfg.n_spheres, 3))
    radii = rng.uniform(cfg.r_lo, cfg.r_hi, size=cfg.n_spheres)
    g, lo, hi = leaf_aabbs(cfg.reso)
    on = np.zeros(len(g), dtype=bool)
    for c, r in zip(centers, radii):
        d = np.maximum(np.maximum(lo - c, 0), c - hi)
        on |= (d * d).sum(1) <= r * r
    if cfg.shell is not None:
        rad, half = cfg.shell
        dmin = np.sqrt((np.maximum(np.maximum(lo, -hi), 0) ** 2).sum(1))  # to origin
        far_c = np.maximum(np.abs(lo), np.abs(hi))
        dmax = np.sqrt((far_c ** 2).sum(1))
        on |= (dmin <= rad + half) & (dmax >= rad - half)
    nl = (cfg.reso + 7) // 8
    leaf_on = on.reshape(nl, nl, nl)
    m = np.repeat(np.repeat(np.repeat(leaf_on, 8, 0), 8, 1), 8, 2)
    return m[: cfg.reso, : cfg.reso, : cfg.reso].copy()


def voxel_size(cfg: SceneConfig) -> float:
    """((prod(max-min))/num_voxels)^(1/3) with num_voxels = R^3 (dvgo.py:132-137)."""
    return float(np.float32(3.0) / np.float32(cfg.reso))


def grid_values(cfg: SceneConfig, nslots: int):
    """Density ~N(0,2) and colour ~N(0,0.1) per slot (slot order)."""
    rng = np.random.default_rng(1)
    den = rng.normal(0.0, 2.0, size=nslots).astype(np.float32)
    col = rng.normal(0.0, 0.1, size=(nslots, cfg.channels)).astype(np.float32)
    return den, col


def rays(cfg: SceneConfig, n: int | None = None):
    """Origins on a radius-4 sphere (upper hemisphere for c3), aimed at the
    origin with U(-0.2,0.2)^3 jitter, unit directions."""
    n = cfg.n_rays if n is None else n
    rng = np.random.default_rng(2)
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    if cfg.hemisphere:
        v[:, 2] = np.abs(v[:, 2])
    o = 4.0 * v
    tgt = rng.uniform(-0.2, 0.2, size=(n, 3))
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o = o.astype(np.float32)
    d = d.astype(np.float32)
    viewdirs = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    far = None
    if cfg.far_lo is not None:
        far = rng.uniform(cfg.far_lo, cfg.far_hi, size=n).astype(np.float32)
    return o, d, viewdirs, far


def targets(cfg: SceneConfig, n: int | None = None):
    n = cfg.n_rays if n is None else n
    rng = np.random.default_rng(3)
    nch = 3 if cfg.use_mlp else cfg.channels
    return rng.uniform(0.0, 1.0, size=(n, nch)).astype(np.float32)


def mlp_params(cfg: SceneConfig, dcol: int = 12, width: int = 128, n_freq: int = 4):
    """torch.nn.Linear default init bounds (U(+-1/sqrt(fan_in))); the last bias is
    0 (dvgo.py:97-105).  Weights are stored [in, out] (Linear.weight.T)."""
    rng = np.random.default_rng(4)
    din = dcol + 3 + 3 * n_freq * 2
    dims = [(din, width), (width, width), (width, 3)]
    out = {}
    for i, (fi, fo) in enumerate(dims):
        b = 1.0 / math.sqrt(fi)
        out[f"w{i}"] = rng.uniform(-b, b, size=(fi, fo)).astype(np.float32)
        out[f"b{i}"] = rng.uniform(-b, b, size=fo).astype(np.float32)
    out["b2"][:] = 0
    return out


def act_shift(alpha_init: float) -> float:
    """log(1/(1-alpha_init) - 1) (dvgo.py:48)."""
    return float(np.float32(np.log(1 / (1 - alpha_init) - 1)))


def orbit_poses(n: int = 100, radius: float = 4.0, elevation_deg: float = -30.0):
    """Camera-to-world matrices on an orbit (load_blender.py:74 pose_spherical convention)."""
    def trans_t(t):
        return np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, t], [0, 0, 0, 1]], dtype=np.float64)

    def rot_phi(phi):
        return np.array([[1, 0, 0, 0], [0, np.cos(phi), -np.sin(phi), 0],
                         [0, np.sin(phi), np.cos(phi), 0], [0, 0, 0, 1]], dtype=np.float64)

    def rot_theta(th):
        return np.array([[np.cos(th), 0, -np.sin(th), 0], [0, 1, 0, 0],
                         [np.sin(th), 0, np.cos(th), 0], [0, 0, 0, 1]], dtype=np.float64)

    poses = []
    for th in np.linspace(-180, 180, n + 1)[:-1]:
        c2w = trans_t(radius)
        c2w = rot_phi(elevation_deg / 180.0 * np.pi) @ c2w
        c2w = rot_theta(th / 180.0 * np.pi) @ c2w
        c2w = np.array([[-1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]]) @ c2w
        poses.append(c2w.astype(np.float32))
    return np.stack(poses)


def intrinsics(H: int = 800, W: int = 800, focal: float = 1111.0):
    return np.array([[focal, 0, 0.5 * W], [0, focal, 0.5 * H], [0, 0, 1]], dtype=np.float32)


def make_scene(name: str, n_rays: int | None = None, channels: int | None = None) -> dict:
    """Everything a training step needs, as host numpy arrays.  Grid values are
    given per slot, so the caller needs the topology's slot order: the leaves
    come from `occupancy_mask`, sorted into NanoVDB order by the topology
    builder.  `channels` replaces the colour channel count of a configuration
    without the MLP (3 gives the coarse stage's k0 grid, dvgo.py:72-75)."""
    cfg = CONFIGS[name]
    if channels is not None:
        if cfg.use_mlp:
            raise ValueError("the MLP configurations take 12 colour features")
        cfg = replace(cfg, channels=int(channels))
    mask = occupancy_mask(cfg)
    o, d, vd, far = rays(cfg, n_rays)
    sc = {
        "cfg": cfg,
        "mask": mask,
        "reso": [cfg.reso] * 3,
        "xyz_min": XYZ_MIN.copy(),
        "xyz_max": XYZ_MAX.copy(),
        "rays_o": o,
        "rays_d": d,
        "viewdirs": vd,
        "far_per_ray": far,
        "target": targets(cfg, len(o)),
        "near": cfg.near,
        "stepsize": cfg.stepsize,
        "voxel_size": voxel_size(cfg),
        "voxel_size_ratio": 1.0,
        "act_shift": act_shift(cfg.alpha_init),
        "fast_color_thres": cfg.fast_color_thres,
        "bg": cfg.bg,
    }
    if cfg.use_mlp:
        sc["mlp"] = mlp_params(cfg, cfg.channels)
    return sc


def merged_scene(name: str = "c4", H: int = 800, W: int = 800, focal: float = 1111.0,
                 n_poses: int = 100, t_stop: float = 1e-4) -> dict:
    """Render-only workload (c5 when name is c4): the configuration's topology as a
    merged index grid (index 1 + C-order rank of each active voxel), fp16 data
    (density ~N(0,2), colour ~N(0,0.1), zero row 0), the configuration's MLP,
    and an orbit of camera poses."""
    cfg = CONFIGS[name]
    mask = occupancy_mask(cfg)
    n = int(mask.sum())
    idx = np.zeros(mask.shape, np.int32)
    idx[mask] = np.arange(1, n + 1, dtype=np.int32)
    rng = np.random.default_rng(1)
    den = np.concatenate([[0.0], rng.normal(0.0, 2.0, n)]).astype(np.float16)
    col = np.concatenate([np.zeros((1, 12)), rng.normal(0.0, 0.1, (n, 12))]).astype(np.float16)
    vs = voxel_size(cfg)
    return {
        "cfg": cfg, "mask": mask, "idx": idx, "den": den, "col": col,
        "mlp": mlp_params(cfg, 12),
        "K": intrinsics(H, W, focal), "poses": orbit_poses(n_poses), "H": H, "W": W,
        "reso": [cfg.reso] * 3, "xyz_min": XYZ_MIN.copy(), "xyz_max": XYZ_MAX.copy(),
        "near": cfg.near, "far": cfg.far, "stepdist": float(np.float32(cfg.stepsize * vs)),
        "act_shift": act_shift(cfg.alpha_init), "interval": cfg.stepsize,
        "fast_color_thres": 1e-4, "bg": cfg.bg, "inverse_y": False, "t_stop": t_stop,
    }
