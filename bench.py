"""Benchmark of the PlenVDB hot path on one MI355X.

One step = one training iteration of the 160^3 configuration (64 spheres,
8192 rays, 12-channel features + view-dependent MLP): ray marching with
empty-space skip, density/colour trilinear gathers, alpha compositing, MLP,
MSE loss, the full backward and the Adam steps (VDB grids + MLP).

    python bench.py [--steps K] [--warmup W] [--path batched|plain] [--no-cpu]

Prints one JSON line.  `value` is ray-samples per second of the fused batched
path (a sample is a point whose density gather is evaluated: inside the bbox,
in an active voxel, before the alpha/weight culls).  The plain FP32 path (the
stage-per-kernel structure of the reference) is timed beside it, and the CPU
baseline is the oracle's FP32 restatement of the reference step timed on the
host cores on a bounded sample of the same workload.  The `render` object is
the merged-VDB renderer on the 800x800 render-only workload (frames/s and
ray-samples/s).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "samples/s fwd+bwd, BF16+batched vs FP32 plain @<=1e-3 err; 800^2 fps"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--path", default="auto", choices=["auto", "batched", "plain"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="built-in launch shapes instead of the autotuner's choice")
    ap.add_argument("--launch", default=None,
                    help="fixed launch shapes of the headline step instead of tuning: "
                         "mlp_fwd_workgroups,scatter_workgroups,scatter_chunk_records"
                         "[,ray_waves]")
    ap.add_argument("--configs", default="c1,c3,c4",
                    help="other workloads whose fused step is timed beside the headline one "
                         "(comma list, empty for none)")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip timing the plain FP32 path beside the batched one")
    ap.add_argument("--cpu-rays", type=int, default=8192)
    ap.add_argument("--render-frames", type=int, default=100,
                    help="poses of the merged-VDB render benchmark (0 skips it)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run only a device copy of known size (counter calibration)")
    return ap.parse_args()


def unique_corner_voxels(xs, ys, zs) -> int:
    """Number of distinct voxels among the 8 trilinear corners of the points."""
    i0 = np.floor(np.stack([xs, ys, zs], 1)).astype(np.int64)
    keys = []
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                c = i0 + np.array([dx, dy, dz])
                keys.append(((c[:, 0] + 16) << 40) | ((c[:, 1] + 16) << 20) | (c[:, 2] + 16))
    return int(np.unique(np.concatenate(keys)).size)


def kernel_bytes(name: str, n: int, uniq: int, channels: int) -> float:
    """Algorithmic HBM bytes of one launch: each input/output element once, each
    distinct grid voxel once (read; read+write for the atomic scatter)."""
    c = channels
    if name == "pvdb_trilinear_fwd":
        return n * (12 + 4 * c) + uniq * 4 * c
    if name == "pvdb_trilinear_bwd":
        return n * (12 + 4 * c) + uniq * 4 * c * 2
    return float("nan")


def copy_probe():
    import torch

    n = 256 << 20
    a = torch.empty(n, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n}))


def bench_plain(args, scene):
    import torch

    from plenvdb_amd import lib
    from plenvdb_amd.trainer import PlainTrainer

    tr = PlainTrainer(scene)
    for _ in range(args.warmup):
        tr.step()
    torch.cuda.synchronize()
    samples = []
    with lib.CallTimer() as timer:
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        t0.record()
        for _ in range(args.steps):
            _, out = tr.step()
            samples.append(int(out["n_samples"]))
        t1.record()
        torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / args.steps
    totals = timer.totals()
    # per-launch workload of the colour scatter / gathers, from one more step
    pts = tr.model.sample_ray(tr.rays_o, tr.rays_d, **tr.render_kwargs)[0]
    keep = tr.model.mask_cache(pts)
    idx = tr.model.density.wld2idx(pts[keep]).cpu().numpy()
    final = out["ray_id"].numel()
    return {
        "ms_per_step": ms,
        "samples_per_step": float(np.mean(samples)),
        "totals": totals,
        "points_density": idx,
        "n_color": final,
        "n_rays": int(tr.rays_o.shape[0]),
    }


MFMA_BF16_PEAK_TFLOPS = 2500.0
# MLP FLOPs per sample / per ray (PE-factorised first layer): forward 36,608 +
# 6,912/ray; backward 73,216 + 6,912/ray.
MLP_FWD_FLOP, MLP_BWD_FLOP, MLP_RAY_FLOP = 36608, 73216, 6912


def bench_batched(args, scene):
    import torch

    from plenvdb_amd import fused

    # the first forward+backward of a fresh trainer, kept for the accuracy check
    # that runs beside the CPU baseline
    twin = fused.FusedTrainer(scene)
    twin.forward_backward()
    torch.cuda.synchronize()
    first = {"rgb": twin.ray_out[:, :3].cpu().numpy(), "loss": float(twin.loss.item()),
             "grad_density": twin.grad_density.cpu().numpy(),
             "grad_color": twin.grad_color.cpu().numpy(),
             "grad_mlp": twin.mlp_grad.cpu().numpy(), "mlp": twin.mlp.cpu().numpy(),
             "density": twin.density.cpu().numpy(), "color": twin.color.float().cpu().numpy()}
    del twin
    tr = fused.FusedTrainer(scene)
    if args.launch:
        from plenvdb_amd import autotune

        v = [int(x) for x in args.launch.split(",")]
        if len(v) not in (3, 4):
            raise SystemExit("--launch takes three or four integers")
        launch = v + [0] if len(v) == 3 else v
        autotune.apply(tr, launch)
        tuned = {"launch": {autotune.NAMES[i]: launch[i] for i in autotune.NAMES},
                 "source": "given"}
    elif args.no_tune:
        tuned = {"launch": None, "source": "built-in defaults"}
    else:
        tuned = tr.autotune()
    for _ in range(args.warmup):
        tr.step()
    torch.cuda.synchronize()
    fused.stage_timing(1)
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(args.steps):
        tr.step()
    t1.record()
    torch.cuda.synchronize()
    stages = fused.stage_timing(0)
    ms_eager = t0.elapsed_time(t1) / args.steps
    # forward + backward alone (no optimizer step), eager launches
    t0.record()
    for _ in range(args.steps):
        tr.forward_backward()
    t1.record()
    torch.cuda.synchronize()
    ms_fwd_bwd = t0.elapsed_time(t1) / args.steps
    tr.zero_grad()
    # the same steps replayed from one captured hipGraph (the production path)
    tr.capture_graph()
    for _ in range(args.warmup):
        tr.step()
    torch.cuda.synchronize()
    t0.record()
    for _ in range(args.steps):
        tr.step()
    t1.record()
    torch.cuda.synchronize()
    ms_graph = t0.elapsed_time(t1) / args.steps
    ms = ms_graph
    samples = tr.total_records()
    # workload descriptors of one more (untimed) forward/backward
    tr.zero_grad()
    tr.forward_backward()
    torch.cuda.synchronize()
    rec = tr.records()
    flags = rec["ab"][:, 1].view(np.uint32)
    alive_w = (flags & np.uint32(1 << 31)) != 0
    gd = tr.grad_density.cpu().numpy()
    gc = tr.grad_color.cpu().numpy()
    touched_leaf = (gd.reshape(-1, 512) != 0).any(1) | (gc.reshape(-1, 512 * 12) != 0).any(1)
    act = tr.topo.active_slot_mask()
    v_u = int(((gd != 0) & act).sum())
    tr.zero_grad()
    pts = rec["xyze"][:, :3]
    return {"first_step": first,
            **{
        "ms_per_step": ms,
        "ms_per_step_eager": ms_eager,
        "ms_fwd_bwd_eager": ms_fwd_bwd,
        "launch": tuned,
        "samples_per_step": float(samples),
        "stages": stages,
        "n_color": int(alive_w.sum()),
        "uniq_all": unique_corner_voxels(pts[:, 0], pts[:, 1], pts[:, 2]),
        "uniq_w": unique_corner_voxels(pts[alive_w, 0], pts[alive_w, 1], pts[alive_w, 2]),
        "l_t": int(touched_leaf.sum()),
        "v_u": v_u,
        "n_rays": tr.n_rays,
        "color_bytes": 2 if tr.color_bf16 else 4,
        "workload": args.config,
        "n_cu": min(512, torch.cuda.get_device_properties(0).multi_processor_count),
    }}


# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on
# the default workloads, scaled by a 256 MiB copy's counts (FETCH x2.0, WRITE
# x1.0); profiles/pmc_traffic_c2_train_c5_render.md holds the runs.
TRAFFIC_SOURCE = "profiles/pmc_traffic_c2_train_c5_render.md"
MEASURED_TRAFFIC_MB = {
    "c2": {"mlp_bwd": 63.71, "scatter": 85.97, "mlp_fwd": 71.7, "ray_march": 13.9,
           "composite": 15.69, "adam": 122.02, "ray_count": 6.23},
    "c5": {"render_mlp": 656.47},
}


def stage_models(r):
    """Algorithmic work of one launch of each fused stage: FLOPs for the MLP
    stages (bounded by BF16 MFMA), HBM bytes for the others (each input and
    output element once, each distinct grid voxel once)."""
    S, N, sc, nw = r["samples_per_step"], r["n_rays"], r["color_bytes"], r["n_color"]
    ua, uw, L_t, V_u = r["uniq_all"], r["uniq_w"], r["l_t"], r["v_u"]
    P = 22019
    return {
        # rays in, counts, the kept steps' 2-byte indices, the rays' 168-byte
        # encoding rows and 32-byte set-ups out; the MLP parameters packed to
        # bf16 and the MLP gradient zeroed by the same launch
        "ray_count": ("hbm", N * (12 + 12 + 4 + 12 + 168 + 32) + S * 2
                      + P * 4 + 2 * (128 * 64 + 2 * 128 * 128 + 2 * 16 * 128 + 128 * 32) + P * 4),
        # ray set-ups + offsets and the kept step indices in, 24-byte records
        # out (xyze, alpha + flags), distinct density voxels once
        "ray_march": ("hbm", N * (32 + 8) + S * (2 + 24) + ua * 4),
        # 36,608 FLOP per MLP sample + 6,912 per ray (PE part of layer 0)
        "mlp_fwd": ("mfma", nw * MLP_FWD_FLOP + N * MLP_RAY_FLOP),
        # alpha + flags and exp(sigma) in, adjoints out per record; rgb in per
        # weighted record; rays in/out
        "composite": ("hbm", S * (8 + 4 + 16) + nw * 16 + N * (12 + 16 + 16 + 16 + 4)),
        # 73,216 FLOP per MLP sample + 6,912 per ray (dW of the PE block)
        "mlp_bwd": ("mfma", nw * MLP_BWD_FLOP + N * MLP_RAY_FLOP),
        # records + per-record gradients in, distinct voxels' 13 gradients
        # read+written; the same launch reduces the dW partials (one per CU in,
        # MLP gradient read+written) and applies the MLP Adam (parameters and
        # both moments read+written), and sums the per-ray loss terms
        "scatter": ("hbm", S * (16 + 8 + 64) + uw * 13 * 4 * 2
                    + r["n_cu"] * P * 4 + 2 * P * 4 + 6 * P * 4 + N * 4),
        # B_adam: density p,m,v r/w + g r; colour p,m,v r/w + g r
        "adam": ("hbm", V_u * (28 + 12 * (2 * sc + 20)) + L_t * 512 * 13 * 4),
    }


def measured_traffic(workload: str, stage: str):
    mb = MEASURED_TRAFFIC_MB.get(workload, {}).get(stage)
    return None if mb is None else round(mb * 1e6)


def stage_roofline(name, kind, work, avg_s, workload="c2"):
    traffic = measured_traffic(workload, name)
    if kind == "mfma":
        achieved = work / avg_s / 1e12
        out = {"kernel": name, "bound": "mfma", "achieved": round(achieved, 3),
               "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
               "frac": round(achieved / MFMA_BF16_PEAK_TFLOPS, 6), "traffic": traffic,
               "algorithmic_flop": float(work), "avg_launch_ms": round(avg_s * 1e3, 5)}
    else:
        achieved = work / avg_s / 1e9
        out = {"kernel": name, "bound": "hbm", "achieved": round(achieved, 2),
               "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
               "traffic": traffic, "algorithmic_bytes": float(work),
               "avg_launch_ms": round(avg_s * 1e3, 5)}
    if traffic is not None:
        out["traffic_source"] = TRAFFIC_SOURCE
    return out


def batched_roofline(r, steps):
    """Roofline of the dominant stage, every stage's fraction, and the whole-step
    byte model (B_step = B_grid + B_grad + B_adam + B_ray)."""
    st = r["stages"]
    models = stage_models(r)
    table = {}
    for name, (tot, calls) in st.items():
        if calls and name in models:
            kind, work = models[name]
            table[name] = stage_roofline(name, kind, work, tot / calls * 1e-3, r["workload"])
    dom = max((k for k in st if k in table), key=lambda k: st[k][0])
    roof = dict(table[dom])
    N, sc, L_t, V_u = r["n_rays"], r["color_bytes"], r["l_t"], r["v_u"]
    b_step = (L_t * 512 * (4 + 12 * sc) + L_t * 512 * 52 + V_u * (28 + 12 * (2 * sc + 20)) + N * 60)
    step = {"bytes": b_step, "achieved_GBps": round(b_step / (r["ms_per_step"] * 1e-3) / 1e9, 2),
            "frac": round(b_step / (r["ms_per_step"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
            "leaves_touched": L_t, "voxels_updated": V_u}
    stage_fracs = {k: {"unit": v["unit"], "achieved": v["achieved"], "frac": v["frac"]}
                   for k, v in table.items()}
    return roof, step, stage_fracs


def bench_config(name: str, steps: int, warmup: int):
    """The fused step (hipGraph replay) on another workload of the set: its
    default rays, timed with HIP events over `steps` replays."""
    import torch

    from plenvdb_amd import fused, scenes

    sc = scenes.make_scene(name)
    tr = fused.FusedTrainer(sc)
    tuned = tr.autotune()
    tr.capture_graph()
    for _ in range(warmup):
        tr.step()
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(steps):
        tr.step()
    t1.record()
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / steps
    samples = tr.total_records()
    cfg = sc["cfg"]
    out = {"workload": f"{cfg.name}: {cfg.reso}^3 grid, {cfg.n_spheres} spheres"
                       f"{' + shell' if cfg.shell else ''}, {cfg.n_rays} rays"
                       f"{' (far in [%g, %g])' % (cfg.far_lo, cfg.far_hi) if cfg.far_lo else ''}, "
                       f"{'bf16' if tr.color_bf16 else 'fp32'} colour"
                       f"{'' if cfg.use_mlp else ', no MLP (rgb = sigmoid of the %d-ch k0)' % cfg.channels}"
                       ", fwd+bwd+Adam",
           "ms_per_step": round(ms, 4), "samples_per_step": int(samples),
           "samples_per_s": round(samples / (ms * 1e-3), 1), "launch": tuned["launch"]}
    if samples > tr.capacity:
        out["note"] = f"records capped at capacity {tr.capacity}"
    del tr
    torch.cuda.empty_cache()
    return out


def roofline_from(totals, launches_bytes):
    """Pick the kernel with the largest total device time among those with a byte model."""
    best = None
    for name, (tot_ms, cnt) in totals.items():
        if name not in launches_bytes:
            continue
        avg_ms = tot_ms / cnt
        if best is None or tot_ms > best[1]:
            best = (name, tot_ms, avg_ms)
    if best is None:
        return None
    name, _, avg_ms = best
    b = launches_bytes[name]
    achieved = b / (avg_ms * 1e-3) / 1e9
    return {"kernel": name, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
            "algorithmic_bytes": b, "avg_launch_ms": round(avg_ms, 5)}


def bench_render(frames: int, warmup: int):
    """Merged-VDB rendering (c5): c4's topology as an index grid with fp16 data,
    800x800 pinhole camera (focal 1111), orbit poses at radius 4 and -30 deg,
    step 0.5 voxel, early stop at T < 1e-4.  Frames/s over `frames` poses, timed
    with HIP events on the render stream; the per-frame sample counts come from
    an untimed second pass."""
    import torch

    from plenvdb_amd import scenes, vdbio
    from plenvdb_amd.renderer import MGRenderer

    ms = scenes.merged_scene("c4", H=800, W=800, focal=1111.0, n_poses=frames)
    r = MGRenderer(12, 27, 128, 3)
    p = ms["mlp"]
    r.load_params(p["w0"], p["b0"], p["w1"], p["b1"], p["w2"], p["b2"])
    topo, slot_idx = vdbio.index_grid_arrays(ms["idx"])
    den = ms["den"].astype(np.float32)
    r.load_data_grid(den, ms["col"].astype(np.float32).reshape(-1), topo, slot_idx, len(den))
    r.setScene(ms["reso"], ms["K"], ms["xyz_min"], ms["xyz_max"])
    r.setKwargs(ms["near"], ms["far"], ms["stepdist"], ms["act_shift"], ms["interval"],
                ms["fast_color_thres"], ms["bg"], ms["inverse_y"], ms["H"], ms["W"])
    r.set_t_stop(ms["t_stop"])
    poses = ms["poses"]
    # one synchronous render per pose: the record capacity grows to the largest frame
    for i in range(frames):
        r.input_a_c2w(poses[i])
        r.render_an_image()
    for i in range(warmup):
        r.input_a_c2w(poses[i % len(poses)])
        r.render_async()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for i in range(frames):
        r.input_a_c2w(poses[i])
        r.render_async()
    e1.record()
    torch.cuda.synchronize()
    ms_frame = e0.elapsed_time(e1) / frames
    gathers = kept = 0
    stage_ms = np.zeros(3)
    for i in range(frames):
        r.input_a_c2w(poses[i])
        r.render_async()
        g, k = r.sample_counts()
        gathers += g
        kept += k
    # per-kernel times (HIP events around the launches), on a second pass
    r.stage_timing(True)
    for i in range(frames):
        r.input_a_c2w(poses[i])
        r.render_async()
    torch.cuda.synchronize()
    stage_ms = r.stage_timing(False) / frames
    n_px = ms["H"] * ms["W"]
    flop = (kept * MLP_FWD_FLOP + frames * n_px * MLP_RAY_FLOP) / frames
    tflops = flop / (ms_frame * 1e-3) / 1e12
    mlp_ms = float(stage_ms[1])
    mlp_tflops = flop / (mlp_ms * 1e-3) / 1e12 if mlp_ms > 0 else 0.0
    return {
        "workload": "c5: c4 topology (320^3, 1959 leaves) as a merged index grid, fp16 data, "
                    f"800x800, focal 1111, {frames} orbit poses, T<1e-4",
        "frames_per_s": round(1e3 / ms_frame, 2),
        "ms_per_frame": round(ms_frame, 4),
        "samples_per_s": round(gathers / frames / (ms_frame * 1e-3), 1),
        "samples_per_frame": round(gathers / frames, 1),
        "kept_per_frame": round(kept / frames, 1),
        "kernels_ms": {"march": round(float(stage_ms[0]), 4), "mlp": round(mlp_ms, 4),
                       "composite": round(float(stage_ms[2]), 4)},
        "frame_tflops": round(tflops, 3),
        "roofline": {"kernel": "k_render_mlp", "bound": "mfma", "achieved": round(mlp_tflops, 3),
                     "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(mlp_tflops / MFMA_BF16_PEAK_TFLOPS, 6),
                     "traffic": measured_traffic("c5", "render_mlp"),
                     "traffic_source": TRAFFIC_SOURCE,
                     "algorithmic_flop": flop, "avg_launch_ms": round(mlp_ms, 5)},
    }


def bench_tv(cfg_name: str, reps: int = 20):
    """VDB total-variation gradient (pvdb_tv_add_grad, dense mode) on the
    workload's density (1 channel) and colour (12 channel) grids: average launch
    time over `reps` launches with HIP events on the launch stream.  Algorithmic
    bytes per slot in the box: its values read, its gradient read and written
    (12 B per channel); neighbour values are counted once, as own values."""
    import torch

    from plenvdb_amd import scenes
    from plenvdb_amd.plenvdb import ColorVDB, DensityVDB
    from plenvdb_amd.topology import Topology

    cfg = scenes.CONFIGS[cfg_name]
    topo = Topology.from_mask(scenes.occupancy_mask(cfg))
    reso = [cfg.reso] * 3
    out = {"workload": f"{cfg_name}: {cfg.reso}^3 grid, {topo.nslots // 512} leaves, dense mode"}
    for name, vdb in (("density", DensityVDB(reso, 1, topology=topo)),
                      ("color", ColorVDB(reso, 12, topology=topo))):
        g = torch.Generator(device="cuda").manual_seed(0)
        vdb.grid.copy_(torch.randn(vdb.grid.shape, device="cuda", generator=g) * 0.5)
        vdb.grad.copy_(torch.randn(vdb.grad.shape, device="cuda", generator=g))
        for _ in range(3):
            vdb.tv_add_grad(1.0, 1.0, 1.0, True)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            vdb.tv_add_grad(1.0, 1.0, 1.0, True)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        nbytes = topo.nslots * 12 * vdb.ndim
        gbs = nbytes / (ms * 1e-3) / 1e9
        out[name] = {"avg_launch_ms": round(ms, 5), "bytes": nbytes, "achieved": round(gbs, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 5)}
    return out


def fused_accuracy(scene: dict, first: dict) -> dict:
    """Error of the fused step's first forward+backward (rgb, loss, gradients)
    against the oracle's FP64 reference step, evaluated once with the bf16
    weights the step holds and once with the fp32 master weights."""
    import torch

    from oracle import oracle
    from plenvdb_amd.fused import unpack_mlp

    def rel(a, b):
        a, b = np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel()
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))

    origins, vm = oracle.leaves_from_mask(scene["mask"])
    mlp32 = {k: v.astype(np.float64) for k, v in unpack_mlp(first["mlp"]).items()}
    held = dict(mlp32)
    for k in ("w0", "w1", "w2"):
        held[k] = torch.from_numpy(np.ascontiguousarray(mlp32[k], np.float32)).to(
            torch.bfloat16).double().numpy()
    out = {}
    for name, mlp in (("vs_fp64_held_bf16_weights", held), ("vs_fp64_fp32_weights", mlp32)):
        ref = oracle.reference_train_step(
            dict(scene, origins=origins, vmask=vm, mlp=mlp,
                 density=first["density"].astype(np.float64),
                 color=first["color"].astype(np.float64)), prec="f64")
        gm = unpack_mlp(first["grad_mlp"])
        out[name] = {
            "rgb_max_abs": float(np.abs(first["rgb"] - ref["rgb_marched"]).max()),
            "loss_rel": abs(first["loss"] - ref["loss"]) / abs(ref["loss"]),
            "grad_density_rel_l2": rel(first["grad_density"], ref["grad_density"]),
            "grad_color_rel_l2": rel(first["grad_color"], ref["grad_color"]),
            "grad_mlp_rel_l2_max": max(rel(gm[k], ref["grad_mlp"][k]) for k in gm)}
    return out


def cpu_baseline(cfg_name: str, n_rays: int, threads: int | None = None):
    """Oracle FP32 restatement of the reference step (C kernels under OpenMP,
    MLP/composite/loss in torch CPU), forward + backward + Adam, timed on the
    host cores (all of them, or `threads`) on n_rays rays of the same workload."""
    import torch

    from oracle import oracle
    from plenvdb_amd import scenes

    sc = scenes.make_scene(cfg_name, n_rays=n_rays)
    origins, vm = oracle.leaves_from_mask(sc["mask"])
    nslots = len(origins) * 512
    den, col = scenes.grid_values(sc["cfg"], nslots)
    sc.update(origins=origins, vmask=vm, density=den, color=col)
    act = vm.reshape(-1).astype(np.uint8)
    m_d, v_d = np.zeros_like(den), np.zeros_like(den)
    m_c, v_c = np.zeros_like(col), np.zeros_like(col)
    cores = threads or int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    cores = oracle.set_threads(cores)
    times, samples = [], []
    warm, timed = 3, 20
    for it in range(warm + timed):
        t = time.perf_counter()
        out = oracle.reference_train_step(sc, prec="f32", with_mlp=sc["cfg"].use_mlp)
        sz = oracle.adam_stepsz(0.1, 0.9, 0.99, it + 1)
        gd = np.ascontiguousarray(out["grad_density"].astype(np.float32))
        gc = np.ascontiguousarray(out["grad_color"].astype(np.float32))
        oracle.vdb_adam(den, gd, m_d, v_d, act, 1, 1, sz, 1e-8, 0.9, 0.99)
        oracle.vdb_adam(col, gc, m_c, v_c, act, 3, 1, sz, 1e-8, 0.9, 0.99)
        times.append(time.perf_counter() - t)
        samples.append(out["n_mask"])
    dt = float(np.median(times[warm:]))
    return {"value": round(float(np.mean(samples)) / dt, 1), "unit": "samples/s", "cores": cores,
            "kind": "port", "cpu_model": cpu_model(),
            "sample": f"{cfg_name} workload, {n_rays} of its rays, fwd+bwd+VDB Adam, median of "
                      f"{timed} steps after {warm} warm-ups ({int(np.mean(samples))} samples/step, "
                      f"{dt * 1e3:.0f} ms/step)"}


def cpu_model() -> str:
    """The host CPU's model name (from /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); runs over several GPUs "
              "are not implemented", file=sys.stderr)
        sys.exit(2)
    import torch

    if args.copy_probe:
        copy_probe()
        return
    from plenvdb_amd import scenes

    scene = scenes.make_scene(args.config)
    cfg = scene["cfg"]
    path = "batched" if args.path == "auto" else args.path
    res = {}
    if path == "batched" or args.path == "auto":
        res["batched"] = bench_batched(args, scene)
    if path == "plain" or not args.no_plain:
        res["plain"] = bench_plain(args, scene)
    main_res = res[path]
    step_roof = stage_fracs = None
    if path == "batched":
        roof, step_roof, stage_fracs = batched_roofline(main_res, args.steps)
    else:
        p = main_res["points_density"]
        uniq_d = unique_corner_voxels(p[:, 0], p[:, 1], p[:, 2])
        n_col = main_res["n_color"]
        nb = {"pvdb_trilinear_bwd": kernel_bytes("pvdb_trilinear_bwd", n_col, uniq_d, 12),
              "pvdb_trilinear_fwd": kernel_bytes("pvdb_trilinear_fwd", n_col, uniq_d, 12)}
        roof = roofline_from(main_res["totals"], nb)
    value = main_res["samples_per_step"] / (main_res["ms_per_step"] * 1e-3)
    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_res["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if path == "plain" else "bf16",
        "data": "synthetic",
        "config": {"workload": f"{cfg.name}: {cfg.reso}^3 grid, {cfg.n_spheres} spheres, "
                               f"{cfg.n_rays} rays, 12-ch features + MLP 39-128-128-3, "
                               "fwd+bwd+Adam", "path": path,
                   "samples_per_step": round(main_res["samples_per_step"], 1)},
        "roofline": roof,
        "paths": {k: {"ms_per_step": round(v["ms_per_step"], 4),
                      "samples_per_s": round(v["samples_per_step"] / (v["ms_per_step"] * 1e-3), 1)}
                  for k, v in res.items()},
        "kernels_ms": ({k: round(v[0] / args.steps, 4) for k, v in sorted(
            main_res["stages"].items(), key=lambda kv: -kv[1][0]) if v[1] > 0}
            if path == "batched" else
            {k: round(v[0] / args.steps, 4) for k, v in sorted(
                main_res["totals"].items(), key=lambda kv: -kv[1][0])}),
        "step_bytes_model": step_roof,
        "stage_roofline": stage_fracs,
        "rays_per_s": round(main_res["n_rays"] / (main_res["ms_per_step"] * 1e-3), 1),
    }
    if "ms_per_step_eager" in main_res:
        line["ms_per_step_eager"] = round(main_res["ms_per_step_eager"], 4)
        line["ms_fwd_bwd_eager"] = round(main_res["ms_fwd_bwd_eager"], 4)
        line["launch"] = main_res["launch"]
        line["step_launch"] = "hipGraph replay (eager launch timed beside it)"
    if path == "batched":
        line["adam_voxels_per_s"] = round(main_res["v_u"] / (main_res["ms_per_step"] * 1e-3), 1)
    if "batched" in res and "plain" in res:
        line["speedup_batched_vs_plain"] = round(res["plain"]["ms_per_step"] / res["batched"]["ms_per_step"]
                                                 * res["batched"]["samples_per_step"]
                                                 / res["plain"]["samples_per_step"], 3)
    if args.render_frames > 0:
        line["render"] = bench_render(args.render_frames, 3)
    line["tv_add_grad"] = bench_tv(args.config)
    others = [c for c in args.configs.split(",") if c and c != args.config]
    if others:
        line["configs"] = {c: bench_config(c, 5, 2) for c in others}
    if not args.no_cpu:
        if "batched" in res:
            # the oracle as the checker of the measured step's numbers
            line["accuracy"] = fused_accuracy(scene, res["batched"]["first_step"])
        line["cpu_baseline"] = cpu_baseline(args.config, args.cpu_rays)
        # and one host thread, on an eighth of the rays
        line["cpu_baseline_1thread"] = cpu_baseline(args.config, max(args.cpu_rays // 8, 64), 1)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
