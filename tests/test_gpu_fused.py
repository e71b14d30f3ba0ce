"""GPU parity of the fused (batched, BF16-MLP) training step against the FP64
oracle of the reference step.

The MLP weights are held in BF16 on the device (their fp32 masters are what
Adam updates), so the FP64 oracle evaluates the same BF16-valued weights; the
device carries features, activations and gradients as (hi, lo) BF16 pairs with
FP32 accumulation.  Tolerances: rendered RGB within 1e-3 max-abs; loss within
1e-3 relative; gradients (density, colour features, MLP) within 1e-3 relative
L2 error.  Integer work -- the sample set and per-ray counts -- must match
exactly.  The error against the oracle with the fp32 master weights (which adds
the BF16 rounding of the weights themselves) is printed, not asserted.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def bf16_weights(mlp):
    """The weight matrices as the device holds them (BF16); biases stay fp32."""
    out = dict(mlp)
    for k in ("w0", "w1", "w2"):
        out[k] = torch.from_numpy(np.ascontiguousarray(mlp[k], np.float32)).to(
            torch.bfloat16).double().numpy()
    return out


def oracle_for(oracle_mod, tr, sc, held_weights=True):
    origins, vm = oracle_mod.leaves_from_mask(sc["mask"])
    col = tr.color.float().cpu().numpy().astype(np.float64)
    den = tr.density.cpu().numpy().astype(np.float64)
    from plenvdb_amd.fused import unpack_mlp

    mlp = {k: v.astype(np.float64) for k, v in unpack_mlp(tr.mlp.cpu().numpy()).items()}
    if held_weights:
        mlp = bf16_weights(mlp)
    wm, we, wr = tr.weights
    osc = dict(sc, origins=origins, vmask=vm, density=den, color=col, mlp=mlp, weight_main=wm,
               weight_entropy_last=we, weight_rgbper=wr)
    return oracle_mod.reference_train_step(osc, prec="f64", with_mlp=True)


MSE = (1.0, 0.0, 0.0)
# the fine stage's loss weights (configs/default.py:73-74), the only stage with the MLP
FINE = (1.0, 0.001, 0.01)


@pytest.mark.parametrize("name,n_rays,bf16,weights", [
    ("c2", 512, False, MSE), ("c3", 512, False, MSE), ("c2", 384, True, MSE),
    ("c4", 256, True, MSE), ("c2", 384, False, FINE), ("c3", 256, True, FINE)])
def test_fused_forward_backward_vs_fp64_oracle(oracle_mod, name, n_rays, bf16, weights):
    from plenvdb_amd import scenes
    from plenvdb_amd.fused import FusedTrainer, unpack_mlp

    print("[case] fused fwd/bwd", name, n_rays, "bf16" if bf16 else "fp32", "loss weights",
          weights, flush=True)
    sc = scenes.make_scene(name, n_rays=n_rays)
    tr = FusedTrainer(sc, color_bf16=bf16, weight_main=weights[0],
                      weight_entropy_last=weights[1], weight_rgbper=weights[2])
    tr.forward_backward()
    torch.cuda.synchronize()
    ref = oracle_for(oracle_mod, tr, sc)
    recs = tr.records()
    # sample set: per-ray occupancy-kept counts are integer work -> exact
    ref_counts = np.bincount(ref["ray_id_mask"], minlength=n_rays)
    np.testing.assert_array_equal(recs["count"], ref_counts)
    assert recs["total"] == ref["n_mask"]
    rgb = tr.ray_out[:, :3].cpu().numpy()
    err_rgb = float(np.abs(rgb - ref["rgb_marched"]).max())
    loss = float(tr.loss.item())
    err_loss = abs(loss - ref["loss"]) / abs(ref["loss"])
    e_den = rel_l2(tr.grad_density.cpu().numpy(), ref["grad_density"])
    e_col = rel_l2(tr.grad_color.cpu().numpy(), ref["grad_color"])
    gm = unpack_mlp(tr.mlp_grad.cpu().numpy())
    e_mlp = {k: rel_l2(gm[k], ref["grad_mlp"][k]) for k in gm}
    print(f"[result] rgb max-abs {err_rgb:.2e}  loss rel {err_loss:.2e}  grad density rel-L2 "
          f"{e_den:.2e}  grad color rel-L2 {e_col:.2e}  mlp " +
          " ".join(f"{k} {v:.2e}" for k, v in e_mlp.items()), flush=True)
    ref32 = oracle_for(oracle_mod, tr, sc, held_weights=False)
    print(f"[info] against fp32 master weights: rgb max-abs "
          f"{float(np.abs(rgb - ref32['rgb_marched']).max()):.2e}  grad color rel-L2 "
          f"{rel_l2(tr.grad_color.cpu().numpy(), ref32['grad_color']):.2e}  grad density rel-L2 "
          f"{rel_l2(tr.grad_density.cpu().numpy(), ref32['grad_density']):.2e}", flush=True)
    assert err_rgb <= 1e-3
    assert err_loss <= 1e-3
    assert e_den <= 1e-3 and e_col <= 1e-3
    assert max(e_mlp.values()) <= 1e-3


def test_fused_step_applies_adam(oracle_mod):
    from plenvdb_amd import scenes
    from plenvdb_amd.fused import FusedTrainer

    print("[case] fused step adam c2", flush=True)
    sc = scenes.make_scene("c2", n_rays=256)
    tr = FusedTrainer(sc)
    ref = oracle_for(oracle_mod, tr, sc)
    d0 = tr.density.clone()
    c0 = tr.color.clone()
    tr.step()
    torch.cuda.synchronize()
    # first Adam step from m = v = 0 (densityvdb.cu:272-349 with stepmode 1):
    # delta = -stepsz * (0.1 g) / (1e-8 + sqrt(0.01 g^2)); zero gradients are skipped
    f = lambda x: float(np.float32(x))  # noqa: E731
    sz = tr._stepsz(0.1, 1)

    def expected(g):
        return -sz * ((1 - f(0.9)) * g) / (f(1e-8) + np.sqrt((1 - f(0.99)) * g * g))

    dd = (tr.density - d0).cpu().numpy()
    g = ref["grad_density"]
    assert (g != 0).sum() > 100
    e_d = rel_l2(dd, expected(g))
    dc = (tr.color - c0).cpu().numpy()
    e_c = rel_l2(dc, expected(ref["grad_color"]))
    print(f"[result] first-step update rel-L2: density {e_d:.2e} colour {e_c:.2e}", flush=True)
    assert e_d <= 1e-2 and e_c <= 1e-2
    assert np.all(dd[np.asarray(g) == 0] == 0)
    # gradients are re-zeroed for the next step
    assert float(tr.grad_density.abs().max()) == 0.0
    assert float(tr.grad_color.abs().max()) == 0.0
    for _ in range(3):
        tr.step()
    torch.cuda.synchronize()
    assert np.isfinite(float(tr.loss.item()))


@pytest.mark.parametrize("n_rays", [8192, 20001, 40000])
def test_fused_record_offsets_are_the_exclusive_scan(n_rays):
    """Per-ray record offsets are the exclusive prefix sum of the per-ray counts
    (integer work: exact), within one of the scan's 16,384-ray chunks, across
    two and three of them, and with ragged tails."""
    from plenvdb_amd import scenes
    from plenvdb_amd.fused import FusedTrainer

    print("[case] record offsets c2", n_rays, flush=True)
    sc = scenes.make_scene("c2", n_rays=n_rays)
    tr = FusedTrainer(sc)
    tr.forward_backward()
    torch.cuda.synchronize()
    recs = tr.records()
    cnt = recs["count"].astype(np.int64)
    exp = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    np.testing.assert_array_equal(recs["offset"].astype(np.int64), exp)
    assert int(cnt.sum()) == int(tr.total_records())


def test_fused_step_with_too_small_capacity():
    """A capacity below the step's record count: the step stays in bounds, covers
    the first `capacity` records, reports the overflow, and yields finite values."""
    from plenvdb_amd import scenes
    from plenvdb_amd.fused import FusedTrainer

    print("[case] fused step c2/256 rays with capacity 1000", flush=True)
    sc = scenes.make_scene("c2", n_rays=256)
    full = FusedTrainer(sc)
    full.forward_backward()
    torch.cuda.synchronize()
    n = full.total_records()
    assert n > 1000 and not full.overflowed()
    tr = FusedTrainer(sc, capacity=1000)
    tr.step()
    torch.cuda.synchronize()
    print(f"[result] {n} records, capacity 1000, overflowed {tr.overflowed()}", flush=True)
    assert tr.total_records() == n and tr.overflowed()
    assert torch.isfinite(tr.loss).all() and torch.isfinite(tr.ray_out).all()
    assert torch.isfinite(tr.density).all() and torch.isfinite(tr.color.float()).all()
    assert torch.isfinite(tr.mlp).all()


def test_fused_step_mlp_adam_first_step():
    """The MLP Adam of the fused step (stepmode 0, adam_upd_kernel.cu:8-58): from
    m = v = 0 the first step moves every parameter by
    -sz * (0.1 g) / (eps + sqrt(0.01 g^2)), g the step's MLP gradient (the
    reduction is deterministic, so a forward+backward of a twin trainer gives
    the same g)."""
    from plenvdb_amd import scenes
    from plenvdb_amd.fused import FusedTrainer

    print("[case] fused step MLP adam c2", flush=True)
    sc = scenes.make_scene("c2", n_rays=256)
    twin = FusedTrainer(sc)
    twin.forward_backward()
    torch.cuda.synchronize()
    g = twin.mlp_grad.cpu().numpy().astype(np.float64)
    tr = FusedTrainer(sc)
    p0 = tr.mlp.cpu().numpy().astype(np.float64)
    tr.step()
    torch.cuda.synchronize()
    np.testing.assert_array_equal(tr.mlp_grad.cpu().numpy(), twin.mlp_grad.cpu().numpy())
    f = lambda x: float(np.float32(x))  # noqa: E731
    sz = f(tr._stepsz(1e-3, 1))
    exp = p0 - sz * ((1 - f(0.9)) * g) / (f(1e-8) + np.sqrt((1 - f(0.99)) * g * g))
    got = tr.mlp.cpu().numpy().astype(np.float64)
    err = rel_l2(got - p0, exp - p0)
    print(f"[result] MLP first-step update rel-L2 {err:.2e}", flush=True)
    assert (g != 0).sum() > 1000
    assert err <= 1e-5
    assert np.all(got[g == 0] == p0[g == 0])


def test_fused_edge_rays_vs_fp64_oracle(oracle_mod):
    """Ragged and degenerate rays: 257 rays (not a multiple of the 4 rays per
    workgroup), ten of them pointing away from the box (no samples: their
    colour is the background), two axis-parallel; exact per-ray counts, RGB and
    gradients against the FP64 oracle."""
    from plenvdb_amd import scenes
    from plenvdb_amd.fused import FusedTrainer, unpack_mlp

    print("[case] fused edge rays c2/257", flush=True)
    sc = scenes.make_scene("c2", n_rays=257)
    rd = sc["rays_d"].copy()
    rd[:10] = -rd[:10]
    ro = sc["rays_o"].copy()
    # axis-parallel rays (zero direction components) through an active voxel
    idx = np.argwhere(sc["mask"])[len(np.argwhere(sc["mask"])) // 2]
    mn, mx = np.asarray(sc["xyz_min"], np.float64), np.asarray(sc["xyz_max"], np.float64)
    p = mn + idx / (np.asarray(sc["reso"]) - 1) * (mx - mn)
    ro[10] = [4.0, p[1], p[2]]
    rd[10] = [-1.0, 0.0, 0.0]
    ro[11] = [p[0], -4.0, p[2]]
    rd[11] = [0.0, 1.0, 0.0]
    vd = rd / np.linalg.norm(rd, axis=1, keepdims=True)
    sc.update(rays_o=ro.astype(np.float32), rays_d=rd.astype(np.float32),
              viewdirs=vd.astype(np.float32))
    tr = FusedTrainer(sc)
    tr.forward_backward()
    torch.cuda.synchronize()
    ref = oracle_for(oracle_mod, tr, sc)
    counts = tr.records()["count"]
    np.testing.assert_array_equal(counts, np.bincount(ref["ray_id_mask"], minlength=257))
    assert np.all(counts[:10] == 0) and counts[10] > 0 and counts[11] > 0
    rgb = tr.ray_out[:, :3].cpu().numpy()
    np.testing.assert_allclose(rgb[:10], float(sc["bg"]), atol=1e-6)
    err = float(np.abs(rgb - ref["rgb_marched"]).max())
    e_den = rel_l2(tr.grad_density.cpu().numpy(), ref["grad_density"])
    e_col = rel_l2(tr.grad_color.cpu().numpy(), ref["grad_color"])
    gm = unpack_mlp(tr.mlp_grad.cpu().numpy())
    e_mlp = max(rel_l2(gm[k], ref["grad_mlp"][k]) for k in gm)
    print(f"[result] rgb {err:.2e} grad density {e_den:.2e} colour {e_col:.2e} mlp {e_mlp:.2e}",
          flush=True)
    assert err <= 1e-3 and e_den <= 1e-3 and e_col <= 1e-3 and e_mlp <= 1e-3


@pytest.mark.parametrize("name,channels", [("c2", None), ("c1", 3)])
def test_fused_gradient_layouts_agree(name, channels):
    """The step's gradients as rows of one buffer (colour, density, padding:
    FusedTrainer's layout) equal those in separate [slots][C] and [slots]
    arrays (grad_stride 0), up to the order of the scatter's float atomics;
    the Adam step that follows is the same."""
    from plenvdb_amd import scenes
    from plenvdb_amd.fused import FusedTrainer

    print("[case] gradient layouts", name, channels, flush=True)
    sc = scenes.make_scene(name, n_rays=1024, channels=channels)
    tr = FusedTrainer(sc)
    tr.forward_backward()
    torch.cuda.synchronize()
    rows = (tr.grad_density.clone(), tr.grad_color.clone())
    tr.zero_grad()
    gd = torch.zeros_like(rows[0])
    gc = torch.zeros_like(rows[1])
    tr.desc.grad_density, tr.desc.grad_color, tr.desc.grad_stride = gd.data_ptr(), gc.data_ptr(), 0
    tr.forward_backward()
    torch.cuda.synchronize()
    assert float(tr.grad_rows.abs().max()) == 0.0  # nothing went to the rows
    for a, b in ((gd, rows[0]), (gc, rows[1])):
        scale = max(float(b.abs().max()), 1e-30)
        assert float(b.abs().max()) > 0
        assert float((a - b).abs().max()) <= 1e-6 * scale
    # Adam reads either layout: fed the same gradient values, the two leave
    # the same grids, bit for bit (the twin's own backward sets its touched
    # leaves, then its rows take the first run's values)
    twin = FusedTrainer(sc)
    twin.forward_backward()
    torch.cuda.synchronize()
    twin.grad_density.copy_(rows[0])
    twin.grad_color.copy_(rows[1])
    gd.copy_(rows[0])
    gc.copy_(rows[1])
    tr.adam_only()
    twin.adam_only()
    torch.cuda.synchronize()
    assert float(gd.abs().max()) == 0.0 and float(gc.abs().max()) == 0.0
    assert float(twin.grad_rows.abs().max()) == 0.0
    assert torch.equal(tr.density, twin.density) and torch.equal(tr.color, twin.color)
    assert torch.equal(tr.m_color, twin.m_color) and torch.equal(tr.v_density, twin.v_density)
