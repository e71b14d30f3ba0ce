"""GPU parity of the fused step on models without the MLP: rgb = sigmoid(k0)
over the colour grid (dvgo.py:72-75,344-346).  Cases: the reference's coarse
stage, with a 3-channel k0 grid, its loss weights (configs/default.py:52-55)
and VDBAdam stepmodes (density 2 with a per-voxel rate, colour 0;
masked_adam.py:43-68); and the 12-channel c1 configuration.

* Forward/backward against the FP64 oracle (reference_train_step with
  with_mlp=False).  The per-ray sample counts are integer work and must match
  exactly.  The rest is FP32 arithmetic with no BF16 operand, so it is held
  to the plain path's bar: rendered RGB within 1e-5 max-abs, loss within 1e-5
  relative, and density and colour gradients within 1e-5 relative L2.
* Three Adam steps (`adam_only`) fed the device's own accumulated gradients,
  and a fourth on all-zero gradients, must equal the oracle's VDB Adam bit
  for bit, in stepmodes (1, 1), (0, 0) and (2, 0).  On the zero gradients,
  stepmodes 0 and 2 still move the voxels with momentum, and stepmode 1
  moves none.
* 100 fused steps must follow 100 steps of the plain FP32 path on c1 within
  1e-3 relative loss.
"""
import numpy as np
import pytest
import torch

from test_gpu_fused import rel_l2

pytestmark = pytest.mark.gpu

MSE = (1.0, 0.0, 0.0)
COARSE = (1.0, 0.01, 0.1)  # configs/default.py:52-55


def direct_oracle(oracle_mod, tr, sc):
    origins, vm = oracle_mod.leaves_from_mask(sc["mask"])
    wm, we, wr = tr.weights
    osc = dict(sc, origins=origins, vmask=vm, density=tr.density.cpu().numpy().astype(np.float64),
               color=tr.color.cpu().numpy().astype(np.float64), weight_main=wm,
               weight_entropy_last=we, weight_rgbper=wr)
    return oracle_mod.reference_train_step(osc, prec="f64", with_mlp=False)


@pytest.mark.parametrize("n_rays,channels,weights", [
    (1024, 12, MSE), (4096, 12, MSE), (2048, 3, COARSE), (4096, 3, MSE)])
def test_direct_fused_vs_fp64_oracle(oracle_mod, n_rays, channels, weights):
    from plenvdb_amd import scenes
    from plenvdb_amd.fused import FusedTrainer

    print("[case] fused without MLP c1", n_rays, "rays,", channels, "channels, loss weights",
          weights, flush=True)
    sc = scenes.make_scene("c1", n_rays=n_rays, channels=channels)
    tr = FusedTrainer(sc, weight_main=weights[0], weight_entropy_last=weights[1],
                      weight_rgbper=weights[2])
    tr.forward_backward()
    torch.cuda.synchronize()
    recs = tr.records()
    ref = direct_oracle(oracle_mod, tr, sc)
    np.testing.assert_array_equal(recs["count"], np.bincount(ref["ray_id_mask"], minlength=n_rays))
    assert recs["total"] == ref["n_mask"]
    out = tr.ray_out.cpu().numpy()
    assert out.shape == (n_rays, channels + 1)
    err_rgb = float(np.abs(out[:, :channels] - ref["rgb_marched"]).max())
    err_tl = float(np.abs(out[:, channels] - ref["alphainv_last"]).max())
    err_loss = abs(float(tr.loss.item()) - ref["loss"]) / abs(ref["loss"])
    gd, gc = tr.grad_density.cpu().numpy(), tr.grad_color.cpu().numpy()
    e_den, e_col = rel_l2(gd, ref["grad_density"]), rel_l2(gc, ref["grad_color"])
    print(f"[result] {recs['total']} records  rgb max-abs {err_rgb:.2e}  T_last max-abs "
          f"{err_tl:.2e}  loss rel {err_loss:.2e}  grad density rel-L2 {e_den:.2e}  grad colour "
          f"rel-L2 {e_col:.2e}", flush=True)
    assert np.count_nonzero(gc) > 1000 and np.count_nonzero(gd) > 1000
    assert err_rgb <= 1e-5 and err_tl <= 1e-5 and err_loss <= 1e-5
    assert e_den <= 1e-5 and e_col <= 1e-5


def test_direct_long_rays_vs_fp64_oracle(oracle_mod):
    """Rays with more occupancy-kept steps than the counting march stores per
    ray (256 step indices), which the forward march marches again, beside rays
    that read the stored indices, in one launch: a fully active c1 grid (3
    channels) with each ray's far bound drawn between 40% and all of its path
    through the box, so the kept counts spread from ~100 to ~390 steps.  Exact per-ray counts
    and the FP32 bars of the direct path against the FP64 oracle."""
    from plenvdb_amd import scenes
    from plenvdb_amd.fused import FusedTrainer

    print("[case] fused without MLP, long rays: c1/256, 3 channels, fully active grid",
          flush=True)
    n = 256
    sc = scenes.make_scene("c1", n_rays=n, channels=3)
    sc["mask"] = np.ones_like(sc["mask"])
    stepdist = float(np.float32(sc["stepsize"]) * np.float32(sc["voxel_size"]))
    t0, t1, _ = oracle_mod.ray_bounds(sc["rays_o"], sc["rays_d"], sc["xyz_min"], sc["xyz_max"],
                                      sc["near"], 1e9, stepdist)
    u = np.random.default_rng(3).uniform(0.4, 1.0, n)
    sc["far_per_ray"] = (np.asarray(t0) + u * (np.asarray(t1) - np.asarray(t0))).astype(np.float32)
    tr = FusedTrainer(sc)
    tr.forward_backward()
    torch.cuda.synchronize()
    ref = direct_oracle(oracle_mod, tr, sc)
    counts = tr.records()["count"]
    np.testing.assert_array_equal(counts, np.bincount(ref["ray_id_mask"], minlength=n))
    n_long = int((counts > 256).sum())
    out = tr.ray_out.cpu().numpy()
    err_rgb = float(np.abs(out[:, :3] - ref["rgb_marched"]).max())
    err_loss = abs(float(tr.loss.item()) - ref["loss"]) / abs(ref["loss"])
    e_den = rel_l2(tr.grad_density.cpu().numpy(), ref["grad_density"])
    e_col = rel_l2(tr.grad_color.cpu().numpy(), ref["grad_color"])
    print(f"[result] {n_long} of {n} rays keep more than 256 steps (max {int(counts.max())}); "
          f"rgb {err_rgb:.2e} loss {err_loss:.2e} grad density {e_den:.2e} colour {e_col:.2e}",
          flush=True)
    assert 30 <= n_long <= n - 30
    assert err_rgb <= 1e-5 and err_loss <= 1e-5 and e_den <= 1e-5 and e_col <= 1e-5


@pytest.mark.parametrize("channels,modes", [(12, (1, 1)), (3, (0, 0)), (3, (2, 0))])
def test_direct_adam_bit_exact_in_each_stepmode(oracle_mod, channels, modes):
    from plenvdb_amd import scenes
    from plenvdb_amd.fused import FusedTrainer

    print("[case] fused Adam without MLP c1/1024", channels, "channels, stepmodes", modes,
          flush=True)
    sc = scenes.make_scene("c1", n_rays=1024, channels=channels)
    tr = FusedTrainer(sc, stepmode_density=min(modes[0], 1), stepmode_color=modes[1])
    origins, vm = oracle_mod.leaves_from_mask(sc["mask"])
    active = oracle_mod.SlotDomain(origins, vm).active().astype(np.uint8)
    rate = None
    if modes[0] == 2:
        # a view-count grid: small integers over the dense resolution
        counts = np.random.default_rng(7).integers(0, 9, size=sc["reso"]).astype(np.float32)
        tr.set_pervoxel_lr(torch.from_numpy(counts))
        rate = tr.density_lr.cpu().numpy().copy()
        assert tr.desc.stepmode_density == 2 and rate.max() == 1.0
    f32 = lambda t: t.detach().float().cpu().numpy().copy()  # noqa: E731
    for step in range(4):
        if step < 3:  # the fourth step runs on the re-zeroed gradients
            tr.forward_backward()
        torch.cuda.synchronize()
        gd, gc = f32(tr.grad_density), f32(tr.grad_color)
        den, md, vd = f32(tr.density)[:, None], f32(tr.m_density)[:, None], f32(tr.v_density)[:, None]
        col, mc, vc = f32(tr.color), f32(tr.m_color), f32(tr.v_color)
        den0, col0 = den.copy(), col.copy()
        sz = tr._stepsz(tr.lr_grid, tr.step_count + 1)
        oracle_mod.vdb_adam(den, np.ascontiguousarray(gd[:, None]), md, vd, active, 1, modes[0], sz,
                            1e-8, 0.9, 0.99, per_lr=rate)
        oracle_mod.vdb_adam(col, gc, mc, vc, active, 3, modes[1], sz, 1e-8, 0.9, 0.99)
        tr.adam_only()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(f32(tr.density), den[:, 0])
        np.testing.assert_array_equal(f32(tr.m_density), md[:, 0])
        np.testing.assert_array_equal(f32(tr.v_density), vd[:, 0])
        np.testing.assert_array_equal(f32(tr.color), col)
        np.testing.assert_array_equal(f32(tr.m_color), mc)
        np.testing.assert_array_equal(f32(tr.v_color), vc)
        assert float(tr.grad_density.abs().max()) == 0 and float(tr.grad_color.abs().max()) == 0
        moved_d = int((den != den0).sum())
        moved_c = int((col != col0).any(1).sum())
        print(f"[result] step {step}: {moved_d} density and {moved_c} colour voxels moved "
              f"({int((gd != 0).sum())} nonzero density gradients)", flush=True)
        if step == 3:
            assert not gd.any() and not gc.any()
            assert (moved_d > 1000) == (modes[0] != 1) and (moved_d == 0) == (modes[0] == 1)
            assert (moved_c > 1000) == (modes[1] != 1) and (moved_c == 0) == (modes[1] == 1)


def test_direct_training_curve_follows_plain_path():
    from plenvdb_amd import scenes
    from plenvdb_amd.fused import FusedTrainer
    from plenvdb_amd.trainer import PlainTrainer

    print("[case] training curve c1/2048 without MLP, 100 steps, fused vs plain", flush=True)
    sc = scenes.make_scene("c1", n_rays=2048)
    fused = FusedTrainer(sc)
    lf = []
    for _ in range(100):
        fused.step()
        lf.append(fused.loss.clone())
    lf = np.array([float(x) for x in torch.stack(lf).cpu()])
    plain = PlainTrainer(sc)
    lp = np.array([float(plain.step()[0].detach()) for _ in range(100)])
    rel = np.abs(lf - lp) / lp
    print("[result] step: plain / fused loss  " + "  ".join(
        f"{i}: {lp[i]:.6f}/{lf[i]:.6f}" for i in range(0, 100, 20)) +
          f"  99: {lp[-1]:.6f}/{lf[-1]:.6f}  max rel gap {rel.max():.2e}", flush=True)
    assert lf[-1] < 0.9 * lf[0]
    assert rel.max() <= 1e-3
