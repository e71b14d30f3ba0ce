"""GPU parity of the reference-structured FP32 training step against the
oracle's composed reference step (FP32 reference order and FP64)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _oracle_scene(oracle_mod, sc, den, col):
    origins, vm = oracle_mod.leaves_from_mask(sc["mask"])
    out = dict(sc)
    out.update(origins=origins, vmask=vm, density=den, color=col)
    return out


@pytest.mark.parametrize("name,n_rays,weights", [
    ("c1", 512, (1.0, 0.0, 0.0)), ("c2", 512, (1.0, 0.0, 0.0)), ("c3", 384, (1.0, 0.0, 0.0)),
    ("c2", 384, (1.0, 0.001, 0.01)), ("c1", 384, (1.0, 0.01, 0.1)), ("c1", 256, (0.5, 0.3, 0.5))])
def test_plain_step_matches_oracle(oracle_mod, name, n_rays, weights):
    """weights: (main, entropy_last, rgbper) loss weights of run.py:551-574: MSE
    only, the fine stage's (configs/default.py:73-74), the coarse stage's
    (:52-55, no MLP as in c1) and a stress setting."""
    from plenvdb_amd import scenes
    from plenvdb_amd.trainer import PlainTrainer, init_values

    print("[case] plain step", name, n_rays, "loss weights", weights, flush=True)
    sc = scenes.make_scene(name, n_rays=n_rays)
    tr = PlainTrainer(sc, weight_main=weights[0], weight_entropy_last=weights[1],
                      weight_rgbper=weights[2])
    den, col = init_values(sc, tr.topo)
    loss, out = tr.forward_backward()
    torch.cuda.synchronize()
    osc = _oracle_scene(oracle_mod, sc, den, col)
    osc.update(weight_main=weights[0], weight_entropy_last=weights[1], weight_rgbper=weights[2])
    ref = oracle_mod.reference_train_step(osc, prec="f32", with_mlp=sc["cfg"].use_mlp)
    ref64 = oracle_mod.reference_train_step(
        dict(osc, density=den.astype(np.float64), color=col.astype(np.float64)), prec="f64",
        with_mlp=sc["cfg"].use_mlp)
    # identical sample sets (integer work: bit-exact)
    assert out["n_samples"] == ref["n_mask"]
    np.testing.assert_array_equal(out["ray_id"].cpu().numpy(), ref["ray_id"])
    # rendered colour and loss within 1e-5 relative of the FP32 reference order
    rgb = out["rgb_marched"].detach().cpu().numpy()
    np.testing.assert_allclose(rgb, ref["rgb_marched"], rtol=1e-5, atol=1e-6)
    assert abs(float(loss) - ref["loss"]) <= 1e-5 * abs(ref["loss"])
    np.testing.assert_allclose(rgb, ref64["rgb_marched"], rtol=0, atol=1e-4)
    # per-voxel gradients (scatter order differs: compare to the FP64 sum)
    gd = tr.model.density.grid.grad.cpu().numpy()[:, 0]
    gc = tr.model.k0.grid.grad.cpu().numpy()
    for got, r in ((gd, ref64["grad_density"]), (gc, ref64["grad_color"])):
        scale = np.abs(r).max()
        np.testing.assert_allclose(got, r, rtol=1e-3, atol=1e-4 * scale)
    if sc["cfg"].use_mlp:
        lin = [m for m in tr.model.rgbnet.modules() if isinstance(m, torch.nn.Linear)]
        for i, l in enumerate(lin):
            gw = l.weight.grad.t().cpu().numpy()
            r = ref64["grad_mlp"][f"w{i}"]
            np.testing.assert_allclose(gw, r, rtol=1e-3, atol=1e-4 * np.abs(r).max())


def test_plain_step_runs_optimizer(oracle_mod):
    from plenvdb_amd import scenes
    from plenvdb_amd.trainer import PlainTrainer

    print("[case] plain optimizer c2", flush=True)
    sc = scenes.make_scene("c2", n_rays=256)
    tr = PlainTrainer(sc)
    before = tr.model.density.grid.grid.clone()
    l0, _ = tr.step()
    l1, _ = tr.step()
    torch.cuda.synchronize()
    assert torch.isfinite(l0) and torch.isfinite(l1)
    changed = (tr.model.density.grid.grid != before).sum().item()
    assert changed > 0
    assert tr.vdbopt.densityOpt.getStep() == 2
