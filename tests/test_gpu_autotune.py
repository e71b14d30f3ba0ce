"""The launch-shape autotuner: its choices come from the candidate sets, and
the step computes the same values under every launch shape (the loss bit for
bit; the grid and MLP gradients up to the order of float atomics)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _run(tr):
    tr.zero_grad()
    tr.forward_backward()
    torch.cuda.synchronize()
    return (float(tr.loss.item()), tr.grad_density.clone(), tr.grad_color.clone(),
            tr.mlp_grad.clone())


def test_autotune_choices_and_invariance():
    from plenvdb_amd import autotune, scenes
    from plenvdb_amd.fused import FusedTrainer

    print("[case] autotune c2 2048 rays", flush=True)
    tr = FusedTrainer(scenes.make_scene("c2", n_rays=2048))
    base = _run(tr)
    out = tr.autotune(reps=1, force=True)
    print("[result]", out["launch"], flush=True)
    for i, name in autotune.NAMES.items():
        assert out["launch"][name] in autotune.CANDIDATES[i]
    assert list(tr.desc.launch) == [out["launch"][autotune.NAMES[i]] for i in range(4)]
    # every candidate of each choice gives the same values
    for idx, cands in autotune.CANDIDATES.items():
        for v in cands:
            saved = list(tr.desc.launch)
            tr.desc.launch[idx] = v
            loss, gd, gc, gm = _run(tr)
            tr.desc.launch[idx] = saved[idx]
            assert loss == base[0]
            for a, b in ((gd, base[1]), (gc, base[2]), (gm, base[3])):
                d = float((a - b).abs().max())
                assert d <= 1e-6 * max(float(b.abs().max()), 1e-30), (idx, v, d)
    # a second call reuses the choice
    again = tr.autotune()
    assert again["source"] == "cache" and again["launch"] == out["launch"]


def test_launch_shapes_are_checked():
    from plenvdb_amd import scenes
    from plenvdb_amd.fused import FusedTrainer

    print("[case] invalid launch shapes", flush=True)
    tr = FusedTrainer(scenes.make_scene("c2", n_rays=256))
    tr.desc.launch[2] = 48
    with pytest.raises(RuntimeError, match="32 or 64"):
        tr.forward_backward()
    tr.desc.launch[2] = 0
    tr.desc.launch[0] = -1
    with pytest.raises(RuntimeError, match="launch"):
        tr.forward_backward()
    tr.desc.launch[0] = 0
    tr.desc.launch[3] = 3
    with pytest.raises(RuntimeError, match="rays per workgroup"):
        tr.forward_backward()
    tr.desc.launch[3] = 0
    tr.forward_backward()
    torch.cuda.synchronize()
    assert np.isfinite(float(tr.loss.item()))
