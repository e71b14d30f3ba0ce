"""The fused training step captured in a hipGraph: every replay matches the
eager step (same scene, same initial state) to 1e-5 relative, with eager torch
work and synchronisations between the replays.  The per-step Adam step sizes
reach the replayed graph through a device buffer."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a = a.detach().double().cpu()
    b = b.detach().double().cpu()
    return float((a - b).norm() / max(float(b.norm()), 1e-30))


def test_graph_replay_matches_eager():
    from plenvdb_amd import fused, scenes

    print("[case] graph replay c2 2048 rays", flush=True)
    sc = scenes.make_scene("c2", n_rays=2048)
    ta = fused.FusedTrainer(sc)
    tb = fused.FusedTrainer(sc)
    ta.step()
    tb.step()
    torch.cuda.synchronize()
    tb.capture_graph()
    for k in range(4):
        la = float(ta.step())
        lb = float(tb.step())  # graph replay
        torch.cuda.synchronize()
        # unrelated eager work between replays
        x = torch.randn(1 << 16, device="cuda")
        assert torch.isfinite(x.sum())
        torch.cuda.synchronize()
        assert abs(la - lb) <= 1e-5 * abs(la), (k, la, lb)
        assert _rel(tb.density, ta.density) < 1e-5
        assert _rel(tb.color.float(), ta.color.float()) < 1e-5
        assert _rel(tb.mlp, ta.mlp) < 1e-5
    assert tb.step_count == ta.step_count == 5
