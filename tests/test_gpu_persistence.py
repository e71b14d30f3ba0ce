"""Checkpoints: save_to / load_from of the grids and their optimizers
(plenvdb.h:126-148, 211-276, 720-736), VDBAdam's vdbopt.tar
(masked_adam.py:70-100), pruneGrid and evalActiveVoxelDim on load, and a
resumed training run (run.py:404-424, 611-626) against an uninterrupted one.
Values are compared bit for bit where nothing is recomputed; a resumed run
differs from the uninterrupted one only by the order of the gradient
scatter's float atomics, so its losses are compared to 1e-5 relative."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _c1_grid(cls, n, seed):
    from plenvdb_amd import scenes
    from plenvdb_amd.plenvdb import ColorVDB, DensityVDB
    from plenvdb_amd.topology import Topology

    mask = scenes.occupancy_mask(scenes.CONFIGS["c1"]).copy()
    mask[100:] = False  # an active extent below the resolution on two axes
    mask[:, :, :20] = False
    topo = Topology.from_mask(mask)
    g = (DensityVDB if cls == "density" else ColorVDB)([128] * 3, n, topology=topo)
    rng = np.random.default_rng(seed)
    vals = rng.normal(size=(topo.nslots, n)).astype(np.float32)
    g.grid.copy_(torch.from_numpy(vals.reshape(g.grid.shape)))
    return g, mask


@pytest.mark.parametrize("cls,n", [("density", 1), ("color", 12)])
def test_vdb_save_load_round_trip(tmp_path, cls, n):
    from plenvdb_amd.plenvdb import ColorVDB, DensityVDB

    print("[case] save/load", cls, flush=True)
    g, mask = _c1_grid(cls, n, 3)
    path = str(tmp_path / f"{cls}.vdb")
    g.save_to(path)
    h = (DensityVDB if cls == "density" else ColorVDB)([128] * 3, n)
    h.load_from(path)
    o1, m1 = g.topo.leaves()
    o2, m2 = h.topo.leaves()
    np.testing.assert_array_equal(o1, o2)
    np.testing.assert_array_equal(m1, m2)
    np.testing.assert_array_equal(h.grid.cpu().numpy(), g.grid.cpu().numpy())
    # reso <- evalActiveVoxelDim: the extent of the active voxels
    act = np.argwhere(mask)
    assert h.reso == [int(v) for v in act.max(0) - act.min(0) + 1]
    assert h.reso != [128] * 3


def test_prune_on_load_drops_background_leaves(tmp_path):
    """Leaves with no active voxel whose values are all 0 become background
    under pruneGrid and are dropped on load; a leaf with no active voxel but
    other values, and a fully active uniform leaf, keep their values (the
    latter would be an active tile in OpenVDB).  Lookups are unchanged."""
    from plenvdb_amd.plenvdb import DensityVDB
    from plenvdb_amd.topology import Topology

    print("[case] prune on load", flush=True)
    rng = np.random.default_rng(4)
    o = np.array([[0, 0, 0], [0, 0, 8], [0, 8, 0], [8, 0, 0], [8, 8, 8], [16, 0, 0]], np.int32)
    m = np.zeros((6, 8), np.uint64)
    m[0] = rng.integers(0, 2 ** 63, 8, dtype=np.uint64)  # partly active
    m[4] = np.uint64(2 ** 64 - 1)                       # fully active
    topo = Topology.from_leaves(o, m, [24, 16, 16])
    g = DensityVDB([24, 16, 16], 1, topology=topo)
    vals = np.zeros((6, 512), np.float32)
    vals[0] = rng.normal(size=512)
    vals[2, 7] = 0.5      # inactive leaf with a non-zero value: kept
    vals[4] = 1.25        # active uniform leaf: kept
    # leaves 1, 3, 5: inactive, all zero -> background
    order, _ = topo.leaves()
    perm = [int(np.nonzero((o == x).all(1))[0][0]) for x in order]
    g.grid.copy_(torch.from_numpy(vals[perm]).reshape(g.grid.shape))
    path = str(tmp_path / "density.vdb")
    g.save_to(path)
    h = DensityVDB([24, 16, 16], 1)
    h.load_from(path)
    assert h.topo.nleaf == 3
    kept, _ = h.topo.leaves()
    assert {tuple(x) for x in kept.tolist()} == {(0, 0, 0), (0, 8, 0), (8, 8, 8)}
    h.setReso(24, 16, 16)  # load_from set it to the active extent
    np.testing.assert_array_equal(h.get_dense_grid(), g.get_dense_grid())
    p = rng.uniform(0, 15, (4000, 3)).astype(np.float32)
    a = g.forward(p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy())
    b = h.forward(p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy())
    np.testing.assert_array_equal(a, b)


def test_optimizer_moments_follow_a_pruned_parameter(tmp_path):
    """DensityOpt's moments saved on the unpruned leaf set land on the same
    voxels of the pruned parameter (the reference pairs them by leaf index,
    which shifts once pruneGrid removed leaves)."""
    from plenvdb_amd.plenvdb import DensityOpt, DensityVDB
    from plenvdb_amd.topology import Topology

    print("[case] moments follow pruned parameter", flush=True)
    o = np.array([[0, 0, 0], [0, 0, 8], [8, 0, 0]], np.int32)
    m = np.zeros((3, 8), np.uint64)
    m[0] = m[2] = np.uint64(2 ** 64 - 1)  # leaf (0,0,8) is background -> pruned
    topo = Topology.from_leaves(o, m, [16, 8, 16])
    g = DensityVDB([16, 8, 16], 1, topology=topo)
    rng = np.random.default_rng(5)
    order, _ = topo.leaves()
    bg = int(np.nonzero((order == [0, 0, 8]).all(1))[0][0])
    v = rng.normal(size=(3, 512)).astype(np.float32)
    v[bg] = 0
    g.grid.copy_(torch.from_numpy(v).reshape(g.grid.shape))
    opt = DensityOpt(g, 0.1, 1e-8, 0.9, 0.99)
    mom = rng.normal(size=(3, 512)).astype(np.float32) ** 2
    opt.exp_avg.copy_(torch.from_numpy(mom).reshape(opt.exp_avg.shape))
    opt.exp_avg_sq.copy_(torch.from_numpy(2 * mom).reshape(opt.exp_avg_sq.shape))
    g.save_to(str(tmp_path / "density.vdb"))
    opt.save_to(str(tmp_path / "density_"))
    h = DensityVDB([16, 8, 16], 1)
    h.load_from(str(tmp_path / "density.vdb"))
    assert h.topo.nleaf == 2
    hopt = DensityOpt(h, 0.1, 1e-8, 0.9, 0.99)
    hopt.load_from(str(tmp_path / "density_"))
    ho, _ = h.topo.leaves()
    got = hopt.exp_avg.cpu().numpy().reshape(2, 512)
    for j, x in enumerate(ho.tolist()):
        i = int(np.nonzero((order == x).all(1))[0][0])
        np.testing.assert_array_equal(got[j], mom[i])
    np.testing.assert_array_equal(hopt.exp_avg_sq.cpu().numpy().reshape(-1), 2 * got.reshape(-1))


def test_vdbadam_save_load(tmp_path):
    from plenvdb_amd import scenes
    from plenvdb_amd.masked_adam import VDBAdam
    from plenvdb_amd.trainer import PlainTrainer

    print("[case] VDBAdam save/load", flush=True)
    sc = scenes.make_scene("c2", n_rays=256)
    tr = PlainTrainer(sc)
    for _ in range(3):
        tr.step()
    prefix = str(tmp_path) + "/"
    tr.model.save_to(prefix)
    tr.vdbopt.save_to(prefix)
    st = torch.load(prefix + "vdbopt.tar", weights_only=True)
    assert set(st) == {f"{k}{n}" for k in "dc" for n in ("lr", "eps", "beta0", "beta1", "step")}
    assert st["dstep"] == 3 and st["cstep"] == 3
    tr2 = PlainTrainer(sc)
    tr2.model.load_from(prefix)
    opt2 = VDBAdam([{"params": tr2.model.density, "lr": 1.0, "skip_zero_grad": True},
                    {"params": tr2.model.k0, "lr": 1.0, "skip_zero_grad": True}])
    opt2.load_from(prefix)
    for a, b in ((tr.vdbopt.densityOpt, opt2.densityOpt), (tr.vdbopt.colorOpt, opt2.colorOpt)):
        assert (a.getLr(), a.getEps(), a.getBeta0(), a.getBeta1(), a.getStep()) == \
            (b.getLr(), b.getEps(), b.getBeta0(), b.getBeta1(), b.getStep())
        np.testing.assert_array_equal(a.exp_avg.cpu().numpy(), b.exp_avg.cpu().numpy())
        np.testing.assert_array_equal(a.exp_avg_sq.cpu().numpy(), b.exp_avg_sq.cpu().numpy())
    np.testing.assert_array_equal(tr.model.k0.grid.grid.cpu().numpy(),
                                  tr2.model.k0.grid.grid.cpu().numpy())


def test_resume_continues_the_loss_trajectory(tmp_path):
    from plenvdb_amd import scenes
    from plenvdb_amd.trainer import PlainTrainer

    print("[case] resume plain c2/512", flush=True)
    sc = scenes.make_scene("c2", n_rays=512)
    a = PlainTrainer(sc)
    for _ in range(4):
        a.step()
    prefix = str(tmp_path) + "/fine_"
    a.save_checkpoint(prefix, 4)
    la = [float(a.step()[0].detach()) for _ in range(4)]
    b = PlainTrainer(sc)
    assert b.load_checkpoint(prefix) == 4
    lb = [float(b.step()[0].detach()) for _ in range(4)]
    print(f"[result] uninterrupted {la}  resumed {lb}", flush=True)
    np.testing.assert_allclose(lb, la, rtol=1e-5)


def test_fused_checkpoint_interchanges_with_the_plain_path(tmp_path):
    """The fused trainer writes the plain path's file set: a plain trainer
    loading it renders the same loss (within the fused step's bf16 MLP
    weights), and a fused trainer loading it continues the fused run."""
    from plenvdb_amd import scenes
    from plenvdb_amd.fused import FusedTrainer
    from plenvdb_amd.trainer import PlainTrainer

    print("[case] fused checkpoint -> plain and fused", flush=True)
    sc = scenes.make_scene("c2", n_rays=512)
    f = FusedTrainer(sc)
    for _ in range(5):
        f.step()
    prefix = str(tmp_path) + "/fine_"
    f.save_to(prefix)
    nxt = [float(f.step().item()) for _ in range(3)]
    g = FusedTrainer(sc)
    assert g.load_from(prefix) == 5
    got = [float(g.step().item()) for _ in range(3)]
    np.testing.assert_allclose(got, nxt, rtol=1e-5)
    p = PlainTrainer(sc)
    assert p.load_checkpoint(prefix) == 5
    loss_plain = float(p.forward_backward()[0].detach())
    h = FusedTrainer(sc)
    h.load_from(prefix)
    h.forward_backward()
    loss_fused = float(h.loss.item())
    print(f"[result] fused next {nxt} resumed {got}; loss at the checkpoint: plain "
          f"{loss_plain:.6f} fused {loss_fused:.6f}", flush=True)
    assert abs(loss_plain - loss_fused) <= 1e-3 * loss_plain
