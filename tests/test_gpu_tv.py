"""GPU parity of the VDB total-variation gradient (pvdb_tv_add_grad, through the
C ABI) against the oracle's restatement of the dense kernel on the grid's
dense view."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _case(topo, reso, C, dense, seed, oracle_mod):
    from plenvdb_amd.plenvdb import ColorVDB, DensityVDB

    print("[case] tv_add_grad reso", reso, "C", C, "dense", dense, flush=True)
    o, m = topo.leaves()
    bits = np.unpackbits(m.view(np.uint8).reshape(-1, 64), axis=1, bitorder="little").astype(bool)
    dom = oracle_mod.SlotDomain(o, bits)
    rng = np.random.default_rng(seed)
    vals = (rng.normal(size=(topo.nslots, C)) * 0.6).astype(np.float32)  # some differences clamp
    grad = (rng.normal(size=(topo.nslots, C)) * (rng.random((topo.nslots, C)) < 0.5)).astype(np.float32)
    vdb = DensityVDB(list(reso), 1, topology=topo) if C == 1 else ColorVDB(list(reso), C, topology=topo)
    vdb.grid.copy_(torch.from_numpy(vals))
    vdb.grad.copy_(torch.from_numpy(grad))
    w = (0.4, 0.9, 1.7)
    vdb.tv_add_grad(*w, dense)
    torch.cuda.synchronize()
    got = vdb.grad.cpu().numpy()
    ref = oracle_mod.tv_add_grad_vdb(dom, vals, grad, reso, *w, dense)
    # same FP32 operations in the same order (no contraction): bit-exact
    # (-0.0 == 0.0 holds, so a zero sum's sign does not matter)
    np.testing.assert_array_equal(got, ref)
    changed = got != grad
    assert changed.any()
    if not dense:
        assert not changed[grad == 0].any()


@pytest.mark.parametrize("C,dense", [(1, True), (1, False), (12, True), (12, False)])
def test_tv_add_grad_vs_oracle_c1(C, dense, oracle_mod):
    from plenvdb_amd import scenes
    from plenvdb_amd.topology import Topology

    topo = Topology.from_mask(scenes.occupancy_mask(scenes.CONFIGS["c1"]))
    _case(topo, (128, 128, 128), C, dense, 11 + C, oracle_mod)


def test_tv_add_grad_two_level_layout(oracle_mod):
    """c1's leaves behind the two-level leaf tables: the leaf-face lookups go
    through the lower-node table."""
    from plenvdb_amd import scenes
    from plenvdb_amd.topology import Topology

    one = Topology.from_mask(scenes.occupancy_mask(scenes.CONFIGS["c1"]))
    o, m = one.leaves()
    topo = Topology.from_leaves(o, m, [128, 128, 128], layout=2)
    assert topo.two_level
    _case(topo, (128, 128, 128), 12, False, 21, oracle_mod)


def test_tv_add_grad_box_not_multiple_of_8(oracle_mod):
    """Leaves reach past the box: their outside voxels keep their gradient and
    the box faces contribute no terms."""
    from plenvdb_amd.topology import Topology

    R = (45, 37, 29)
    g = np.stack(np.meshgrid(*[np.arange(r) for r in R], indexing="ij"), -1).astype(np.float64)
    mask = (np.linalg.norm(g - np.array([20.0, 18.0, 14.0]), axis=-1) < 15) | (g[..., 0] > 40)
    topo = Topology.from_mask(mask)
    _case(topo, R, 3, True, 5, oracle_mod)
