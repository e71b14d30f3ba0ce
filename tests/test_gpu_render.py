"""Merged-VDB renderer (MGRenderer) against the oracle's restatement of the
reference renderer (two marches, renderer.cu:222-367, then the MLP and the
composite in numpy).  Tolerance: rendered RGB within 1e-3 max-abs of the FP64
oracle.  The sample set is integer work: per pixel, the first march's kept
count and the number of samples the second march writes equal the FP32
oracle's exactly (both marches run the reference's float arithmetic,
including the second march's restart from the first kept t minus one step).
The oracle's own marcher is pinned by tests/test_oracle_math.py and
tests/test_oracle_pins.py."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _renderer(ms, H, W, K, t_stop):
    from plenvdb_amd import vdbio
    from plenvdb_amd.renderer import MGRenderer

    r = MGRenderer(12, 27, 128, 3)
    p = ms["mlp"]
    r.load_params(p["w0"], p["b0"], p["w1"], p["b1"], p["w2"], p["b2"])
    topo, slot_idx = vdbio.index_grid_arrays(ms["idx"])
    den = ms["den"].astype(np.float32)
    col = ms["col"].astype(np.float32)
    r.load_data_grid(den, col.reshape(-1), topo, slot_idx, len(den))
    r.setScene(ms["reso"], K, ms["xyz_min"], ms["xyz_max"])
    r.setKwargs(ms["near"], ms["far"], ms["stepdist"], ms["act_shift"], ms["interval"],
                ms["fast_color_thres"], ms["bg"], ms["inverse_y"], H, W)
    r.set_t_stop(t_stop)
    return r


def _oracle(oracle_mod, ms, c2w, K, H, W, t_stop, prec, inverse_y=False):
    return oracle_mod.reference_render(
        ms["idx"], ms["den"].astype(np.float32), ms["col"].astype(np.float32), ms["mlp"], c2w, K,
        ms["xyz_min"], ms["xyz_max"], H, W, ms["near"], ms["stepdist"], ms["act_shift"],
        ms["interval"], ms["fast_color_thres"], ms["bg"], inverse_y=inverse_y, t_stop=t_stop,
        prec=prec)


@pytest.mark.parametrize("cfg,H,pose,t_stop,inverse_y", [
    ("c2", 72, 0, 1e-3, False),
    ("c2", 72, 37, 1e-4, False),
    ("c1", 64, 80, 1e-3, True),
])
def test_render_matches_oracle(oracle_mod, cfg, H, pose, t_stop, inverse_y):
    from plenvdb_amd import scenes

    W = H + 8  # not a multiple of the 8x8 tile in both directions
    ms = scenes.merged_scene(cfg, H=H, W=W, focal=1111.0 * H / 800, n_poses=100)
    ms["inverse_y"] = inverse_y
    c2w = ms["poses"][pose]
    if inverse_y:  # camera looking down +z (OpenCV convention): flip the y and z axes
        c2w = (c2w @ np.diag([1.0, -1.0, -1.0, 1.0])).astype(np.float32)
    K = ms["K"]
    r = _renderer(ms, H, W, K, t_stop)
    r.input_a_c2w(c2w)
    _check_frame(oracle_mod, r, ms, c2w, K, H, W, t_stop, inverse_y, f"{cfg} pose {pose}")


def _check_frame(oracle_mod, r, ms, c2w, K, H, W, t_stop, inverse_y, label, background=True):
    print(f"[case] render {label} {H}x{W} t_stop {t_stop}", flush=True)
    r.render_an_image()
    img = r.output_an_image()
    gathers, kept = r.sample_counts()
    count, written = r.pixel_counts()
    ref64, info = _oracle(oracle_mod, ms, c2w, K, H, W, t_stop, "f64", inverse_y)
    ref32, info32 = _oracle(oracle_mod, ms, c2w, K, H, W, t_stop, "f32", inverse_y)
    assert img.shape == (H * W * 3,)
    err64 = np.abs(img - ref64).max()
    err32 = np.abs(img - ref32).max()
    n_ref = int(info32["written"].sum())
    bad = np.nonzero((count != info32["count"]) | (written != info32["written"]))[0]
    print(f"[result] max|gpu-f64| {err64:.3g}  max|gpu-f32| {err32:.3g}  gathers {gathers} "
          f"written {kept} oracle written {n_ref}  pixels with other counts {len(bad)} "
          f"{bad[:8].tolist()}", flush=True)
    assert n_ref > 100
    np.testing.assert_array_equal(count, info32["count"])
    np.testing.assert_array_equal(written, info32["written"])
    assert kept == n_ref
    assert err64 <= 1e-3
    # the picture is not trivial: some pixels are covered and (unless the window
    # lies inside the scene's silhouette) some are background
    cover = 1.0 - info["t_last"]
    assert (cover > 0.5).sum() > 10
    if background:
        assert (cover < 1e-6).sum() > 10
    else:
        assert cover.max() - cover.min() > 0.1


@pytest.mark.parametrize("pose", [0, 61])
def test_render_c5_window_matches_oracle(oracle_mod, pose):
    """The benchmarked render workload: c4's 320^3 topology as a merged index
    grid with the fp16 payload, the 800x800 focal-1111 camera and T < 1e-4, on
    a 96 x 104 window at the image centre (its principal point shifted by the
    window's offset)."""
    from plenvdb_amd import scenes

    H, W = 96, 104
    ms = scenes.merged_scene("c4", H=800, W=800, focal=1111.0, n_poses=100)
    K = ms["K"].copy()
    K[0, 2] -= (800 - W) / 2
    K[1, 2] -= (800 - H) / 2
    c2w = ms["poses"][pose]
    r = _renderer(ms, H, W, K, ms["t_stop"])
    assert not r.payload_f32
    r.input_a_c2w(c2w)
    _check_frame(oracle_mod, r, ms, c2w, K, H, W, ms["t_stop"], False, f"c5 window pose {pose}",
                 background=False)


def test_render_fp32_payload_and_repeat(oracle_mod):
    """Data that fp16 cannot hold exactly goes to the fp32 payload; a second
    render of another pose reuses the renderer."""
    from plenvdb_amd import scenes

    H = W = 48
    ms = scenes.merged_scene("c2", H=H, W=W, focal=1111.0 * H / 800, n_poses=100)
    rng = np.random.default_rng(7)
    den = ms["den"].astype(np.float32) + rng.normal(0, 1e-3, ms["den"].shape).astype(np.float32)
    den[0] = 0
    ms["den"] = den
    r = _renderer(ms, H, W, ms["K"], 1e-3)
    assert r.payload_f32
    for pose in (5, 55):
        c2w = ms["poses"][pose]
        r.input_a_c2w(c2w)
        r.render_an_image()
        img = r.output_an_image()
        ref, _ = oracle_mod.reference_render(
            ms["idx"], den, ms["col"].astype(np.float32), ms["mlp"], c2w, ms["K"], ms["xyz_min"],
            ms["xyz_max"], H, W, ms["near"], ms["stepdist"], ms["act_shift"], ms["interval"],
            ms["fast_color_thres"], ms["bg"], t_stop=1e-3, prec="f64")
        assert np.abs(img - ref).max() <= 1e-3


def test_render_needs_all_setters():
    from plenvdb_amd.renderer import MGRenderer

    r = MGRenderer(12, 27, 128, 3)
    r.render_an_image()  # silently does nothing, like the reference
    assert r.output_an_image() is None
    with pytest.raises(RuntimeError):
        r.render_async()
    with pytest.raises(ValueError):
        MGRenderer(16, 27, 128, 3)


def test_merged_files_round_trip(tmp_path):
    from plenvdb_amd import merged, scenes, vdbio
    from plenvdb_amd.renderer import MGRenderer

    ms = scenes.merged_scene("c1", H=32, W=32, focal=1111.0 * 32 / 800, n_poses=4)
    mask = ms["mask"]
    rng = np.random.default_rng(3)
    den_d = rng.normal(0, 2, mask.shape).astype(np.float32)
    col_d = rng.normal(0, 0.1, mask.shape + (12,)).astype(np.float32)
    idx, den, col = merged.export_merged(mask, den_d, col_d)
    assert idx.max() == mask.sum() and den.dtype == np.float16 and den[0] == 0
    assert np.array_equal(den[idx[mask]], den_d[mask].astype(np.float16))
    merged.save_merged(str(tmp_path), idx, den, col)
    d32, c32, N = merged.load_merged_data(str(tmp_path))
    topo, slot_idx, reso = vdbio.load_index_grid(str(tmp_path / "mergedidxs.vdb"))
    assert reso == list(mask.shape)
    s, act = topo.slots(np.argwhere(mask))
    assert act.all()
    assert np.array_equal(slot_idx.cpu().numpy()[s], idx[mask])
    r = MGRenderer(12, 27, 128, 3)
    r.load_data(d32, c32, str(tmp_path / "mergedidxs.vdb"), N)
    assert r.flags[0] and not r.payload_f32
