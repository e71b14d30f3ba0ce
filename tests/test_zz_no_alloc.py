"""pvdb_render_image_noalloc: the renderer with its frame workspace in the
library's fixed device scratch instead of a caller-provided buffer.  Called
through the C ABI, its image equals pvdb_render_image's bit for bit (the same
kernels on the same inputs; the composite adds each pixel's contributions in
sample order, so the frame does not depend on where its records land), on two
calls in a row, the second with a smaller frame, so nothing the first call
left in the scratch reaches the second.  A frame whose workspace exceeds the
scratch is refused with PVDB_EINVAL before any device work."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _renderer(ms, H, W, K):
    from plenvdb_amd import vdbio
    from plenvdb_amd.renderer import MGRenderer

    r = MGRenderer(12, 27, 128, 3)
    p = ms["mlp"]
    r.load_params(p["w0"], p["b0"], p["w1"], p["b1"], p["w2"], p["b2"])
    topo, slot_idx = vdbio.index_grid_arrays(ms["idx"])
    r.load_data_grid(ms["den"].astype(np.float32), ms["col"].astype(np.float32).reshape(-1),
                     topo, slot_idx, len(ms["den"]))
    r.setScene(ms["reso"], K, ms["xyz_min"], ms["xyz_max"])
    r.setKwargs(ms["near"], ms["far"], ms["stepdist"], ms["act_shift"], ms["interval"],
                ms["fast_color_thres"], ms["bg"], ms["inverse_y"], H, W)
    r.set_t_stop(ms["t_stop"])
    return r


def _lib_i64():
    """The library with the int64 return types of the two size queries set."""
    from plenvdb_amd import lib

    L = lib.load()
    L.pvdb_render_workspace_bytes.restype = C.c_int64
    L.pvdb_render_workspace_bytes.argtypes = [C.c_int32, C.c_int32, C.c_int64]
    L.pvdb_render_noalloc_scratch_bytes.restype = C.c_int64
    L.pvdb_render_noalloc_scratch_bytes.argtypes = []
    return L


def _noalloc_frame(r):
    """The renderer's last frame again, through pvdb_render_image_noalloc with
    no workspace in the descriptor; returns the image."""
    from plenvdb_amd import lib

    L = _lib_i64()
    d = r.desc
    out = torch.full_like(r.image, float("nan"))
    d.out = out.data_ptr()
    d.workspace = None
    d.workspace_bytes = 0
    d.capacity = r.capacity
    lib.call("pvdb_render_image_noalloc", r.topo.handle, C.byref(d), lib.stream())
    torch.cuda.synchronize()
    assert L.pvdb_render_noalloc_scratch_bytes() == 64 << 20
    return out.cpu().numpy()


def test_noalloc_render_equals_workspace_render():
    from plenvdb_amd import scenes

    L = _lib_i64()
    scratch = L.pvdb_render_noalloc_scratch_bytes()
    ms = scenes.merged_scene("c2", H=96, W=104, focal=1111.0 * 96 / 800, n_poses=100)
    for H, W, pose in ((96, 104, 0), (40, 56, 37)):
        print(f"[case] noalloc render c2 {H}x{W} pose {pose}", flush=True)
        K = ms["K"].copy()
        K[0, 2] -= (104 - W) / 2
        K[1, 2] -= (96 - H) / 2
        r = _renderer(ms, H, W, K)
        r.input_a_c2w(ms["poses"][pose])
        r.render_an_image()  # sizes the record capacity for this pose
        torch.cuda.synchronize()
        ref = r.output_an_image()
        need = L.pvdb_render_workspace_bytes(H, W, r.capacity)
        assert 0 < need <= scratch
        img = _noalloc_frame(r)
        print(f"[result] {H}x{W}: capacity {r.capacity}, workspace {need} B, "
              f"max|noalloc - workspace| {np.abs(img - ref).max():.3g}", flush=True)
        assert np.isfinite(img).all()
        np.testing.assert_array_equal(img, ref)
        assert (ref != ms["bg"]).sum() > 100  # the frame is not background only


def test_noalloc_render_refuses_frames_beyond_the_scratch():
    from plenvdb_amd import lib, scenes

    L = lib.load()
    ms = scenes.merged_scene("c2", H=32, W=32, focal=1111.0 * 32 / 800, n_poses=4)
    r = _renderer(ms, 32, 32, ms["K"])
    r.input_a_c2w(ms["poses"][0])
    d = r.desc
    d.payload, d.weights, d.out = r.payload.data_ptr(), r.weights.data_ptr(), r.image.data_ptr()
    d.payload_f32 = int(r.payload_f32)
    d.workspace, d.workspace_bytes = None, 0
    d.capacity = 2 * 1024 * 1024  # 96 MiB of records
    rc = L.pvdb_render_image_noalloc(r.topo.handle, C.byref(d), None)
    assert rc == -1 and b"64 MiB" in L.pvdb_last_error()
