"""The C-ABI library loads and exports every symbol include/*.h declares; host
argument checks fail with codes and messages (no device work)."""
import ctypes as C
import glob
import os
import re

import pytest

from conftest import ROOT


def header_symbols():
    syms = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        syms |= set(re.findall(r"\b(pvdb_[a-z0-9_]+)\s*\(", text))
    return syms


def test_library_exports_all_declared_symbols():
    from plenvdb_amd import lib

    L = lib.load()
    syms = header_symbols()
    assert len(syms) >= 20
    missing = [s for s in sorted(syms) if not hasattr(L, s)]
    assert not missing, missing
    # the Python binding declares exactly the header's functions
    assert set(lib.declared_symbols()) <= syms


def test_argument_errors_are_reported():
    from plenvdb_amd import lib

    L = lib.load()
    h = C.c_void_p()
    rc = L.pvdb_topo_create_from_mask(None, 0, 4, 4, C.byref(h))
    assert rc == -1
    assert b"resolution" in L.pvdb_last_error()
    rc = L.pvdb_trilinear_fwd(None, None, 1, 0, None, None, None, 0, None, None)
    assert rc == -3 and b"topology" in L.pvdb_last_error()
    with pytest.raises(RuntimeError, match="mode must be"):
        lib.call("pvdb_dense_adam", None, None, None, None, None, 10, 7, 0.1, 1e-8, 0.9, 0.99, None)
    with pytest.raises(RuntimeError, match="null topology"):
        lib.call("pvdb_vdb_adam", None, None, None, None, None, None, 12, 3, 0, 0.1, 1e-8, 0.9,
                 0.99, None)
    assert L.pvdb_abi_version() == 1


def test_new_entry_points_check_arguments():
    """Host argument checks of the camera-ray, FP64 adjoint and fused-step entry
    points (all before any device work)."""
    import numpy as np

    from plenvdb_amd import lib

    K = np.array([[50.0, 0, 4], [0, 50.0, 4], [0, 0, 1]])
    c2w = np.eye(4, dtype=np.float32)[:3].copy()
    kp, cp = K.ctypes.data_as(C.c_void_p), c2w.ctypes.data_as(C.c_void_p)
    with pytest.raises(RuntimeError, match="mode must be"):
        lib.call("pvdb_rays_of_view", 4, 4, kp, cp, 0, 0, 0, 0, 3, None, None, None, None, None)
    with pytest.raises(RuntimeError, match="jitter"):
        lib.call("pvdb_rays_of_view", 4, 4, kp, cp, 0, 0, 0, 0, 2, 1, 1, 1, None, None)
    with pytest.raises(RuntimeError, match="null pointer"):
        lib.call("pvdb_rays_of_view", 4, 4, kp, cp, 0, 0, 0, 0, 0, None, None, None, None, None)
    Kz = np.zeros((3, 3))
    with pytest.raises(RuntimeError, match="focal"):
        lib.call("pvdb_rays_of_view", 4, 4, Kz.ctypes.data_as(C.c_void_p), cp, 0, 0, 0, 0, 0,
                 1, 1, 1, None, None)
    lib.call("pvdb_rays_of_view", 0, 4, kp, cp, 0, 0, 0, 0, 0, None, None, None, None, None)
    with pytest.raises(RuntimeError, match="null topology"):
        lib.call("pvdb_trilinear_fwd_f64", None, None, 1, None, None, None, 4, None, None)
    with pytest.raises(RuntimeError, match="n must be"):
        lib.call("pvdb_raw2alpha_f64", None, -1, 0.0, 0.5, None, None, None)
    with pytest.raises(RuntimeError, match="null pointer"):
        lib.call("pvdb_alpha2weight_bwd_f64", None, None, None, None, None, None, 3, None, None,
                 None, None)
    with pytest.raises(RuntimeError, match="null topology"):
        lib.call("pvdb_tv_add_grad", None, None, None, 1, 4, 4, 4, 1.0, 1.0, 1.0, 1, None)
    from plenvdb_amd.fused import StepDesc

    d = StepDesc()
    with pytest.raises(RuntimeError, match="null topology"):
        lib.call("pvdb_fused_step", None, C.byref(d), None)
    assert lib.load().pvdb_fused_workspace_bytes(-1, 1, 1) == -1


def test_fused_step_rejects_capacity_and_misaligned_buffers():
    """pvdb_fused_step checks the record capacity (offsets are int32, so at most
    2^31 - 2) and the alignment of the buffers it moves as 16-byte vectors
    before it reads the topology or launches anything.  The topology argument
    is a zero-filled stand-in buffer: none of these checks reads it."""
    from plenvdb_amd import lib
    from plenvdb_amd.fused import StepDesc

    fake_topo = C.create_string_buffer(4096)
    base = 1 << 20  # a 256-byte aligned fake device address, never dereferenced

    def desc(**kw):
        d = StepDesc()
        for i, f in enumerate(("rays_o", "rays_d", "viewdirs", "target", "density",
                               "grad_density", "color", "grad_color", "mlp", "mlp_grad",
                               "density_m", "density_v", "color_m", "color_v", "mlp_m",
                               "mlp_v")):
            setattr(d, f, base + 4096 * (i + 1))
        d.n_rays, d.capacity, d.stepdist = 16, 1024, 0.01
        d.workspace, d.workspace_bytes = base, 1 << 40
        d.weight_main = 1.0
        for k, v in kw.items():
            setattr(d, k, v)
        return d

    cases = [
        (dict(capacity=2 ** 31 - 1), "capacity"),
        (dict(capacity=2 ** 40), "capacity"),
        (dict(capacity=0), "capacity"),
        (dict(grad_color=base + 4100), "16-byte"),
        (dict(color_m=base + 8), "16-byte"),
        (dict(color=base + 8), "16-byte"),
        (dict(color=base + 4, color_bf16=1), "8-byte"),
        (dict(workspace=base + 64), "256-byte"),
        (dict(ray_out=base + 8), "ray_out"),
        (dict(ray_out=base + 2, k0_dim=3), "ray_out"),
        (dict(k0_dim=5), "k0_dim"),
        (dict(k0_dim=3, color_bf16=1), "fp32 colour grid"),
        (dict(mlp=None), "null MLP"),
        (dict(stepmode_density=3), "stepmode"),
        (dict(stepmode_color=2), "stepmode"),
        (dict(stepmode_density=2), "density_lr"),
        (dict(grad_stride=8), "grad_stride"),
        (dict(grad_stride=14), "grad_stride"),
        (dict(k0_dim=3, grad_stride=2), "grad_stride"),
    ]
    for kw, msg in cases:
        with pytest.raises(RuntimeError, match=msg):
            lib.call("pvdb_fused_step", C.addressof(fake_topo), C.byref(desc(**kw)), None)
    with pytest.raises(RuntimeError, match="256-byte"):
        lib.call("pvdb_fused_init", C.addressof(fake_topo), C.byref(desc(workspace=base + 16)),
                 None)
