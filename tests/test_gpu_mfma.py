"""Device check of the MFMA fragment layouts the MLP kernels rely on."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def bf16_bits(x):
    return (torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(torch.bfloat16)
            .view(torch.int16).numpy().view(np.uint16))


def test_mfma_layouts():
    from plenvdb_amd import lib

    print("[case] mfma layout", flush=True)
    rng = np.random.default_rng(0)
    A = rng.integers(-8, 8, size=(16, 32)).astype(np.float32)
    B = rng.integers(-8, 8, size=(32, 16)).astype(np.float32)
    A16 = rng.integers(-8, 8, size=(16, 16)).astype(np.float32)
    B16 = rng.integers(-8, 8, size=(16, 16)).astype(np.float32)
    dev = [torch.from_numpy(bf16_bits(m).astype(np.int16)).cuda() for m in (A, B, A16, B16)]
    D = torch.zeros(16, 16, device="cuda")
    D16 = torch.zeros(16, 16, device="cuda")
    lib.call("pvdb_mfma_selftest", lib.ptr(dev[0]), lib.ptr(dev[1]), lib.ptr(D), lib.ptr(dev[2]),
             lib.ptr(dev[3]), lib.ptr(D16), lib.stream())
    np.testing.assert_array_equal(D.cpu().numpy(), A @ B)
    np.testing.assert_array_equal(D16.cpu().numpy(), A16 @ B16)


def test_mfma_f16_layout():
    from plenvdb_amd import lib, renderer  # noqa: F401  (declares the f16 self test)

    print("[case] mfma f16 layout", flush=True)
    rng = np.random.default_rng(1)
    A = rng.integers(-8, 8, size=(16, 32)).astype(np.float32)
    B = rng.integers(-8, 8, size=(32, 16)).astype(np.float32)
    dev = [torch.from_numpy(m.astype(np.float16).view(np.int16)).cuda() for m in (A, B)]
    D = torch.zeros(16, 16, device="cuda")
    lib.call("pvdb_mfma_selftest_f16", lib.ptr(dev[0]), lib.ptr(dev[1]), lib.ptr(D), lib.stream())
    np.testing.assert_array_equal(D.cpu().numpy(), A @ B)
