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
