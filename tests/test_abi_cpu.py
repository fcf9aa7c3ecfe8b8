"""The C ABI: libpffrg.so loads without a GPU and exports every entry point that
include/pffrg.h declares, and the Python binding covers exactly that set."""
import ctypes
import re
from pathlib import Path

import pytest

from pffrg import native

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "pffrg.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pffrg_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "pffrg_compute_dgamma_1l" in syms and "pffrg_compute_gamma" in syms
    assert len(syms) >= 30


@pytest.mark.skipif(not native.LIB_PATH.exists(), reason="libpffrg.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(native.LIB_PATH))
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_matches_header():
    assert sorted(native.exported_symbols()) == declared_symbols()


@pytest.mark.skipif(not native.LIB_PATH.exists(), reason="libpffrg.so not built")
def test_create_without_device_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    lib = native.load()
    dims = native.Dims(4, 10, 10, 10, 0.5, 2, 0)
    h = ctypes.c_void_p()
    rc = lib.pffrg_create(ctypes.byref(dims), ctypes.byref(h))
    assert rc != 0


@pytest.mark.skipif(not native.LIB_PATH.exists(), reason="libpffrg.so not built")
@pytest.mark.parametrize("dims", [
    (0, 10, 10, 10, 0.5, 2, 0),    # no sites
    (65, 10, 10, 10, 0.5, 2, 0),   # more sites than lanes of a wavefront
    (4, 1, 10, 10, 0.5, 2, 0),     # a one-point Σ mesh
    (4, 10, 65, 10, 0.5, 2, 0),    # more Ω points than the LDS tables hold
    (4, 10, 10, 65, 0.5, 2, 0),    # more ν points than the LDS tables hold
    (4, 10, 10, 10, 0.5, 0, 0),    # no slots
    (4, 10, 10, 10, 0.0, 2, 0),    # S = 0
])
def test_create_rejects_dims_out_of_range(dims):
    # the size limits are checked before any device call, so they hold without a GPU
    lib = native.load()
    h = ctypes.c_void_p()
    assert lib.pffrg_create(ctypes.byref(native.Dims(*dims)), ctypes.byref(h)) == -4  # PFFRG_ERR_LIMIT
    assert not h
    assert lib.pffrg_create(None, ctypes.byref(h)) == -1  # PFFRG_ERR_ARG
