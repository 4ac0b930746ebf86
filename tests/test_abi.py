"""The C-ABI library loads and exports every function include/daggerfft_hip.h
declares; argument validation fails cleanly without touching a GPU."""
import ctypes

import pytest


def test_library_exports_every_declared_symbol():
    from daggerfft import _lib as L
    lib = L.lib()
    names = L.header_symbols()
    assert len(names) >= 16
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_strerror_and_version():
    from daggerfft import _lib as L
    assert L.strerror(0) == "success"
    assert "unsupported" in L.strerror(L.DFFT_ERR_UNSUPPORTED_SIZE)
    assert L.strerror(12345) == "unknown status"
    a, b = ctypes.c_int(), ctypes.c_int()
    assert L.lib().dfft_version(ctypes.byref(a), ctypes.byref(b)) == 0


@pytest.mark.parametrize("rank,n,kind,sign,dims", [
    (0, (8,), 0, -1, (0,)),            # rank out of range
    (3, (8, 8, 8), 0, 0, (0, 1, 2)),   # bad sign
    (3, (8, 8, 8), 7, -1, (0, 1, 2)),  # bad kind
    (3, (8, 8, 8), 0, -1, (0, 0, 1)),  # repeated dim
    (3, (8, 8, 8), 0, -1, (0, 3)),     # dim out of range
    (3, (8, 0, 8), 0, -1, (0,)),       # zero length
    (3, (8, 8, 8), 1, +1, (0,)),       # R2C must be forward
    (3, (8, 8, 8), 2, -1, (0,)),       # C2R must be inverse
    (3, (8, 8, 8), 3, -1, (0,)),       # R2R without kinds
])
def test_plan_create_rejects_bad_arguments(rank, n, kind, sign, dims):
    from daggerfft import _lib as L
    h = ctypes.c_void_p()
    st = L.lib().dfft_plan_create(ctypes.byref(h), rank, L.i64(n), kind, sign, None, 1, 0, L.i32(dims),
                                  len(dims), 0)
    assert st == L.DFFT_ERR_ARG
    assert not h.value


def test_null_handles_rejected():
    from daggerfft import _lib as L
    lib = L.lib()
    assert lib.dfft_execute(None, None, None, None, None) == L.DFFT_ERR_ARG
    assert lib.dfft_plan_destroy(None) == L.DFFT_ERR_ARG
    assert lib.dfft_box_copy(None, None, None, None, None, None, None, 8, None) == L.DFFT_ERR_ARG
    assert lib.dfft_fill_uniform(None, 4, 1, 1, 0, None) == L.DFFT_ERR_ARG


def test_julia_shim_binds_declared_symbols():
    """Every entry point the Julia shim ccalls is declared in the header (the
    shim cannot run here: no Julia in this image)."""
    import os
    import re

    from conftest import ROOT
    from daggerfft import _lib as L
    jl = open(os.path.join(ROOT, "daggerfft.jl_amd", "julia", "DaggerFFTHIP.jl")).read()
    called = set(re.findall(r"\(:(dfft_[a-z0-9_]+),\s*libdfft\)", jl))
    assert called, "no ccall found"
    missing = called - set(L.header_symbols())
    assert not missing, missing
    for name in ("dfft_plan_create", "dfft_execute", "dfft_batch_execute", "dfft_server_submit"):
        assert name in called
