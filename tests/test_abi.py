"""The C ABI library: it loads and exports every symbol include/finchgpu.h declares.

No compute call is made here (these tests run without a GPU)."""
import ctypes
import os
import re

import pytest

from conftest import REPO
from finchgpu import _lib


def declared_symbols():
    text = open(os.path.join(REPO, "include", "finchgpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fgx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    syms = declared_symbols()
    for name in ("fgx_spmv_csr", "fgx_spmv_csc", "fgx_spmm_csr", "fgx_sddmm_csr",
                 "fgx_ewise_symbolic", "fgx_ewise_numeric", "fgx_spgemm_symbolic",
                 "fgx_spgemm_numeric", "fgx_batch_plan_create", "fgx_serve_run"):
        assert name in syms


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    # The Python binding covers the same set.
    assert sorted(_lib.SIGNATURES) == declared_symbols()


def test_julia_binding_calls_every_declared_symbol():
    # The Julia glue a Finch maintainer would load binds the same C ABI: each
    # declared entry point appears in one of its ccall targets.
    text = open(os.path.join(REPO, "finch.jl_amd", "julia", "FinchGPU.jl")).read()
    called = set(re.findall(r"ccall\(\(:(fgx_[a-z0-9_]+)", text))
    missing = [s for s in declared_symbols() if s not in called]
    assert not missing, missing


def test_version_and_error_paths_without_device():
    lib = _lib.load()
    assert lib.fgx_version() == 1
    rc = lib.fgx_spmv_csr(None, None, None, None, None)
    assert rc == _lib.EINVAL
    assert b"NULL ctx" in lib.fgx_last_error()
    with pytest.raises(_lib.FinchGPUError):
        _lib.check(rc, "probe")


def test_struct_layouts_match_header():
    # fgx_compressed: 3 x int64, 4 x int32, 3 pointers = 24 + 16 + 24 bytes.
    assert ctypes.sizeof(_lib.Compressed) == 64
    assert ctypes.sizeof(_lib.BatchItem) == 64 + 24
    assert ctypes.sizeof(_lib.Request) == 48
    assert ctypes.sizeof(_lib.ServeConfig) == 24
    assert ctypes.sizeof(_lib.RequestResult) == 24
