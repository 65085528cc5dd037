"""The C-ABI library loads without a GPU and exports every function that
include/*.h declares (no compute calls)."""
import ctypes
import os
import re

import pytest

import libapr_amd as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header=None):
    """functions declared by include/*.h (or by one header)"""
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h") or (header and fn != header):
            continue
        txt = open(os.path.join(ROOT, "include", fn)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b(aprh_[a-z0-9_]+)\s*\(", txt):
            names.add(m.group(1))
    names.discard("aprh_params_default")      # static inline in the header
    return sorted(names)


def test_library_exports_all_declared():
    if not os.path.exists(A.LIB_PATH):
        pytest.skip("libaprh.so not built")
    lib = ctypes.CDLL(A.LIB_PATH)
    io = set(declared_functions("aprh_io.h"))
    missing = [n for n in declared_functions() if n not in io and not hasattr(lib, n)]
    assert not missing, missing
    assert len(declared_functions()) > 20


def test_io_library_exports_all_declared():
    if not os.path.exists(A.IO_LIB_PATH):
        pytest.skip("libaprh_io.so not built")
    A.lib()
    lib = ctypes.CDLL(A.IO_LIB_PATH)
    missing = [n for n in declared_functions("aprh_io.h") if not hasattr(lib, n)]
    assert not missing, missing


def test_version_and_errors_without_gpu():
    if not os.path.exists(A.LIB_PATH):
        pytest.skip("libaprh.so not built")
    L = A.lib()
    assert L.aprh_version() >= 1
    # argument checks fire before any device call
    rc = L.aprh_get_apr_u16(None, None, 1, 1, 1, None, None)
    assert rc == A.APRH_ERR_INVALID_ARGUMENT
    assert b"NULL" in L.aprh_last_error() or b"context" in L.aprh_last_error()


def test_host_synth_is_deterministic():
    if not os.path.exists(A.LIB_PATH):
        pytest.skip("libaprh.so not built")
    a = A.synth_blobs((8, 16, 24), 5, 7)
    b = A.synth_blobs((8, 16, 24), 5, 7)
    assert (a == b).all() and a.max() > 100
