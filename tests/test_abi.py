"""The C-ABI libraries load and export every function their headers
declare (no compute calls: runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    txt = open(os.path.join(ROOT, header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\**\s*(axbc_[a-z_0-9]+|[a-z_0-9]+cuda[a-z_0-9]*|get_[a-z]+|fincuda|initcuda|setcuda|assignDeviceToProcess|CudaMallocHostSet|CudaFreeHost)\s*\(", txt, flags=re.M)))


@pytest.mark.parametrize("header,lib", [("include/axbc.h", "axbc_amd/lib/libaxbc_hip.so")])
def test_exports(header, lib):
    path = os.path.join(ROOT, lib)
    assert os.path.exists(path), f"{lib} not built"
    L = ctypes.CDLL(path)
    names = declared(header)
    assert len(names) > 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_error_paths_without_gpu():
    from axbc_amd import _ffi
    L = _ffi.hip()
    # null context and bad arguments are rejected before any device call
    assert L.axbc_pivot(None, 0, 1.0, None) == -1
    assert L.axbc_sync(None) == -1
    assert b"null" in L.axbc_last_error()
