"""The C-ABI libraries load and export every declared entry point (CPU)."""
import ctypes as C
import os
import re

import pytest

from conftest import ROOT

HEADERS = {
    "include/h2hip.h": "h2lib_amd/libh2hip.so",
}


def declared(header):
    text = open(os.path.join(ROOT, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(h2\w+)\s*\(", text)
    return sorted(set(names))


@pytest.mark.parametrize("header", sorted(HEADERS))
def test_library_exports_declared_symbols(header):
    so = os.path.join(ROOT, HEADERS[header])
    assert os.path.exists(so), "build the library first"
    lib = C.CDLL(so, mode=C.RTLD_LOCAL)
    names = declared(header)
    assert len(names) > 5
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_import_without_gpu_fails_loudly_only_on_use():
    import h2lib_amd
    lib = h2lib_amd.hip()
    # no device in the build container: surface creation must fail with a
    # clear error, never fall back to a CPU path
    if lib.h2hip_device_count() == 0:
        import numpy as np
        with pytest.raises(h2lib_amd.H2HipError):
            h2lib_amd.Surface3d(np.eye(3), np.array([[0, 1, 2]]))
