"""The C-ABI libraries load without a GPU and export every function that
include/*.h declares (no compute calls here)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = {"rtm_hip.h": "librtm_hip.so", "rtm_host.h": "librtm_host.so", "rtm_cnn.h": "librtm_hip.so"}


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    text = re.sub(r"^\s*#[^\n]*", " ", text, flags=re.M)
    names = re.findall(r"\b(rtm\w*)\s*\([^;{]*\)\s*;", text)
    return sorted(set(names))


@pytest.mark.parametrize("header", sorted(HEADERS))
def test_library_exports_declared_functions(header):
    names = _declared(header)
    assert len(names) >= 9, names
    lib = C.CDLL(os.path.join(ROOT, "rtm_gpu_amd", "lib", HEADERS[header]))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"{HEADERS[header]} lacks {missing}"


def test_least_squares_entry_points_declared():
    names = _declared("rtm_hip.h")
    for n in ("rtm_set_reflectivity", "rtm_born_begin", "rtm_born_advance", "rtm_adjoint_begin",
              "rtm_adjoint_advance", "rtm_adjoint_gradient"):
        assert n in names
