"""The C-ABI libraries load and export every function include/*.h declares (no GPU
calls), and the Python binding lists them all."""
import ctypes
import os
import re
import subprocess

import pytest

heom_amd = pytest.importorskip("heom_amd")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char\*)\s+(heom_\w+)\s*\(", text, flags=re.M)))


@pytest.mark.parametrize("lib,header,listed", [
    (heom_amd.LIB_PATH, heom_amd.HEADER, heom_amd.EXPORTED),
    (heom_amd.HOST_LIB_PATH, heom_amd.HOST_HEADER, heom_amd.HOST_EXPORTED),
])
def test_library_exports_every_declared_symbol(lib, header, listed):
    assert os.path.exists(lib), "build the libraries first (make -C dm-heom_amd)"
    names = declared(header)
    assert len(names) >= 8
    handle = ctypes.CDLL(lib)
    for n in names:
        assert hasattr(handle, n), n
    assert sorted(listed) == names
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (heom_\w+)", out))
    assert set(names) <= exported


def test_abi_version_and_error_path_without_gpu():
    L = heom_amd.load()
    assert L.heom_abi_version() == 2
    # a NULL descriptor is rejected before any device call
    h = ctypes.c_void_p()
    rc = L.heom_create(None, ctypes.byref(h))
    assert rc == heom_amd.HEOM_E_INVALID
    assert b"NULL" in L.heom_last_error()


def test_apps_print_usage():
    for app in ("heom_population_dynamics", "heom_linear_absorption", "heom_thermal_state_search"):
        exe = os.path.join(ROOT, "dm-heom_amd", "bin", app)
        assert os.path.exists(exe)
        r = subprocess.run([exe, "--help"], capture_output=True, text=True, timeout=60)
        assert r.returncode != 0
        assert "Usage" in r.stderr


def test_search_handle_checks_without_gpu():
    """heom_create_tss refuses bad arguments before any device call: a NULL S_u array
    (read on the host for eps - S_u), FP32, a zero epsilon."""
    L = heom_amd.load()
    N = 2
    c = (ctypes.c_double * (2 * N * N))(1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0)
    ev = (ctypes.c_double * (2 * N))()
    eps = (ctypes.c_double * 2)(1e14, 0.0)
    h = ctypes.c_void_p()
    d = heom_amd.Desc(states=N, matsubaras=1, baths=1, baths_max_per_state=1, width=1,
                      precision=heom_amd.HEOM_PRECISION_FP64, matrices=3)
    rc = L.heom_create_tss(ctypes.byref(d), c, ev, eps, ctypes.byref(h))
    assert rc == heom_amd.HEOM_E_INVALID and b"pf_same_sum" in L.heom_last_error()
    assert not h.value
    d.precision = heom_amd.HEOM_PRECISION_FP32
    assert L.heom_create_tss(ctypes.byref(d), c, ev, eps, ctypes.byref(h)) == heom_amd.HEOM_E_UNSUPPORTED
    d.precision = heom_amd.HEOM_PRECISION_FP64
    pss = (ctypes.c_double * 6)()
    d.pf_same_sum = ctypes.cast(pss, ctypes.c_void_p)
    zero = (ctypes.c_double * 2)()
    assert L.heom_create_tss(ctypes.byref(d), c, ev, zero, ctypes.byref(h)) == heom_amd.HEOM_E_INVALID
    assert b"epsilon" in L.heom_last_error()
