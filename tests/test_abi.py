"""The C-ABI library loads and exports every entry point include/*.h declares."""
import ctypes as C
import glob
import os
import re

from kratos_polrad_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b(kp_[a-z0-9_]+)\s*\(", txt):
            names.add(m.group(1))
    return names


def test_every_declared_symbol_is_exported():
    lib = C.CDLL(_abi.LIB_PATH)
    names = _declared()
    assert "kp_forward" in names and "kp_mesh_create" in names
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_view_covers_the_header():
    # every declared entry point has a ctypes signature in the Python view
    assert _declared() <= set(_abi.SIGNATURES)


def test_last_error_roundtrip():
    L = _abi.lib()
    h = C.c_void_p()
    st = L.kp_mesh_create((C.c_int32 * 3)(1, 1, 1), 3, (C.c_double * 3)(0, 0, 0),
                          (C.c_double * 3)(1, 1, 1), None, C.byref(h))
    assert st == _abi.KP_ERR_ARG
    assert b"even" in L.kp_last_error()
    st = L.kp_mesh_create((C.c_int32 * 3)(1, 1, 1), 4, (C.c_double * 3)(0, 0, 0),
                          (C.c_double * 3)(1, 1, 1), None, C.byref(h))
    assert st == _abi.KP_OK and L.kp_last_error() == b""
    L.kp_mesh_destroy(h)


def test_null_arguments_rejected():
    L = _abi.lib()
    assert L.kp_forward(None, None, None, None, None) == _abi.KP_ERR_ARG
    assert L.kp_set_media(None, None, None, None) == _abi.KP_ERR_ARG
    assert L.kp_mesh_finalize(None) == _abi.KP_ERR_ARG
