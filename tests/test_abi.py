"""CPU checks of the C-ABI library: it loads without a GPU and exports every entry
point include/negf_gpu.h declares, with the Python binding table in step."""
import os
import re

from negftrans_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "negf_gpu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return set(re.findall(r"\b(negf_[a-z0-9_]+)\s*\(", txt))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared()
    assert len(names) >= 14
    for n in names:
        assert hasattr(lib, n), n


def test_binding_table_matches_header():
    assert declared() == set(_lib.SIGNATURES)


def test_no_gpu_calls_fail_cleanly():
    lib = _lib.load()
    if lib.negf_device_count() > 0:
        return
    import ctypes
    h = ctypes.c_void_p()
    import numpy as np
    nx = np.zeros(1, dtype=np.int32)
    v = np.zeros(2, dtype=np.float64)
    st = lib.negf_ctx_create(ctypes.byref(h), 1, 2, _lib.iptr(nx), _lib.iptr(nx), 1.0, 1.0, 1.0,
                             0.0, 0.0, 0.0, _lib.dptr(v), 0, 0, 0)
    assert st < 0 and b"GPU" in lib.negf_last_error()
