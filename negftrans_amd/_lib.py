"""ctypes binding of libnegf_gpu.so (the C-ABI of include/negf_gpu.h).

The library is built in-tree by negftrans_amd.build (or __graft_entry__.build()).
There is no fallback: if the shared object is missing, importing the product API
raises, so nothing can silently run on the CPU.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnegf_gpu.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p

# name -> (restype, argtypes); must match include/negf_gpu.h
SIGNATURES = {
    "negf_version": (ctypes.c_char_p, []),
    "negf_last_error": (ctypes.c_char_p, []),
    "negf_device_count": (ctypes.c_int, []),
    "negf_ctx_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, _ip, _ip,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "negf_ctx_destroy": (ctypes.c_int, [_vp]),
    "negf_transmission_batch": (ctypes.c_int, [_vp, _dp, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int,
                                               _dp, _dp]),
    "negf_rgf_tri_solve": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _dp, _dp, _dp]),
    "negf_density_batch": (ctypes.c_int, [_vp, _dp, _dp, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, _dp]),
    "negf_ldos_batch": (ctypes.c_int, [_vp, _dp, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, _dp]),
    "negf_neq_density_batch": (ctypes.c_int, [_vp, _dp, _dp, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, _dp]),
    "negf_pw_to_real": (ctypes.c_int, [ctypes.c_int, _ip, _ip, ctypes.c_int, _dp, ctypes.c_int,
                                       ctypes.c_int, _dp]),
    "negf_vhat_from_real": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int,
                                           ctypes.c_int, ctypes.c_int, _dp]),
    "negf_nonlocal_vhat_from_real": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int,
                                                    _dp]),
    "negf_inverse_batch": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, _dp]),
    "negf_inverse_batch_noalloc": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp]),
    "negf_matrix_product_batch": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]),
    "negf_last_kernel_seconds": (ctypes.c_double, [_vp]),
    "negf_last_fallbacks": (ctypes.c_int, [_vp]),
    "negf_last_variant": (ctypes.c_int, [_vp]),
    "negf_tune_seconds": (ctypes.c_double, [_vp, ctypes.c_int]),
    "negf_ctx_set_potential": (ctypes.c_int, [_vp, _dp]),
    "negf_multi_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, ctypes.POINTER(_vp)]),
    "negf_multi_transmission": (ctypes.c_int, [_vp, ctypes.POINTER(_dp), _ip, ctypes.POINTER(_dp),
                                               ctypes.POINTER(_dp), ctypes.POINTER(_dp), _ip,
                                               ctypes.POINTER(_dp)]),
    "negf_multi_replay": (ctypes.c_int, [_vp, ctypes.POINTER(_dp)]),
    "negf_multi_ldos": (ctypes.c_int, [_vp, ctypes.POINTER(_dp), _ip, ctypes.POINTER(_dp), ctypes.POINTER(_dp),
                                       ctypes.POINTER(_dp), _ip, ctypes.c_int, ctypes.POINTER(_ip),
                                       ctypes.POINTER(_dp)]),
    "negf_multi_last_seconds": (ctypes.c_double, [_vp]),
    "negf_multi_variant": (ctypes.c_int, [ctypes.c_int]),
    "negf_multi_destroy": (ctypes.c_int, [_vp]),
    # device-pointer entry points of the adaptive-grid network (pointers as integers)
    "negf_mlp_gemm": (ctypes.c_int, [ctypes.c_int] * 3 + [_vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int,
                                     ctypes.c_int, _vp, ctypes.c_int, _vp, _vp, ctypes.c_int, ctypes.c_int, _vp]),
    "negf_mlp_gelu_backward": (ctypes.c_int, [ctypes.c_int, _vp, _vp, _vp, _vp]),
    "negf_mlp_colsum": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int, _vp, _vp]),
    "negf_mlp_nll": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp, _vp]),
    "negf_mlp_adam": (ctypes.c_int, [ctypes.c_int, _vp, _vp, _vp, _vp, ctypes.c_float, ctypes.c_float,
                                     ctypes.c_float, ctypes.c_float, _vp]),
    "negf_dev_alloc": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_size_t]),
    "negf_dev_free": (ctypes.c_int, [_vp]),
    "negf_dev_copy": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t, ctypes.c_int]),
    "negf_hamiltonian_direct": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _ip, _ip, ctypes.c_double,
                                               ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp,
                                               ctypes.c_int, ctypes.c_int, _dp]),
    "negf_e_minus_h_from_h": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_double,
                                             ctypes.c_double, _ip, _ip] + [ctypes.c_double] * 7 + [_dp]),
    "negf_transmission_trace": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_double, _ip, _ip]
                                + [ctypes.c_double] * 7 + [_dp]),
    "negf_hamiltonian": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, _dp]),
    "negf_e_minus_h": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, _dp]),
}

_lib = None


class NegfError(RuntimeError):
    pass


def load():
    """Load the in-tree shared library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NegfError(f"{LIB_PATH} is missing: build it with `python -m negftrans_amd.build`")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(status, what):
    """Raise on negative status; return positive (zero-pivot slice) status unchanged."""
    if status < 0:
        msg = load().negf_last_error().decode(errors="replace")
        raise NegfError(f"{what} failed ({status}): {msg}")
    return status


def dptr(a):
    return a.ctypes.data_as(_dp)


def iptr(a):
    return a.ctypes.data_as(_ip)


def as_c128_fortran(a):
    return np.asfortranarray(np.asarray(a, dtype=np.complex128))
