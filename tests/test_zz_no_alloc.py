"""The allocation-free batched inverse (negf_inverse_batch_noalloc) against the
allocating one (negf_inverse_batch), through the C-ABI: the same kernel on the same
inputs, so the inverses and statuses must agree bit for bit.  Two calls in a row, the
second with a smaller input, since the work array keeps the previous call's contents;
a batch beyond the work array's size is rejected with an error."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    from negftrans_amd import _lib, negf
    if negf.device_count() < 1:
        pytest.fail("no GPU visible to libnegf_gpu.so")
    return _lib.load()


class DevBuf:
    def __init__(self, lib, nbytes):
        self.lib, self.p = lib, ctypes.c_void_p()
        assert lib.negf_dev_alloc(ctypes.byref(self.p), nbytes) == 0

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert self.lib.negf_dev_copy(self.p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 0) == 0

    def get(self, a):
        assert self.lib.negf_dev_copy(a.ctypes.data_as(ctypes.c_void_p), self.p, a.nbytes, 1) == 0
        return a

    def free(self):
        self.lib.negf_dev_free(self.p)


def batch(rng, N, count):
    A = rng.standard_normal((N, N, count)) + 1j * rng.standard_normal((N, N, count))
    A += 4.0 * N ** 0.5 * np.eye(N)[:, :, None]
    if count > 1:                                   # one matrix that needs row interchanges
        A[:, :, 1] = A[:, :, 1] * 1e-3 + np.eye(N)[::-1]
    return np.asfortranarray(A)


def noalloc_inverse(lib, A):
    N, _, count = A.shape
    dA, dI, dS = DevBuf(lib, A.nbytes), DevBuf(lib, A.nbytes), DevBuf(lib, 4 * count)
    try:
        dA.put(A.ravel(order="F"))
        st = lib.negf_inverse_batch_noalloc(N, count, dA.p, dI.p, dS.p, None)
        assert st == 0, lib.negf_last_error()
        out = dI.get(np.empty(N * N * count, dtype=np.complex128))
        status = dS.get(np.empty(count, dtype=np.int32))
    finally:
        for b in (dA, dI, dS):
            b.free()
    return out.reshape((N, N, count), order="F"), status


@pytest.mark.parametrize("N1,c1,N2,c2", [(317, 24, 61, 5), (100, 40, 100, 3), (600, 3, 17, 2)])
def test_noalloc_matches_allocating_entry(lib, N1, c1, N2, c2):
    from negftrans_amd import negf
    rng = np.random.default_rng(N1 + 7 * c1)
    for N, c in ((N1, c1), (N2, c2)):               # second call smaller than the first
        A = batch(rng, N, c)
        got, status = noalloc_inverse(lib, A)
        ref = negf.inverse(A)
        assert np.all(status == 0)
        assert np.array_equal(got, ref), f"N={N} count={c}: max diff {np.max(np.abs(got - ref)):.3e}"
        assert np.allclose(got[:, :, 0] @ A[:, :, 0], np.eye(N), atol=1e-10)


def test_noalloc_rejects_batch_beyond_scratch(lib):
    N = 317                                          # 1.97 MB of work area per matrix
    big = ctypes.c_void_p(16)                        # never dereferenced: rejected first
    st = lib.negf_inverse_batch_noalloc(N, 40, big, big, big, None)
    assert st < 0 and b"scratch" in lib.negf_last_error()
