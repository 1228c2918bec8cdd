"""Kernels of the learned adaptive grid (negf_mlp_*, extension) against the CPU oracle.

The oracle rounds the same float32 inputs to bfloat16 as the kernels do, so products
agree to FP32 accumulation error: bar 1e-4 relative to the largest entry for single
kernels.  A whole forward/backward pass also rounds intermediate activations to bfloat16,
where FP32 (GPU) and float64 (oracle) values a rounding step apart can land on different
sides: bar 1e-2 relative L2 per gradient, 1e-4 on the loss.
"""
import numpy as np
import pytest

from oracle import mlp_oracle as MO

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    from negftrans_amd import _lib, negf
    if negf.device_count() < 1:
        pytest.fail("no GPU visible to libnegf_gpu.so")
    return _lib.load()


def dev(x):
    from negftrans_amd.mlp import DeviceArray
    x = np.asarray(x, dtype=np.float32)
    return DeviceArray(x.shape, x)


def empty(shape):
    from negftrans_amd.mlp import DeviceArray
    return DeviceArray(shape)


def close(got, ref, rel):
    got, ref = np.asarray(got, dtype=np.float64), np.asarray(ref, dtype=np.float64)
    assert np.max(np.abs(got - ref)) <= rel * max(np.abs(ref).max(), 1e-30), np.max(np.abs(got - ref))


@pytest.mark.parametrize("M,N,K,ta,tb,mode,acc", [
    (70, 130, 37, 0, 0, 0, 0), (64, 64, 16, 1, 0, 0, 1), (33, 200, 2048, 0, 1, 0, 0),
    (129, 65, 100, 1, 1, 0, 0), (24, 512, 2048, 0, 0, 1, 0), (24, 1024, 512, 0, 0, 2, 0),
    (512, 512, 24, 1, 0, 0, 0)])
def test_gemm_vs_oracle(env, M, N, K, ta, tb, mode, acc):
    lib = env
    rng = np.random.default_rng(M * 7 + N + K)
    A = rng.standard_normal((K, M) if ta else (M, K)).astype(np.float32)
    B = rng.standard_normal((N, K) if tb else (K, N)).astype(np.float32)
    C0 = rng.standard_normal((M, N)).astype(np.float32)
    bias = rng.standard_normal(N).astype(np.float32)
    dA, dB, dC, db = dev(A), dev(B), dev(C0), dev(bias)
    dZ = empty(C0.shape)
    st = lib.negf_mlp_gemm(M, N, K, dA.ptr, A.shape[1], ta, dB.ptr, B.shape[1], tb, dC.ptr, N,
                           db.ptr, dZ.ptr, mode, acc, None)
    assert st == 0
    P = MO.gemm(A, B, bool(ta), bool(tb))
    got = dC.numpy()
    if mode == 0:
        close(got, P + (C0 if acc else 0), 1e-4)
    elif mode == 2:
        close(got, P + bias, 1e-4)
    else:
        close(dZ.numpy(), P + bias, 1e-4)
        close(got, MO.gelu(P + bias), 1e-4)


def test_elementwise_kernels_vs_oracle(env):
    lib = env
    rng = np.random.default_rng(3)
    Z = (3 * rng.standard_normal((17, 33))).astype(np.float32)
    dY = rng.standard_normal((17, 33)).astype(np.float32)
    # every device input is held in a variable until its kernel has run: a temporary
    # tensor's memory could be handed to the next allocation before the launch
    tZ, tdY = dev(Z), dev(dY)
    dZ = empty((17, 33))
    assert lib.negf_mlp_gelu_backward(Z.size, tZ.ptr, tdY.ptr, dZ.ptr, None) == 0
    close(dZ.numpy(), dY * MO.gelu_grad(Z), 1e-5)
    out = empty((33,))
    assert lib.negf_mlp_colsum(17, 33, tdY.ptr, 33, out.ptr, None) == 0
    close(out.numpy(), dY.astype(np.float64).sum(axis=0), 1e-5)
    # NLL, including clamped log-variances
    B, P = 9, 40
    O = rng.standard_normal((B, 2 * P)).astype(np.float32)
    O[0, P:P + 3] = [25.0, -25.0, 19.0]
    T = rng.random((B, P)).astype(np.float32)
    tO, tT = dev(O), dev(T)
    G = empty((B, 2 * P))
    loss = empty((1,))
    assert lib.negf_mlp_nll(B, P, tO.ptr, tT.ptr, G.ptr, loss.ptr, None) == 0
    lr, Gr = MO.nll(O.astype(np.float64), T.astype(np.float64))
    np.testing.assert_allclose(loss.numpy()[0], lr, rtol=1e-5)
    close(G.numpy(), Gr, 1e-5)
    # Adam, three steps
    w = rng.standard_normal(300).astype(np.float32)
    dw, dm, dv = dev(w), empty((300,)), empty((300,))
    wr, mr, vr = w.astype(np.float64), np.zeros(300), np.zeros(300)
    for step in (1, 2, 3):
        g = rng.standard_normal(300).astype(np.float32)
        tg = dev(g)
        lr_t = 1e-2 * np.sqrt(1 - 0.999 ** step) / (1 - 0.9 ** step)
        assert lib.negf_mlp_adam(300, dw.ptr, tg.ptr, dm.ptr, dv.ptr, lr_t,
                                 0.9, 0.999, 1e-8, None) == 0
        wr, mr, vr = MO.adam(wr, g.astype(np.float64), mr, vr, 1e-2, 0.9, 0.999, 1e-8, step)
    close(dw.numpy(), wr, 1e-5)


def test_forward_backward_vs_oracle(env):
    """One full loss-and-gradient pass of a small network, GPU kernels against the oracle
    (same bfloat16 operand rounding), and the training loop reduces the loss."""
    lib = env
    from negftrans_amd.mlp import TransmissionMLP
    rng = np.random.default_rng(5)
    P = 20
    net = TransmissionMLP(48, P, hidden=64, depth=3, seed=2)
    X = rng.standard_normal((24, 48)).astype(np.float32)
    T = (3 * rng.random((24, P))).astype(np.float32)
    tX, tT = dev(X), dev(T)
    loss, grads = net.loss_and_grads(tX, tT)
    params = net.params_numpy()
    lr, gr = MO.loss_and_grads(params, X, T.astype(np.float64), rounding=True)
    np.testing.assert_allclose(loss.numpy()[0], lr, rtol=1e-4)
    for (dW, db), (rW, rb) in zip(grads, gr):
        for g, r in ((dW.numpy(), rW), (db.numpy(), rb)):
            assert np.linalg.norm(g - r) <= 1e-2 * np.linalg.norm(r), np.linalg.norm(g - r) / np.linalg.norm(r)
    hist = net.fit(X, T, epochs=60, batch=8, lr=3e-3)
    assert hist[-1] < hist[0] - 1.0
