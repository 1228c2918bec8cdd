"""CPU oracle of the adaptive-grid network kernels (negftrans_amd/csrc/negf_mlp.hip).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP kernels; the
product path never calls it.

No reference counterpart: NEGFtrans has no learned component.  The restatement follows
the kernels' stated arithmetic: products take both operands rounded to bfloat16 (round to
nearest even, the rounding the kernels apply when staging in LDS) and are accumulated
here in float64 (the kernels accumulate in FP32, so the two agree to FP32 accumulation
error).  Element-wise parts are float64 here, FP32 on the GPU.

Pins (tests/test_mlp_oracle.py): with rounding switched off, every analytic gradient of
loss_and_grads matches central finite differences in float64; bf16_round matches the
bit pattern rule on hand-picked values (ties to even, NaN, overflow to inf); GELU matches
its closed form and its derivative a finite difference.
"""
from __future__ import annotations

import math

import numpy as np

SMAX = 20.0


def bf16_round(x):
    """float32 values rounded to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, (u >> 16) | 0x40, r)
    out = (r.astype(np.uint32) << 16).view(np.float32)
    return out.astype(np.float64).reshape(a.shape)


def gemm(A, B, transA=False, transB=False, rounding=True):
    """op(A) op(B) with bf16-rounded operands (rounding=True), float64 accumulation."""
    a = A.T if transA else A
    b = B.T if transB else B
    if rounding:
        a, b = bf16_round(a), bf16_round(b)
    return np.asarray(a, dtype=np.float64) @ np.asarray(b, dtype=np.float64)


_erf = np.vectorize(math.erf)


def gelu(z):
    z = np.asarray(z, dtype=np.float64)
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    z = np.asarray(z, dtype=np.float64)
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def nll(O, T):
    """Gaussian NLL of T (B x P) under O = [mu | log-variance] (B x 2P) and its gradient."""
    B, P = T.shape
    mu, sr = O[:, :P], O[:, P:]
    s = np.clip(sr, -SMAX, SMAX)
    r = T - mu
    ex = np.exp(-s)
    sc = 1.0 / (B * P)
    loss = float(np.sum(0.5 * (r * r * ex + s)) * sc)
    G = np.empty_like(O, dtype=np.float64)
    G[:, :P] = -r * ex * sc
    G[:, P:] = np.where((sr > -SMAX) & (sr < SMAX), 0.5 * (1.0 - r * r * ex) * sc, 0.0)
    return loss, G


def forward(params, X, rounding=True):
    """params: list of (W (K x N), b (N,)); GELU after every layer but the last.
    Returns (output, cache of (input, pre-activation) per layer)."""
    h = np.asarray(X, dtype=np.float64)
    cache = []
    for li, (W, b) in enumerate(params):
        z = gemm(h, W, rounding=rounding) + b
        cache.append((h, z))
        h = gelu(z) if li < len(params) - 1 else z
    return h, cache


def loss_and_grads(params, X, T, rounding=True):
    """NLL loss and its gradient with respect to every W and b (same arithmetic as the
    kernels' backward pass: dX = dZ W^T, dW = X^T dZ, db = column sums of dZ)."""
    out, cache = forward(params, X, rounding)
    loss, g = nll(out, T)
    grads = [None] * len(params)
    for li in range(len(params) - 1, -1, -1):
        h_in, z = cache[li]
        dz = g if li == len(params) - 1 else g * gelu_grad(z)
        W, _ = params[li]
        grads[li] = (gemm(h_in, dz, transA=True, rounding=rounding), dz.sum(axis=0))
        if li > 0:
            g = gemm(dz, W, transB=True, rounding=rounding)
    return loss, grads


def adam(w, g, m, v, lr, b1, b2, eps, step):
    """One Adam update (bias correction folded into the learning rate, as k_adam)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    lr_t = lr * math.sqrt(1 - b2 ** step) / (1 - b1 ** step)
    return w - lr_t * m / (np.sqrt(v) + eps), m, v
