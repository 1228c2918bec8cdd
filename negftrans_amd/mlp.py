"""The network of the learned adaptive energy grid (extension; NEGFtrans has no learned
component).

A multilayer perceptron maps low-order Fourier samples of a device potential to T(E) on a
fixed energy grid plus a per-bin log-variance.  Every product, activation, loss and
optimizer step runs in the library's HIP kernels (negf_mlp_* in libnegf_gpu.so): products
on the matrix cores with bfloat16 operands and FP32 accumulation, GELU and its derivative,
the Gaussian negative log-likelihood and Adam in FP32.  Device buffers come from the
library as well (negf_dev_alloc / negf_dev_copy), so the network shares the library's
HIP runtime and needs nothing else on the GPU.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib
from ._lib import check


class DeviceArray:
    """A float32 row-major matrix (or vector) in device memory.  `rows(a, b)` is a view
    of rows a..b-1 that shares the allocation."""

    def __init__(self, shape, data=None, _view_of=None, _ptr=None):
        self._lib = _lib.load()
        self.shape = tuple(int(x) for x in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.nbytes = 4 * int(np.prod(self.shape))
        self._owner = _view_of
        if _view_of is not None:
            self.ptr = _ptr
            return
        p = ctypes.c_void_p()
        check(self._lib.negf_dev_alloc(ctypes.byref(p), self.nbytes), "negf_dev_alloc")
        self.ptr = p.value
        self.upload(np.zeros(self.shape, dtype=np.float32) if data is None else data)

    def upload(self, data):
        a = np.ascontiguousarray(data, dtype=np.float32)
        if a.size * 4 != self.nbytes:
            raise ValueError(f"upload of {a.shape} into {self.shape}")
        check(self._lib.negf_dev_copy(self.ptr, a.ctypes.data, self.nbytes, 0), "negf_dev_copy")

    def numpy(self):
        a = np.empty(self.shape, dtype=np.float32)
        check(self._lib.negf_dev_copy(a.ctypes.data, self.ptr, self.nbytes, 1), "negf_dev_copy")
        return a

    def rows(self, a, b):
        b = min(b, self.shape[0])
        cols = self.shape[1] if len(self.shape) > 1 else 1
        shape = (b - a,) + self.shape[1:]
        return DeviceArray(shape, _view_of=self, _ptr=self.ptr + 4 * a * cols)

    @property
    def ld(self):
        return self.shape[1] if len(self.shape) > 1 else 1

    def __del__(self):
        if getattr(self, "_owner", 1) is None and getattr(self, "ptr", None):
            try:
                self._lib.negf_dev_free(self.ptr)
            except Exception:
                pass
            self.ptr = None


class TransmissionMLP:
    """sizes: [d_in, hidden x depth, 2 P].  GELU after every hidden layer; the last layer
    is linear and its output is [mean of T per bin | log-variance per bin].  Work is
    queued on the default stream; reading a result back waits for it."""

    def __init__(self, d_in: int, n_bins: int, hidden: int = 512, depth: int = 3, seed: int = 0):
        self._lib = _lib.load()
        self.n_bins = n_bins
        self.sizes = [d_in] + [hidden] * depth + [2 * n_bins]
        rng = np.random.default_rng(seed)
        self.W, self.b, self.mW, self.vW, self.mb, self.vb = [], [], [], [], [], []
        for li, (a, c) in enumerate(zip(self.sizes[:-1], self.sizes[1:])):
            scale = math.sqrt(2.0 / a) if li < depth else 0.1 / math.sqrt(a)
            self.W.append(DeviceArray((a, c), rng.standard_normal((a, c)) * scale))
            self.b.append(DeviceArray((c,)))
            self.mW.append(DeviceArray((a, c)))
            self.vW.append(DeviceArray((a, c)))
            self.mb.append(DeviceArray((c,)))
            self.vb.append(DeviceArray((c,)))
        self.steps = 0
        self._bufs = {}

    def _buf(self, key, shape):
        """Working buffer reused across steps (one per role and shape).  All work runs on
        the default stream, so a step's kernels follow the previous step's in order."""
        k = (key, tuple(shape))
        if k not in self._bufs:
            self._bufs[k] = DeviceArray(shape)
        return self._bufs[k]

    # -------------------------------------------------------------- kernels
    def _gemm(self, A, B, C, transA=0, transB=0, bias=None, Z=None, mode=0, accumulate=0):
        M = A.shape[1] if transA else A.shape[0]
        K = A.shape[0] if transA else A.shape[1]
        N = B.shape[0] if transB else B.shape[1]
        check(self._lib.negf_mlp_gemm(M, N, K, A.ptr, A.ld, transA, B.ptr, B.ld, transB, C.ptr, C.ld,
                                      None if bias is None else bias.ptr, None if Z is None else Z.ptr, mode,
                                      accumulate, None), "negf_mlp_gemm")

    # -------------------------------------------------------------- passes
    def forward(self, X):
        """X: (B, d_in) DeviceArray.  Returns (output (B, 2P), cache)."""
        h = X
        cache = []
        L = len(self.W)
        for li in range(L):
            out = self._buf(("out", li), (X.shape[0], self.sizes[li + 1]))
            if li < L - 1:
                z = self._buf(("z", li), out.shape)
                self._gemm(h, self.W[li], out, bias=self.b[li], Z=z, mode=1)
            else:
                z = out
                self._gemm(h, self.W[li], out, bias=self.b[li], mode=2)
            cache.append((h, z))
            h = out
        return h, cache

    def backward(self, cache, G):
        """G = d loss / d output.  Returns [(dW, db)] per layer."""
        grads = [None] * len(self.W)
        g = G
        for li in range(len(self.W) - 1, -1, -1):
            h_in, z = cache[li]
            if li == len(self.W) - 1:
                dz = g
            else:
                dz = self._buf(("dz", li), z.shape)
                check(self._lib.negf_mlp_gelu_backward(int(np.prod(z.shape)), z.ptr, g.ptr, dz.ptr, None),
                      "negf_mlp_gelu_backward")
            dW = self._buf(("dW", li), self.W[li].shape)
            self._gemm(h_in, dz, dW, transA=1)
            db = self._buf(("db", li), self.b[li].shape)
            check(self._lib.negf_mlp_colsum(dz.shape[0], dz.shape[1], dz.ptr, dz.ld, db.ptr, None), "negf_mlp_colsum")
            grads[li] = (dW, db)
            if li > 0:
                gn = self._buf(("g", li), (dz.shape[0], self.sizes[li]))
                self._gemm(dz, self.W[li], gn, transB=1)
                g = gn
        return grads

    def loss_and_grads(self, X, T, loss=None):
        """Gaussian NLL of T (B, P) and its parameter gradients (held in the model's working
        buffers until the next pass).  The loss is added into `loss` (a one-element
        DeviceArray, allocated zeroed when not given)."""
        out, cache = self.forward(X)
        G = self._buf("G", out.shape)
        if loss is None:
            loss = DeviceArray((1,))
        check(self._lib.negf_mlp_nll(T.shape[0], T.shape[1], out.ptr, T.ptr, G.ptr, loss.ptr, None), "negf_mlp_nll")
        return loss, self.backward(cache, G)

    def adam_step(self, grads, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.steps += 1
        lr_t = lr * math.sqrt(1 - b2 ** self.steps) / (1 - b1 ** self.steps)
        for li, (dW, db) in enumerate(grads):
            for w, g, m, v in ((self.W[li], dW, self.mW[li], self.vW[li]), (self.b[li], db, self.mb[li], self.vb[li])):
                check(self._lib.negf_mlp_adam(int(np.prod(w.shape)), w.ptr, g.ptr, m.ptr, v.ptr, lr_t, b1, b2, eps,
                                              None), "negf_mlp_adam")

    def fit(self, X, T, epochs=200, batch=64, lr=1e-3, seed=0, log_every=0):
        """Minibatch Adam on (X, T) (numpy).  Each epoch uploads a host-shuffled copy of the
        set, so minibatches are contiguous row views; per-batch losses land in a device
        buffer read once per epoch.  Returns the mean loss per epoch."""
        X = np.asarray(X, dtype=np.float32)
        T = np.asarray(T, dtype=np.float32)
        rng = np.random.default_rng(seed)
        n = X.shape[0]
        nb = (n + batch - 1) // batch
        Xd, Td = DeviceArray(X.shape), DeviceArray(T.shape)
        hist = []
        for ep in range(epochs):
            perm = rng.permutation(n)
            Xd.upload(X[perm])
            Td.upload(T[perm])
            losses = DeviceArray((nb,))
            for k in range(nb):
                s = k * batch
                _, grads = self.loss_and_grads(Xd.rows(s, s + batch), Td.rows(s, s + batch), losses.rows(k, k + 1))
                self.adam_step(grads, lr=lr)
            hist.append(float(losses.numpy().mean()))
            if log_every and (ep + 1) % log_every == 0:
                print(f"# epoch {ep + 1}: loss {hist[-1]:.5f}", flush=True)
        return hist

    def predict(self, X):
        """(mean T, log-variance) per bin, numpy float64 arrays of shape (B, P)."""
        X = np.asarray(X, dtype=np.float32)
        out, _ = self.forward(DeviceArray(X.shape, X))
        o = out.numpy().astype(np.float64)
        return o[:, :self.n_bins], o[:, self.n_bins:]

    def params_numpy(self):
        return [(W.numpy().astype(np.float64), b.numpy().astype(np.float64)) for W, b in zip(self.W, self.b)]
