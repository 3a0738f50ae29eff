"""Python handle over the warm-start network C ABI (include/chemsim.h, cs_cnn_*)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from .capi import call
from .lattice import Context, _ptr


class Net:
    def __init__(self, ctx: Context, height: int, width: int, base: int = 32, batch: int = 8, seed: int = 0):
        self.ctx = ctx
        self.H, self.W, self.base, self.batch = int(height), int(width), int(base), int(batch)
        h = C.c_void_p()
        call("cs_cnn_create", ctx.handle, self.H, self.W, self.base, self.batch, int(seed) & 0xffffffff, C.byref(h))
        self.handle = h

    @property
    def param_count(self) -> int:
        n = C.c_int64()
        call("cs_cnn_param_count", self.handle, C.byref(n))
        return n.value

    def get_params(self) -> np.ndarray:
        out = np.empty(self.param_count, np.float32)
        call("cs_cnn_get_params", self.handle, _ptr(out))
        return out

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.param_count
        call("cs_cnn_set_params", self.handle, _ptr(p))

    def get_grads(self) -> np.ndarray:
        out = np.empty(self.param_count, np.float32)
        call("cs_cnn_get_grads", self.handle, _ptr(out))
        return out

    def forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        n = x.shape[0]
        assert x.shape[1:] == (self.H, self.W, 8)
        out = np.empty((n, self.H, self.W, 3), np.float32)
        for s in range(0, n, self.batch):
            k = min(self.batch, n - s)
            xs = np.ascontiguousarray(x[s:s + k])
            ys = np.empty((k, self.H, self.W, 3), np.float32)
            call("cs_cnn_forward", self.handle, k, _ptr(xs), _ptr(ys))
            out[s:s + k] = ys
        return out

    def loss_grad(self, x, t, w) -> float:
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        w = np.ascontiguousarray(w, dtype=np.float32)
        loss = C.c_float()
        call("cs_cnn_loss_grad", self.handle, x.shape[0], _ptr(x), _ptr(t), _ptr(w), C.byref(loss))
        return loss.value

    def adam_step(self, lr: float, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8) -> None:
        call("cs_cnn_adam_step", self.handle, float(lr), float(beta1), float(beta2), float(eps))

    def set_dataset(self, x, t, w) -> None:
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        w = np.ascontiguousarray(w, dtype=np.float32)
        call("cs_cnn_set_dataset", self.handle, x.shape[0], _ptr(x), _ptr(t), _ptr(w))

    def train_epoch(self, order, lr: float) -> float:
        order = np.ascontiguousarray(order, dtype=np.int32)
        loss = C.c_float()
        call("cs_cnn_train_epoch", self.handle, _ptr(order), int(order.size), float(lr), C.byref(loss))
        return loss.value

    def time_step(self, iters: int = 5):
        ms = C.c_float()
        fl = C.c_double()
        call("cs_cnn_time_step", self.handle, int(iters), C.byref(ms), C.byref(fl))
        return ms.value, fl.value

    def close(self):
        from . import capi
        if getattr(self, "handle", None):
            capi.lib.cs_cnn_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gemm_bf16(ctx: Context, A_bf16: np.ndarray, B_bf16: np.ndarray):
    """C = A . B^T on the MFMA tile kernel; A [M][K], B [N][K] as uint16 bf16 bits."""
    A = np.ascontiguousarray(A_bf16, dtype=np.uint16)
    B = np.ascontiguousarray(B_bf16, dtype=np.uint16)
    M, K = A.shape
    N = B.shape[0]
    Cm = np.empty((M, N), np.float32)
    ms = C.c_float()
    call("cs_gemm_bf16", ctx.handle, M, N, K, _ptr(A), _ptr(B), _ptr(Cm), C.byref(ms))
    return Cm, ms.value
