"""The parameter-search surrogate (include/axicor_surrogate.h) from Python.

``Mlp`` wraps one network handle: forward passes, Adam steps, gradients and
parameter I/O in the library's padded layout. ``Scaler`` standardises
features and targets. All compute runs in libaxicor_surrogate.so's HIP
kernels; device memory is DeviceBuffer (gpu_malloc of the Day-Phase library).
Nothing falls back to the CPU: a missing library raises.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._ffi import ptr, surrogate_lib
from .engine import DeviceBuffer, GpuError

D_IN, HIDDEN, DEPTH, D_OUT = 16, 512, 3, 8


def _check(err: int, what: str) -> None:
    if err:
        raise GpuError(f"{what} failed (hipError={err})")


class Mlp:
    """d_in -> hidden x depth -> d_out, GELU, bf16 weights, fp32 accumulation."""

    def __init__(self, d_in: int = D_IN, hidden: int = HIDDEN, depth: int = DEPTH, d_out: int = D_OUT,
                 max_batch: int = 4096, seed: int = 1):
        self.lib = surrogate_lib()
        h = C.c_void_p()
        _check(self.lib.axc_mlp_create(d_in, hidden, depth, d_out, max_batch, seed, C.byref(h)), "axc_mlp_create")
        self.h = h.value
        self.d_in, self.hidden, self.depth, self.d_out, self.max_batch = d_in, hidden, depth, d_out, max_batch
        nl = C.c_int32()
        ins = np.zeros(depth + 1, np.int32)
        outs = np.zeros(depth + 1, np.int32)
        npar = C.c_int64()
        _check(self.lib.axc_mlp_layout(self.h, C.byref(nl), ptr(ins), ptr(outs), C.byref(npar)), "axc_mlp_layout")
        self.ins, self.outs, self.n_params = ins.tolist(), outs.tolist(), npar.value
        self._bufs: dict = {}

    # -- device staging ---------------------------------------------------
    def _buf(self, key: str, nbytes: int) -> DeviceBuffer:
        b = self._bufs.get(key)
        if b is None or b.nbytes < nbytes:
            if b is not None:
                b.free()
            b = DeviceBuffer(max(nbytes, 256))
            self._bufs[key] = b
        return b

    def _stage(self, key: str, a: np.ndarray) -> DeviceBuffer:
        a = np.ascontiguousarray(a, np.float32)
        b = self._buf(key, a.nbytes)
        b.upload(a)
        return b

    # -- compute ----------------------------------------------------------
    def forward_device(self, d_x: int, n: int, d_y: int, stream=None) -> None:
        _check(self.lib.axc_mlp_forward(self.h, d_x, n, d_y, stream), "axc_mlp_forward")

    def set_fused(self, enable: bool) -> bool:
        """Forward passes in one fused kernel (default) or layer by layer;
        returns whether this network's shape has the fused kernel."""
        r = self.lib.axc_mlp_set_fused_forward(self.h, 1 if enable else 0)
        if r < 0 or r > 1:
            raise RuntimeError(f"axc_mlp_set_fused_forward failed ({r})")
        return r == 1

    def predict(self, X: np.ndarray) -> np.ndarray:
        X = np.ascontiguousarray(X, np.float32)
        n = X.shape[0]
        dx = self._stage("x", X)
        dy = self._buf("y", n * self.d_out * 4)
        self.forward_device(dx.ptr, n, dy.ptr)
        return dy.download(np.float32, n * self.d_out).reshape(n, self.d_out)

    def train_step_device(self, d_x: int, d_t: int, n: int, lr: float, want_loss: bool = False, stream=None):
        loss = C.c_float()
        _check(self.lib.axc_mlp_train_step(self.h, d_x, d_t, n, lr, C.byref(loss) if want_loss else None, stream),
               "axc_mlp_train_step")
        return loss.value if want_loss else None

    def train_step(self, X: np.ndarray, T: np.ndarray, lr: float) -> float:
        dx = self._stage("x", X)
        dt = self._stage("t", T)
        return self.train_step_device(dx.ptr, dt.ptr, X.shape[0], lr, want_loss=True)

    def gradients(self, X: np.ndarray, T: np.ndarray) -> tuple[np.ndarray, float]:
        dx = self._stage("x", X)
        dt = self._stage("t", T)
        g = np.zeros(self.n_params, np.float32)
        loss = C.c_float()
        _check(self.lib.axc_mlp_gradients(self.h, dx.ptr, dt.ptr, X.shape[0], ptr(g), C.byref(loss), None),
               "axc_mlp_gradients")
        return g, loss.value

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.n_params, np.float32)
        _check(self.lib.axc_mlp_get_params(self.h, ptr(p)), "axc_mlp_get_params")
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n_params
        _check(self.lib.axc_mlp_set_params(self.h, ptr(p)), "axc_mlp_set_params")

    def reset_optimizer(self) -> None:
        _check(self.lib.axc_mlp_reset_optimizer(self.h), "axc_mlp_reset_optimizer")

    def destroy(self) -> None:
        for b in self._bufs.values():
            b.free()
        self._bufs.clear()
        if self.h:
            self.lib.axc_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def gemm(A: np.ndarray, B: np.ndarray, ta: bool = False, tb: bool = False) -> np.ndarray:
    """op(A) op(B) on the surrogate's MFMA GEMM kernel. A and B are given as
    bf16 bit patterns (uint16) in their stored layout; the result is fp32."""
    A = np.ascontiguousarray(A, np.uint16)
    B = np.ascontiguousarray(B, np.uint16)
    M = A.shape[1] if ta else A.shape[0]
    K = A.shape[0] if ta else A.shape[1]
    N = B.shape[0] if tb else B.shape[1]
    assert (B.shape[1] if tb else B.shape[0]) == K
    da, db = DeviceBuffer.from_array(A), DeviceBuffer.from_array(B)
    dc = DeviceBuffer(M * N * 4)
    try:
        _check(surrogate_lib().axc_mlp_gemm(M, N, K, int(ta), int(tb), da.ptr, A.shape[1], db.ptr, B.shape[1],
                                            dc.ptr, N, None), "axc_mlp_gemm")
        return dc.download(np.float32, M * N).reshape(M, N)
    finally:
        da.free()
        db.free()
        dc.free()


class Scaler:
    """Per-column standardisation fitted on a sample."""

    def __init__(self, X: np.ndarray):
        self.mu = X.mean(axis=0)
        self.sd = X.std(axis=0) + 1e-9

    def fwd(self, X: np.ndarray) -> np.ndarray:
        return ((X - self.mu) / self.sd).astype(np.float32)

    def inv(self, Z: np.ndarray) -> np.ndarray:
        return Z * self.sd + self.mu
