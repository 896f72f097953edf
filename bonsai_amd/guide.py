"""Path-guiding MLP on the GPU (no counterpart in the reference crate).

``GuideMLP`` owns the network's FP32 master parameters, its Adam state and the
BF16 copies the matrix-core kernels read. ``eval`` returns the 40 raw outputs
per sample (8 lobes x (logit, direction 3-vector, log-concentration));
``train`` takes one Adam step on the batch mean of ``-weight * log p(dir)``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .gpu import Context

N_PARAMS = N.BG_GUIDE_PARAMS
N_OUT = N.BG_GUIDE_OUTPUTS


def pack_samples(p, wo, n) -> np.ndarray:
    """(S, 12) float32 rows (p.xyz, 0, wo.xyz, 0, n.xyz, 0) as the C ABI takes them."""
    S = len(p)
    out = np.zeros((S, 12), np.float32)
    out[:, 0:3] = p
    out[:, 4:7] = wo
    out[:, 8:11] = n
    return out


def init_params(seed: int) -> np.ndarray:
    """Seeded initial parameters (bg_guide_init_params): He-uniform weights, zero biases."""
    out = np.empty(N_PARAMS, np.float32)
    N.check(N.lib().bg_guide_init_params(int(seed), out.ctypes.data))
    return out


class GuideMLP:
    def __init__(self, ctx: Context, params: np.ndarray | int, max_batch: int):
        """params: BG_GUIDE_PARAMS floats, or an int seed for init_params."""
        if isinstance(params, (int, np.integer)):
            params = init_params(int(params))
        params = np.ascontiguousarray(params, np.float32)
        if params.shape != (N_PARAMS,):
            raise ValueError(f"params must have shape ({N_PARAMS},)")
        self.ctx = ctx
        self.max_batch = int(max_batch)
        h = C.c_void_p()
        N.check(N.lib().bg_guide_create(ctx.h, params.ctypes.data, self.max_batch, C.byref(h)), ctx.h)
        self.h = h

    def close(self):
        if self.h:
            N.lib().bg_guide_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params: np.ndarray) -> None:
        params = np.ascontiguousarray(params, np.float32)
        N.check(N.lib().bg_guide_set_params(self.h, params.ctypes.data), self.ctx.h)

    def params(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(master parameters, Adam m, Adam v), FP32."""
        p, m, v = (np.empty(N_PARAMS, np.float32) for _ in range(3))
        N.check(N.lib().bg_guide_get_params(self.h, p.ctypes.data, m.ctypes.data, v.ctypes.data), self.ctx.h)
        return p, m, v

    def snapshot(self) -> bytes:
        """Parameters, Adam state, position box and replay ring (bg_guide_snapshot)."""
        size = C.c_uint64(0)
        N.check(N.lib().bg_guide_snapshot(self.h, None, C.byref(size)), self.ctx.h)
        buf = C.create_string_buffer(size.value)
        N.check(N.lib().bg_guide_snapshot(self.h, buf, C.byref(size)), self.ctx.h)
        return buf.raw[:size.value]

    def restore(self, snap: bytes) -> None:
        """Continue from a snapshot (bg_guide_restore): later calls give bit-identical results."""
        buf = C.create_string_buffer(snap, len(snap))
        N.check(N.lib().bg_guide_restore(self.h, buf, len(snap)), self.ctx.h)

    def set_bounds(self, lo, hi) -> None:
        lo = np.ascontiguousarray(lo, np.float32)
        hi = np.ascontiguousarray(hi, np.float32)
        N.check(N.lib().bg_guide_set_bounds(self.h, lo.ctypes.data, hi.ctypes.data), self.ctx.h)

    def eval(self, samples: np.ndarray) -> np.ndarray:
        samples = np.ascontiguousarray(samples, np.float32)
        n = len(samples)
        out = np.empty((n, N_OUT), np.float32)
        N.check(N.lib().bg_guide_eval(self.h, samples.ctypes.data, n, out.ctypes.data), self.ctx.h)
        return out

    def train(self, samples: np.ndarray, targets: np.ndarray, lr: float = 1e-2,
              want_grad: bool = False) -> tuple[float, np.ndarray | None]:
        """One Adam step. targets: (S, 4) = (direction.xyz, weight). Returns (mean loss before the
        step, gradient or None)."""
        samples = np.ascontiguousarray(samples, np.float32)
        targets = np.ascontiguousarray(targets, np.float32)
        n = len(samples)
        loss = C.c_double(0.0)
        grad = np.empty(N_PARAMS, np.float32) if want_grad else None
        N.check(N.lib().bg_guide_train(self.h, samples.ctypes.data, targets.ctypes.data, n, float(lr),
                                       C.byref(loss), grad.ctypes.data if want_grad else None), self.ctx.h)
        return loss.value, grad

    def eval_device(self, d_samples: int, n: int, d_out: int) -> float:
        """Raw outputs of n device-resident samples into device memory; returns the kernel ms."""
        ms = C.c_double(0.0)
        N.check(N.lib().bg_guide_eval_device(self.h, d_samples, n, d_out, C.byref(ms)), self.ctx.h)
        return ms.value

    def train_device(self, d_samples: int, d_targets: int, n: int, lr: float = 1e-2) -> float:
        """One Adam step on device-resident samples and targets; returns the kernel ms."""
        ms = C.c_double(0.0)
        N.check(N.lib().bg_guide_train_device(self.h, d_samples, d_targets, n, float(lr), C.byref(ms)), self.ctx.h)
        return ms.value
