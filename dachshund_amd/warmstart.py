"""Learned initial guess of the Wiener-filter solve (ds_ws_* of the C ABI).

A point network (162 -> 64 -> 64 -> 64 -> 1, FP32 weights, BF16
activations) predicts each pixel's solution value from its 32 nearest
neighbours; the prediction becomes the solve's initial guess (the x
in/out argument of WFVector*::solve_pcg, lib/pcg.h:98).  Training and
inference run in the library's own HIP kernels.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib
from .wiener import default_context


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class WarmStart:
    def __init__(self, seed: int = 0, ctx=None, quantized=True):
        self.ctx = ctx or default_context()
        h = C.c_void_p()
        check(lib().ds_ws_create(self.ctx.h, seed, C.byref(h)), self.ctx.h,
              "ws_create")
        self.h = h
        self.num_params = lib().ds_ws_num_params()
        self.feature_dim = lib().ds_ws_feature_dim()
        self.set_quantized(quantized)

    def close(self):
        if getattr(self, "h", None):
            lib().ds_ws_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_quantized(self, on: bool):
        check(lib().ds_ws_set_quantized(self.h, int(bool(on))))

    @property
    def params(self) -> np.ndarray:
        P = np.empty(self.num_params, dtype=np.float32)
        check(lib().ds_ws_get_params(self.h, _fp(P), self.num_params),
              self.ctx.h)
        return P

    @params.setter
    def params(self, P):
        P = np.ascontiguousarray(P, dtype=np.float32)
        if P.shape != (self.num_params,):
            raise ValueError(f"expected {self.num_params} parameters")
        check(lib().ds_ws_set_params(self.h, _fp(P), self.num_params),
              self.ctx.h)

    def save(self, path):
        np.savez(path, params=self.params)

    def load(self, path):
        self.params = np.load(path, allow_pickle=False)["params"]

    def features(self, wf) -> np.ndarray:
        out = np.empty((wf.n, self.feature_dim), dtype=np.float32)
        check(lib().ds_ws_features(self.h, wf.h, _fp(out)), self.ctx.h,
              "ws_features")
        return out

    def predict_into(self, wf):
        """Write the predicted initial guess into wf's resident x (enqueued
        on the context's stream, ahead of any later call on it; the host
        does not wait: ctx.synchronize() does)."""
        check(lib().ds_ws_predict(self.h, wf.h), self.ctx.h, "ws_predict")

    def add_sample(self, wf, seed: int = 0):
        """Add wf (whose resident x holds its solution) to the training
        store."""
        check(lib().ds_ws_add_sample(self.h, wf.h, seed), self.ctx.h,
              "ws_add_sample")

    def num_samples(self) -> int:
        c = C.c_longlong()
        check(lib().ds_ws_num_samples(self.h, C.byref(c)))
        return c.value

    def clear_samples(self):
        check(lib().ds_ws_clear_samples(self.h))

    def train(self, epochs, batch=16384, lr=1e-3, lr_decay=0.9, seed=0):
        loss = np.zeros(max(epochs, 1))
        check(lib().ds_ws_train(self.h, epochs, batch, lr, lr_decay, seed,
                                loss.ctypes.data_as(C.POINTER(C.c_double))),
              self.ctx.h, "ws_train")
        return loss[:epochs]

    def eval_batch(self, feat, target):
        """Predictions, gradient of the mean squared error, loss."""
        feat = np.ascontiguousarray(feat, dtype=np.float32)
        target = np.ascontiguousarray(target, dtype=np.float32)
        n = len(feat)
        pred = np.empty(n, dtype=np.float32)
        grad = np.empty(self.num_params, dtype=np.float32)
        loss = C.c_double()
        check(lib().ds_ws_eval_batch(self.h, n, _fp(feat), _fp(target),
                                     _fp(pred), _fp(grad), C.byref(loss)),
              self.ctx.h, "ws_eval_batch")
        return pred, grad, loss.value
