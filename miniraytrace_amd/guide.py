"""Standalone path-guiding network (libmrt.so, mini_raytrace_ext.h section 5).

The same HIP kernels the engine uses for guided rendering and online
training, driven from host arrays: BF16 MFMA forward / backward with FP32
accumulation, FP32 master weights and Adam.

Records are float32 arrays of shape (N, 16):
    pos01.xyz, weight, n.xyz, 0, wo.xyz, 0, wi.xyz, 0
(`make_records` builds them).  Raw outputs are (N, 40): per lobe k the mean
direction (unnormalised) at 5k..5k+2, log kappa at 5k+3, weight logit at 5k+4.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L

PARAM_COUNT = L.MRTX_GUIDE_PARAM_COUNT
OUTPUTS = L.MRTX_GUIDE_OUTPUTS


def make_records(pos01, n, wo, wi=None, weight=None) -> np.ndarray:
    pos01 = np.asarray(pos01, np.float32)
    count = pos01.shape[0]
    r = np.zeros((count, 16), np.float32)
    r[:, 0:3] = pos01
    r[:, 4:7] = np.asarray(n, np.float32)
    r[:, 8:11] = np.asarray(wo, np.float32)
    if wi is not None:
        r[:, 12:15] = np.asarray(wi, np.float32)
    r[:, 3] = 1.0 if weight is None else np.asarray(weight, np.float32)
    return r


class GuideNetwork:
    def __init__(self, max_batch: int = 65536, seed: int = 1, gpu_index1: int = 0):
        h = C.c_void_p()
        L.check(L.lib().mrtxCreateGuideNetwork(gpu_index1, max_batch, seed, C.byref(h)), "mrtxCreateGuideNetwork")
        self._h = h

    def close(self) -> None:
        if self._h:
            L.lib().mrtxDestroyGuideNetwork(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self) -> np.ndarray:
        out = np.empty(PARAM_COUNT, np.float32)
        L.check(L.lib().mrtxGuideNetworkGetParams(self._h, out.ctypes.data, PARAM_COUNT), "mrtxGuideNetworkGetParams")
        return out

    @params.setter
    def params(self, value) -> None:
        p = np.ascontiguousarray(value, np.float32)
        if p.size != PARAM_COUNT:
            raise ValueError(f"expected {PARAM_COUNT} parameters")
        L.check(L.lib().mrtxGuideNetworkSetParams(self._h, p.ctypes.data, PARAM_COUNT), "mrtxGuideNetworkSetParams")

    def adam_state(self):
        m = np.empty(PARAM_COUNT, np.float32)
        v = np.empty(PARAM_COUNT, np.float32)
        t = C.c_uint32()
        L.check(L.lib().mrtxGuideNetworkGetAdam(self._h, m.ctypes.data, v.ctypes.data, PARAM_COUNT, C.byref(t)),
                "mrtxGuideNetworkGetAdam")
        return m, v, t.value

    def infer(self, records: np.ndarray) -> np.ndarray:
        r = np.ascontiguousarray(records, np.float32)
        out = np.empty((r.shape[0], OUTPUTS), np.float32)
        L.check(L.lib().mrtxGuideNetworkInfer(self._h, r.ctypes.data, r.shape[0], out.ctypes.data),
                "mrtxGuideNetworkInfer")
        return out

    def gradient(self, records: np.ndarray):
        r = np.ascontiguousarray(records, np.float32)
        g = np.empty(PARAM_COUNT, np.float32)
        loss = C.c_float()
        L.check(L.lib().mrtxGuideNetworkGradient(self._h, r.ctypes.data, r.shape[0], g.ctypes.data, PARAM_COUNT,
                                                 C.byref(loss)), "mrtxGuideNetworkGradient")
        return float(loss.value), g

    def train_step(self, records: np.ndarray, lr: float = 1e-3) -> float:
        r = np.ascontiguousarray(records, np.float32)
        loss = C.c_float()
        L.check(L.lib().mrtxGuideNetworkTrain(self._h, r.ctypes.data, r.shape[0], lr, C.byref(loss)),
                "mrtxGuideNetworkTrain")
        return float(loss.value)

    def train_step_no_alloc(self, records_ptr: int, count: int, lr: float = 1e-3, stream: int = 0,
                            want_loss: bool = True):
        """mrtxGuideNetworkTrainNoAlloc: one step on `count` device-resident
        records at address records_ptr (16 floats each), enqueued on the HIP
        stream handle `stream` (0 = the network's own stream).  Returns the
        loss (the call then waits for the stream) or None."""
        loss = C.c_float()
        L.check(L.lib().mrtxGuideNetworkTrainNoAlloc(self._h, C.c_void_p(records_ptr), count, lr,
                                                     C.c_void_p(stream) if stream else None,
                                                     C.byref(loss) if want_loss else None),
                "mrtxGuideNetworkTrainNoAlloc")
        return float(loss.value) if want_loss else None

    def sample(self, raw: np.ndarray, u: np.ndarray) -> np.ndarray:
        """(N, 4): sampled unit direction xyz and the mixture pdf of it."""
        raw = np.ascontiguousarray(raw, np.float32)
        u = np.ascontiguousarray(u, np.float32)
        out = np.empty((raw.shape[0], 4), np.float32)
        L.check(L.lib().mrtxGuideNetworkSample(self._h, raw.ctypes.data, u.ctypes.data, raw.shape[0],
                                               out.ctypes.data), "mrtxGuideNetworkSample")
        return out

    def pdf(self, raw: np.ndarray, dirs: np.ndarray) -> np.ndarray:
        raw = np.ascontiguousarray(raw, np.float32)
        d = np.zeros((raw.shape[0], 4), np.float32)
        d[:, :3] = dirs
        out = np.empty(raw.shape[0], np.float32)
        L.check(L.lib().mrtxGuideNetworkPdf(self._h, raw.ctypes.data, d.ctypes.data, raw.shape[0], out.ctypes.data),
                "mrtxGuideNetworkPdf")
        return out
