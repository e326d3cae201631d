"""Python mirror of the radiance-cache entry points of include/bbgpu.h."""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_float, c_int32, c_uint32, c_void_p

import numpy as np

from . import capi

RAW_STRIDE = 16  # floats per input row


class BbNrcDesc(ctypes.Structure):
    _fields_ = [("hidden_width", c_int32), ("hidden_layers", c_int32), ("learning_rate", c_float), ("seed", c_uint32)]


def _declare(lib):
    if getattr(lib, "_nrc_declared", False):
        return lib
    lib.bbgpu_nrc_param_count.restype = c_int32
    lib.bbgpu_nrc_param_count.argtypes = [POINTER(BbNrcDesc)]
    lib.bbgpu_nrc_create.argtypes = [POINTER(BbNrcDesc), POINTER(c_void_p)]
    lib.bbgpu_nrc_destroy.argtypes = [c_void_p]
    lib.bbgpu_nrc_get_weights.argtypes = [c_void_p, POINTER(c_float), c_int32]
    lib.bbgpu_nrc_set_weights.argtypes = [c_void_p, POINTER(c_float), c_int32]
    lib.bbgpu_nrc_query.argtypes = [c_void_p, c_void_p, c_int32, c_void_p, c_void_p]
    lib.bbgpu_nrc_train.argtypes = [c_void_p, c_void_p, c_void_p, c_int32, c_int32, POINTER(c_float), POINTER(c_float),
                                    c_void_p]
    lib.bbgpu_nrc_diverged.argtypes = [c_void_p, POINTER(c_int32)]
    for name in ("bbgpu_nrc_create", "bbgpu_nrc_destroy", "bbgpu_nrc_get_weights", "bbgpu_nrc_set_weights",
                 "bbgpu_nrc_query", "bbgpu_nrc_train", "bbgpu_nrc_diverged"):
        getattr(lib, name).restype = c_int32
    lib._nrc_declared = True
    return lib


class Nrc:
    """The cache on the current HIP device (torch buffers for device data)."""

    def __init__(self, seed=1, learning_rate=1e-3, width=64, layers=4):
        import torch

        self.torch = torch
        self.lib = _declare(capi.load_library())
        d = BbNrcDesc(width, layers, learning_rate, seed)
        h = c_void_p()
        capi.check(self.lib.bbgpu_nrc_create(ctypes.byref(d), ctypes.byref(h)), "bbgpu_nrc_create", self.lib)
        self.h = h
        self.width, self.layers = width, layers
        self.n_params = self.lib.bbgpu_nrc_param_count(ctypes.byref(d))

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.lib.bbgpu_nrc_destroy(self.h)
            self.h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def weights(self) -> np.ndarray:
        w = np.zeros(self.n_params, np.float32)
        capi.check(self.lib.bbgpu_nrc_get_weights(self.h, w.ctypes.data_as(POINTER(c_float)), self.n_params),
                   "bbgpu_nrc_get_weights", self.lib)
        return w

    def set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        capi.check(self.lib.bbgpu_nrc_set_weights(self.h, w.ctypes.data_as(POINTER(c_float)), self.n_params),
                   "bbgpu_nrc_set_weights", self.lib)

    def diverged(self) -> bool:
        """True once a training step or a cached render found the network diverged."""
        d = c_int32()
        capi.check(self.lib.bbgpu_nrc_diverged(self.h, ctypes.byref(d)), "bbgpu_nrc_diverged", self.lib)
        return bool(d.value)

    def _dev(self, a, cols):
        a = np.ascontiguousarray(a, np.float32).reshape(-1, cols)
        return self.torch.from_numpy(a).to("cuda")

    def query(self, raw) -> np.ndarray:
        r = self._dev(raw, RAW_STRIDE)
        out = self.torch.zeros((len(r), 3), dtype=self.torch.float32, device="cuda")
        capi.check(self.lib.bbgpu_nrc_query(self.h, r.data_ptr(), len(r), out.data_ptr(), None), "bbgpu_nrc_query",
                   self.lib)
        return out.cpu().numpy()

    def train(self, raw, target, apply=True, want_grad=False):
        r = self._dev(raw, RAW_STRIDE)
        t = self._dev(target, 3)
        loss = c_float()
        g = np.zeros(self.n_params, np.float32) if want_grad else None
        capi.check(self.lib.bbgpu_nrc_train(self.h, r.data_ptr(), t.data_ptr(), len(r), int(apply), ctypes.byref(loss),
                                            g.ctypes.data_as(POINTER(c_float)) if want_grad else None, None),
                   "bbgpu_nrc_train", self.lib)
        return (loss.value, g) if want_grad else loss.value
