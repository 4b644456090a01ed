"""The swap-ranking surrogate network (include/gg_mlp.h): a small MLP whose
forward pass, backward pass and Adam updates run on the GPU with BF16 MFMA
GEMMs (FP32 accumulate)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L

_f32p = ctypes.POINTER(ctypes.c_float)


def _sigs():
    lib = L.lib()
    if getattr(lib, "_mlp_sigs", False):
        return lib
    for name, args, res in [
        ("gg_mlp_create", [ctypes.c_int, L.i32p, ctypes.c_uint64], ctypes.c_void_p),
        ("gg_mlp_destroy", [ctypes.c_void_p], None),
        ("gg_mlp_param_count", [ctypes.c_void_p], ctypes.c_int),
        ("gg_mlp_layout", [ctypes.c_void_p, L.i32p, L.i32p, L.i32p, L.i32p], ctypes.c_int),
        ("gg_mlp_get_params", [ctypes.c_void_p, _f32p], ctypes.c_int),
        ("gg_mlp_set_params", [ctypes.c_void_p, _f32p], ctypes.c_int),
        ("gg_mlp_forward", [ctypes.c_void_p, _f32p, ctypes.c_int, _f32p], ctypes.c_int),
        ("gg_mlp_gradient", [ctypes.c_void_p, _f32p, _f32p, ctypes.c_int, _f32p, _f32p],
         ctypes.c_int),
        ("gg_mlp_train_step", [ctypes.c_void_p, _f32p, _f32p, ctypes.c_int, ctypes.c_float,
                               _f32p], ctypes.c_int),
    ]:
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    lib._mlp_sigs = True
    return lib


def _fp(a):
    return a.ctypes.data_as(_f32p)


class Mlp:
    def __init__(self, dims=(64, 256, 256, 256, 1), seed: int = 1):
        lib = _sigs()
        d = np.asarray(dims, dtype=np.int32)
        p = lib.gg_mlp_create(len(dims) - 1, L.ptr(d), seed)
        if not p:
            raise L.GGError(-1, (lib.gg_last_error() or b"").decode())
        self._p = ctypes.c_void_p(p)
        self.dims = tuple(dims)
        self.n_params = lib.gg_mlp_param_count(self._p)
        nl = len(dims) - 1
        arrs = [np.zeros(nl, dtype=np.int32) for _ in range(4)]
        lib.gg_mlp_layout(self._p, *[L.ptr(a) for a in arrs])
        self.layout = tuple(a.tolist() for a in arrs)

    def params(self) -> np.ndarray:
        out = np.zeros(self.n_params, dtype=np.float32)
        L.check(_sigs().gg_mlp_get_params(self._p, _fp(out)))
        return out

    def set_params(self, p) -> None:
        a = np.ascontiguousarray(p, dtype=np.float32)
        L.check(_sigs().gg_mlp_set_params(self._p, _fp(a)))

    def forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros(x.shape[0], dtype=np.float32)
        L.check(_sigs().gg_mlp_forward(self._p, _fp(x), x.shape[0], _fp(y)))
        return y

    def gradient(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(self.n_params, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        L.check(_sigs().gg_mlp_gradient(self._p, _fp(x), _fp(t), x.shape[0], _fp(loss), _fp(g)))
        return float(loss[0]), g

    def train_step(self, x, t, lr: float = 1e-3) -> float:
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        L.check(_sigs().gg_mlp_train_step(self._p, _fp(x), _fp(t), x.shape[0], lr, _fp(loss)))
        return float(loss[0])

    def close(self):
        if getattr(self, "_p", None):
            _sigs().gg_mlp_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
