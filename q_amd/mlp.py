"""ctypes binding of the MLP entry points of libqmi.so (include/qmi.h):
the overlap predictor of the learned lambda schedule, trained and evaluated
on the GPU by the library's own kernels."""
from __future__ import annotations

import ctypes as C

import numpy as np

from q_amd.engine import QmiError, load_library

_FP = C.POINTER(C.c_float)


def _bind(lib):
    if getattr(lib, "_mlp_bound", False):
        return lib
    p = C.c_void_p
    sig = {
        "qmi_mlp_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_ulonglong, C.POINTER(p)]),
        "qmi_mlp_num_params": (C.c_longlong, [p]),
        "qmi_mlp_get_params": (C.c_int, [p, _FP, C.c_longlong]),
        "qmi_mlp_set_params": (C.c_int, [p, _FP, C.c_longlong]),
        "qmi_mlp_forward": (C.c_int, [p, _FP, C.c_int, _FP]),
        "qmi_mlp_forward_noalloc": (C.c_int, [p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
        "qmi_mlp_grad": (C.c_int, [p, _FP, _FP, C.c_int, _FP, _FP]),
        "qmi_mlp_train": (C.c_int, [p, _FP, _FP, C.c_int, C.c_int, C.c_int, C.c_float, C.c_ulonglong, _FP]),
        "qmi_mlp_last_error": (C.c_char_p, []),
        "qmi_mlp_destroy": (None, [p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    lib._mlp_bound = True
    return lib


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(_FP)


class MLP:
    """n_in -> hidden (SiLU) -> hidden (SiLU) -> 1 (sigmoid) on HIP device `device`."""

    def __init__(self, n_in: int, hidden: int = 128, seed: int = 0, device: int = 0):
        self.lib = _bind(load_library())
        self._h = C.c_void_p()
        self._check(self.lib.qmi_mlp_create(device, n_in, hidden, seed, C.byref(self._h)))
        self.n_in, self.hidden = n_in, hidden
        self.n_params = int(self.lib.qmi_mlp_num_params(self._h))

    def _check(self, rc):
        if rc < 0:
            raise QmiError(self.lib.qmi_mlp_last_error().decode())
        return rc

    def close(self):
        if self._h:
            self.lib.qmi_mlp_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self) -> np.ndarray:
        out = np.zeros(self.n_params, np.float32)
        self._check(self.lib.qmi_mlp_get_params(self._h, out.ctypes.data_as(_FP), self.n_params))
        return out

    @params.setter
    def params(self, value):
        a, pa = _f32(value)
        self._check(self.lib.qmi_mlp_set_params(self._h, pa, a.size))

    def split(self, flat=None):
        """(W1, b1, W2, b2, W3, b3) views of a flat parameter vector"""
        p = self.params if flat is None else flat
        H, I = self.hidden, self.n_in
        k = 0
        out = []
        for shape in ((H, I), (H,), (H, H), (H,), (1, H), (1,)):
            n = int(np.prod(shape))
            out.append(p[k:k + n].reshape(shape))
            k += n
        return out

    def forward(self, x) -> np.ndarray:
        x, px = _f32(x)
        x = x.reshape(-1, self.n_in)
        y = np.zeros(len(x), np.float32)
        self._check(self.lib.qmi_mlp_forward(self._h, px, len(x), y.ctypes.data_as(_FP)))
        return y

    def forward_device(self, x_ptr: int, batch: int, y_ptr: int, stream: int = 0) -> None:
        """qmi_mlp_forward_noalloc: the forward pass on device buffers (device
        addresses x_ptr [batch x n_in], y_ptr [batch]) queued on `stream`
        (0: the model's stream), without allocating and without waiting."""
        self._check(self.lib.qmi_mlp_forward_noalloc(self._h, C.c_void_p(x_ptr), batch, C.c_void_p(y_ptr),
                                                     C.c_void_p(stream) if stream else None))

    def grad(self, x, t):
        x, px = _f32(x)
        t, pt = _f32(t)
        g = np.zeros(self.n_params, np.float32)
        loss = C.c_float()
        self._check(self.lib.qmi_mlp_grad(self._h, px, pt, len(t), C.byref(loss), g.ctypes.data_as(_FP)))
        return float(loss.value), g

    def train(self, x, t, epochs=50, batch=256, lr=1e-3, seed=0) -> float:
        x, px = _f32(x)
        t, pt = _f32(t)
        loss = C.c_float()
        self._check(self.lib.qmi_mlp_train(self._h, px, pt, len(t), epochs, batch, lr, seed, C.byref(loss)))
        return float(loss.value)
