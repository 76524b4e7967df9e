"""Python front end of the learned K-step jump (include/ppf_unet.h).

``UNet`` owns one network on the device: BF16-MFMA convolutions with FP32
master weights, trained by the library's own backward and Adam kernels.
Fields are float32 arrays shaped [batch, nz, ny, nx, 3] (etaS, etaL, c).
"""
import ctypes

import numpy as np

from . import _lib
from .solver import PpfError

_fp = ctypes.POINTER(ctypes.c_float)
_SIG = {
    "ppf_unet_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_int,
                                       ctypes.c_uint64]),
    "ppf_unet_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "ppf_unet_nparams": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]),
    "ppf_unet_get_params": (ctypes.c_int, [ctypes.c_void_p, _fp]),
    "ppf_unet_set_params": (ctypes.c_int, [ctypes.c_void_p, _fp]),
    "ppf_unet_forward": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_int] * 4 + [_fp, _fp]),
    "ppf_unet_loss_grad": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_int] * 4 +
                           [_fp, _fp, _fp, _fp]),
    "ppf_unet_train_step": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_int] * 4 +
                            [_fp, _fp, ctypes.c_float, _fp]),
    "ppf_unet_apply": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "ppf_unet_last_forward": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                                             ctypes.POINTER(ctypes.c_double)]),
    "ppf_unet_set_conv_kernel": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "ppf_unet_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
}


def lib():
    L = _lib.load("libppf.so")
    if not getattr(L, "_ppf_unet_typed", False):
        for name, (res, args) in _SIG.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        L._ppf_unet_typed = True
    return L


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None and a.shape != shape:
        raise ValueError(f"array shape {a.shape} != {shape}")
    return a


def layer_shapes(base=32):
    """[(name, shape)] of the flat parameter vector, in order."""
    C = base
    geo = [(3, C), (C, C), (C, 2 * C), (2 * C, 2 * C), (2 * C, 4 * C), (4 * C, 4 * C),
           (6 * C, 2 * C), (2 * C, 2 * C), (3 * C, C), (C, C)]
    out = []
    for i, (ci, co) in enumerate(geo):
        out.append((f"w{i}", (27, co, ci)))
        out.append((f"b{i}", (co,)))
    out.append(("head_w", (3, C)))
    out.append(("head_b", (3,)))
    return out


def unflatten(flat, base=32):
    d, o = {}, 0
    for name, shp in layer_shapes(base):
        n = int(np.prod(shp))
        d[name] = np.asarray(flat[o:o + n]).reshape(shp)
        o += n
    assert o == len(flat)
    return d


class UNet:
    def __init__(self, base=32, device=0, seed=1):
        L = lib()
        h = ctypes.c_void_p()
        rc = L.ppf_unet_create(ctypes.byref(h), int(base), int(device), int(seed))
        if rc:
            raise PpfError(rc, L.ppf_unet_last_error(None).decode(errors="replace"))
        self._h = h
        self.base = base
        n = ctypes.c_int64()
        self._check(L.ppf_unet_nparams(h, ctypes.byref(n)))
        self.nparams = n.value

    def _check(self, rc):
        if rc:
            raise PpfError(rc, lib().ppf_unet_last_error(self._h).decode(errors="replace"))

    def close(self):
        if getattr(self, "_h", None):
            lib().ppf_unet_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def get_params(self):
        p = np.empty(self.nparams, np.float32)
        self._check(lib().ppf_unet_get_params(self._h, p.ctypes.data_as(_fp)))
        return p

    def set_params(self, p):
        p = _f32(p, (self.nparams,))
        self._check(lib().ppf_unet_set_params(self._h, p.ctypes.data_as(_fp)))

    def forward(self, x):
        x = _f32(x)
        B, D, H, W, c = x.shape
        assert c == 3
        y = np.empty_like(x)
        self._check(lib().ppf_unet_forward(self._h, B, D, H, W, x.ctypes.data_as(_fp),
                                           y.ctypes.data_as(_fp)))
        return y

    def loss_grad(self, x, y):
        x = _f32(x)
        y = _f32(y, x.shape)
        B, D, H, W, _ = x.shape
        loss = ctypes.c_float()
        g = np.empty(self.nparams, np.float32)
        self._check(lib().ppf_unet_loss_grad(self._h, B, D, H, W, x.ctypes.data_as(_fp),
                                             y.ctypes.data_as(_fp), ctypes.byref(loss),
                                             g.ctypes.data_as(_fp)))
        return loss.value, g

    def train_step(self, x, y, lr):
        x = _f32(x)
        y = _f32(y, x.shape)
        B, D, H, W, _ = x.shape
        loss = ctypes.c_float()
        self._check(lib().ppf_unet_train_step(self._h, B, D, H, W, x.ctypes.data_as(_fp),
                                              y.ctypes.data_as(_fp), float(lr),
                                              ctypes.byref(loss)))
        return loss.value

    def apply(self, solver):
        """Jump the solver's state in place (solute projected back)."""
        self._check(lib().ppf_unet_apply(self._h, solver._h))

    def set_conv_kernel(self, kind):
        """2: row-blocked convolutions; 3, 4: row-blocked with fixed tiles;
        5: kind 2 with materialized upsampling; 6 (default): kind 2 with the
        decoder convolutions in polyphase form at inference; 7: kind 6 with
        the row-blocked polyphase kernel only; 1: LDS-tiled; 0: the direct
        form."""
        self._check(lib().ppf_unet_set_conv_kernel(self._h, int(kind)))

    def last_forward(self):
        ms, fl = ctypes.c_double(), ctypes.c_double()
        self._check(lib().ppf_unet_last_forward(self._h, ctypes.byref(ms), ctypes.byref(fl)))
        return ms.value, fl.value
