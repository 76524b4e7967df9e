"""HDF5 file contract of the stepping program, over ``libppfio.so``.

Mirrors the reference's serial reader/writer (phasefield.c:996-1109) and its
output naming (phasefield.c:117-121, 1193-1206).  Arrays are float64, C order
``[nz, ny, nx]`` (x fastest).
"""
import ctypes

import numpy as np

from . import _lib

ETAL_COMPAT = 1

_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_dbl_p = ctypes.POINTER(ctypes.c_double)


class PpfIOError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ppfio error {code}: {msg}")
        self.code = code


def _lib_io():
    lib = _lib.load("libppfio.so")
    if not getattr(lib, "_ppf_typed", False):
        lib.ppfio_dims.argtypes = [ctypes.c_char_p, _c_int_p]
        lib.ppfio_read.argtypes = [ctypes.c_char_p, _c_int_p, _c_dbl_p, _c_dbl_p, _c_dbl_p,
                                   ctypes.c_int]
        lib.ppfio_write.argtypes = [ctypes.c_char_p, _c_int_p, _c_dbl_p, _c_dbl_p, _c_dbl_p]
        lib.ppfio_write_chunked.argtypes = [ctypes.c_char_p, _c_int_p, _c_dbl_p, _c_dbl_p,
                                            _c_dbl_p, _c_int_p]
        lib.ppfio_layout.argtypes = [ctypes.c_char_p, ctypes.c_char_p, _c_int_p]
        lib.ppfio_job_stem.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
        lib.ppfio_output_name.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p,
                                          ctypes.c_int]
        lib.ppfio_last_error.restype = ctypes.c_char_p
        lib._ppf_typed = True
    return lib


def _check(lib, rc):
    if rc != 0:
        raise PpfIOError(rc, lib.ppfio_last_error().decode(errors="replace"))


def _dptr(a):
    return a.ctypes.data_as(_c_dbl_p)


def read(path, etal_compat=False):
    """Return (etaS, etaL, c) from ``path``; dims come from /etaS."""
    lib = _lib_io()
    dims = (ctypes.c_int * 3)()
    _check(lib, lib.ppfio_dims(str(path).encode(), dims))
    shape = tuple(int(d) for d in dims)
    s, l, c = (np.empty(shape, np.float64) for _ in range(3))
    _check(lib, lib.ppfio_read(str(path).encode(), dims, _dptr(s), _dptr(l), _dptr(c),
                               ETAL_COMPAT if etal_compat else 0))
    return s, l, c


def write(path, s, l, c, chunk=None):
    """Write /etaS, /etaL, /c; ``chunk`` (z, y, x) gives the chunked layout of
    the reference's parallel writer (phasefield.c:576-582), default
    contiguous as its serial one."""
    lib = _lib_io()
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (s, l, c)]
    if not (arrs[0].shape == arrs[1].shape == arrs[2].shape) or arrs[0].ndim != 3:
        raise ValueError("etaS, etaL, c must be 3-D arrays of one shape")
    dims = (ctypes.c_int * 3)(*arrs[0].shape)
    if chunk is None:
        _check(lib, lib.ppfio_write(str(path).encode(), dims, *(_dptr(a) for a in arrs)))
    else:
        cd = (ctypes.c_int * 3)(*(int(v) for v in chunk))
        _check(lib, lib.ppfio_write_chunked(str(path).encode(), dims,
                                            *(_dptr(a) for a in arrs), cd))


def layout(path, dataset="etaS"):
    """Chunk dims (z, y, x) of a dataset, or (0, 0, 0) when contiguous."""
    lib = _lib_io()
    cd = (ctypes.c_int * 3)()
    _check(lib, lib.ppfio_layout(str(path).encode(), dataset.encode(), cd))
    return tuple(int(v) for v in cd)


def job_stem(input_name):
    lib = _lib_io()
    buf = ctypes.create_string_buffer(4096)
    _check(lib, lib.ppfio_job_stem(input_name.encode(), buf, 4096))
    return buf.value.decode()


def output_name(stem, k):
    lib = _lib_io()
    buf = ctypes.create_string_buffer(4096)
    _check(lib, lib.ppfio_output_name(stem.encode(), int(k), buf, 4096))
    return buf.value.decode()
