"""ctypes front end of the C oracle ``oracle/liboracle.so``.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.

The C oracle restates phasefield.c:301-476 and :1112-1189 in the reference's
own layout and operand order; see oracle/pf_oracle.c for the line map.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BC_PERIODIC = 0
BC_NOFLUX = 1
NO_C_ONLY = 100000000  # phasefield.c:105 n_iter_init: the c-only sweep never starts

_dp = ctypes.POINTER(ctypes.c_double)


class OrcParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("MS", "ML", "AS", "AL", "cL0", "cS0", "CS", "CL", "L", "dt", "h2", "m", "gamma",
                 "kappa")]


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run `make -C oracle` (or __graft_entry__.build())")
        L = ctypes.CDLL(path)
        L.orc_run_dense.argtypes = [ctypes.c_int] * 3 + [_dp] * 3 + [ctypes.c_long, ctypes.c_int,
                                                                   ctypes.c_long, _dp]
        L.orc_run_dense_params.argtypes = [ctypes.c_int] * 3 + [_dp] * 3 + [
            ctypes.c_long, ctypes.c_int, ctypes.c_long, ctypes.POINTER(OrcParams), _dp]
        L.orc_increments_dense.argtypes = [ctypes.c_int] * 3 + [_dp] * 3 + [ctypes.c_int] * 2 + [
            _dp] * 3
        L.orc_time_dense.argtypes = [ctypes.c_int] * 3 + [_dp] * 3 + [ctypes.c_long, ctypes.c_int,
                                                                    _dp]
        L.orc_max_threads.restype = ctypes.c_int
        L.orc_params_default.argtypes = [ctypes.POINTER(OrcParams)]
        _lib = L
    return _lib


def default_params():
    p = OrcParams()
    lib().orc_params_default(ctypes.byref(p))
    return p


def _p(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def run(s, l, c, nsteps, bc=BC_PERIODIC, n_iter_init=NO_C_ONLY, params=None):
    """Step copies of (s, l, c) ``nsteps`` times.  ``l=None`` keeps etaL at 1.0
    (the reference's as-shipped reader).  Returns (s, l, c, sums) where sums
    are the save-time sums of the last increments (phasefield.c:228-238)."""
    s = np.array(s, dtype=np.float64, order="C", copy=True)
    c = np.array(c, dtype=np.float64, order="C", copy=True)
    if l is None:
        l = np.ones_like(s)
    else:
        l = np.array(l, dtype=np.float64, order="C", copy=True)
    nz, ny, nx = s.shape
    sums = np.zeros(3)
    if params is None:
        rc = lib().orc_run_dense(nz, ny, nx, _p(s), _p(l), _p(c), int(nsteps), int(bc),
                                 int(n_iter_init), _p(sums))
    else:
        rc = lib().orc_run_dense_params(nz, ny, nx, _p(s), _p(l), _p(c), int(nsteps), int(bc),
                                        int(n_iter_init), ctypes.byref(params), _p(sums))
    if rc != 0:
        raise RuntimeError("oracle run failed")
    return s, l, c, sums


def increments(s, l, c, bc=BC_PERIODIC, c_only=False):
    """(detaS, detaL, dc) of one sweep from state (s, l, c)."""
    s, l, c = (np.ascontiguousarray(a, dtype=np.float64) for a in (s, l, c))
    out = [np.zeros_like(s) for _ in range(3)]
    nz, ny, nx = s.shape
    rc = lib().orc_increments_dense(nz, ny, nx, _p(s), _p(l), _p(c), int(bc), int(bool(c_only)),
                                    *(_p(a) for a in out))
    if rc != 0:
        raise RuntimeError("oracle increments failed")
    return tuple(out)


def time_steps(s, l, c, nsteps, bc=BC_PERIODIC):
    """Wall seconds of ``nsteps`` oracle sweeps (stepping loop only) and the
    OpenMP thread count used."""
    s, l, c = (np.ascontiguousarray(a, dtype=np.float64) for a in (s, l, c))
    nz, ny, nx = s.shape
    sec = ctypes.c_double()
    rc = lib().orc_time_dense(nz, ny, nx, _p(s), _p(l), _p(c), int(nsteps), int(bc),
                              ctypes.byref(sec))
    if rc != 0:
        raise RuntimeError("oracle timing run failed")
    return sec.value, lib().orc_max_threads()
