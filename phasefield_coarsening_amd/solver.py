"""Python front end of ``libppf.so`` (the C ABI of include/ppf.h).

``Solver`` holds one device context: the fields (etaS, etaL, c) resident in
HBM, stepped by the gfx950 kernels.  It mirrors the reference's stepping
loop (phasefield.c:178-258): ``step`` is ghost fill + ``sweep_fields``
(phasefield.c:186-215), ``step_c_only`` is ghost fill + ``sweep_fields_c``
(phasefield.c:217), and ``step_report`` also returns the three save-time sums
of phasefield.c:228-238.

Every call goes through the native library; there is no CPU fallback.  A
failure raises ``PpfError`` with the library's message.
"""
import ctypes

import numpy as np

from . import _lib

BC_PERIODIC = 0
BC_NOFLUX = 1
F64 = 0
F32 = 1
KERNEL_AUTO = 0
KERNEL_NAIVE = 1
KERNEL_TILED = 2
KERNEL_FAST = 3  # FP64, fused multiply-adds and the collapsed AS == AL chemistry

_dp = ctypes.POINTER(ctypes.c_double)


class Params(ctypes.Structure):
    """Arguments of sweep_fields (phasefield.c:309-314), in that order."""
    _fields_ = [(n, ctypes.c_double) for n in
                ("MS", "ML", "AS", "AL", "cL0", "cS0", "CS", "CL", "L", "dt", "h2", "m", "gamma",
                 "kappa")]


class Desc(ctypes.Structure):
    _fields_ = [("nz", ctypes.c_int), ("ny", ctypes.c_int), ("nx", ctypes.c_int),
                ("bc", ctypes.c_int), ("precision", ctypes.c_int), ("kernel", ctypes.c_int),
                ("device", ctypes.c_int), ("reserved", ctypes.c_int), ("params", Params)]


class Stats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("sum_c", "sum_h", "n_solid", "sum_s", "min_c", "max_c", "n_sites",
                 "n_nonfinite", "sum_h1h")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class ParticleStats(ctypes.Structure):
    _fields_ = [("n_particles", ctypes.c_int64), ("solid_sites", ctypes.c_double),
                ("mean_radius", ctypes.c_double), ("max_radius", ctypes.c_double),
                ("mean_volume", ctypes.c_double)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class IncSums(ctypes.Structure):
    _fields_ = [("sum_dc2", ctypes.c_double), ("sum_ds2", ctypes.c_double),
                ("sum_dl2", ctypes.c_double)]


class Diff(ctypes.Structure):
    _fields_ = [("max_abs", ctypes.c_double * 3), ("rms", ctypes.c_double * 3)]


class PpfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libppf error {code}: {msg}")
        self.code = code


# Every symbol include/ppf.h declares, with its ctypes signature.
SIGNATURES = {
    "ppf_params_default": (ctypes.c_int, [ctypes.POINTER(Params)]),
    "ppf_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(Desc)]),
    "ppf_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "ppf_upload": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp]),
    "ppf_download": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp]),
    "ppf_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64]),
    "ppf_step_c_only": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64]),
    "ppf_step_report": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(IncSums)]),
    "ppf_reduce": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Stats)]),
    "ppf_init_spheres": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp, _dp, ctypes.c_double,
                                        ctypes.c_double]),
    "ppf_spheres_generate": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                            ctypes.c_double, _dp, _dp]),
    "ppf_set_tiling": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_int] * 4),
    "ppf_get_tiling": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_int)] * 4),
    "ppf_set_fusion": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_int] * 5),
    "ppf_get_fusion": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_int)] * 5),
    "ppf_particles": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double,
                                     ctypes.POINTER(ParticleStats),
                                     ctypes.POINTER(ctypes.c_int64), ctypes.c_int64]),
    "ppf_set_mass_correction": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "ppf_mass_target": (ctypes.c_int, [ctypes.c_void_p, _dp, ctypes.POINTER(ctypes.c_int64)]),
    "ppf_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp]),
    "ppf_copy_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "ppf_compare": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(Diff)]),
    "ppf_timer_start": (ctypes.c_int, [ctypes.c_void_p]),
    "ppf_timer_stop": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "ppf_sync": (ctypes.c_int, [ctypes.c_void_p]),
    "ppf_steps_taken": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]),
    "ppf_kernel_name": (ctypes.c_char_p, [ctypes.c_void_p]),
    "ppf_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "ppf_sweep_report_noalloc": (ctypes.c_int, [ctypes.c_void_p] * 6 + [ctypes.c_int] * 5 +
                                 [ctypes.POINTER(Params), ctypes.c_int,
                                  ctypes.POINTER(IncSums)]),
    "ppf_ens_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, ctypes.POINTER(Params)]),
    "ppf_ens_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "ppf_ens_upload": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp]),
    "ppf_ens_download": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp]),
    "ppf_ens_download_f32": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int] +
                             [ctypes.POINTER(ctypes.c_float)] * 3),
    "ppf_ens_init_spheres": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, _dp,
                                            ctypes.c_double, ctypes.c_double]),
    "ppf_ens_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64]),
    "ppf_ens_step_c_only": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64]),
    "ppf_ens_set_tiling": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_int] * 3),
    "ppf_ens_get_tiling": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int] +
                           [ctypes.POINTER(ctypes.c_int)] * 2),
    "ppf_ens_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "ppf_ens_set_fusion": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "ppf_ens_get_fusion": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]),
    "ppf_ens_set_tiling2": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_int] * 3),
    "ppf_ens_get_tiling2": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int] +
                            [ctypes.POINTER(ctypes.c_int)] * 2),
    "ppf_ens_reduce": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(Stats)]),
    "ppf_ens_set_mass_correction": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "ppf_ens_mass_target": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp]),
    "ppf_ens_timer_start": (ctypes.c_int, [ctypes.c_void_p]),
    "ppf_ens_timer_stop": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "ppf_ens_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64),
                                    ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    "ppf_ens_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
}


def lib():
    L = _lib.load("libppf.so")
    if not getattr(L, "_ppf_typed", False):
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        L._ppf_typed = True
    return L


def default_params():
    p = Params()
    rc = lib().ppf_params_default(ctypes.byref(p))
    if rc:
        raise PpfError(rc, "ppf_params_default")
    return p


def spheres(seed, n, shape, rmin, rmax):
    """Seeded sphere list: (centers [n,3] as (z, y, x), radii [n]).  Host only."""
    nz, ny, nx = shape
    cen = np.zeros((max(n, 1), 3))
    rad = np.zeros(max(n, 1))
    rc = lib().ppf_spheres_generate(int(seed), int(n), nz, ny, nx, float(rmin), float(rmax),
                                    cen.ctypes.data_as(_dp), rad.ctypes.data_as(_dp))
    if rc:
        raise PpfError(rc, "ppf_spheres_generate: bad arguments")
    return cen[:n], rad[:n]


def _ptr(a):
    return a.ctypes.data_as(_dp)


class Solver:
    """One device-resident grid.  shape = (nz, ny, nx), x fastest."""

    def __init__(self, shape, bc=BC_PERIODIC, precision=F64, kernel=KERNEL_AUTO, device=0,
                 params=None):
        L = lib()
        d = Desc()
        d.nz, d.ny, d.nx = (int(v) for v in shape)
        d.bc, d.precision, d.kernel, d.device = int(bc), int(precision), int(kernel), int(device)
        d.params = params if params is not None else default_params()
        h = ctypes.c_void_p()
        rc = L.ppf_create(ctypes.byref(h), ctypes.byref(d))
        if rc:
            raise PpfError(rc, L.ppf_last_error(None).decode(errors="replace"))
        self._h = h
        self.shape = (d.nz, d.ny, d.nx)
        self.bc = bc
        self.precision = precision

    def _check(self, rc):
        if rc:
            raise PpfError(rc, lib().ppf_last_error(self._h).decode(errors="replace"))

    def close(self):
        if getattr(self, "_h", None):
            lib().ppf_destroy(self._h)
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

    @property
    def kernel_name(self):
        return lib().ppf_kernel_name(self._h).decode()

    def upload(self, s, l, c):
        """Set the state.  l=None leaves etaL = 1.0 everywhere (the state the
        reference's reader produces, phasefield.c:1032)."""
        arrs = []
        for a in (s, l, c):
            if a is None:
                arrs.append(None)
                continue
            a = np.ascontiguousarray(a, dtype=np.float64)
            if a.shape != self.shape:
                raise ValueError(f"field shape {a.shape} != grid {self.shape}")
            arrs.append(a)
        self._check(lib().ppf_upload(self._h, _ptr(arrs[0]),
                                     _ptr(arrs[1]) if arrs[1] is not None else None,
                                     _ptr(arrs[2])))

    def download(self):
        out = [np.empty(self.shape, np.float64) for _ in range(3)]
        self._check(lib().ppf_download(self._h, *(_ptr(a) for a in out)))
        return tuple(out)

    def init_spheres(self, centers, radii, c_liq=0.1, c_span=0.8):
        centers = np.ascontiguousarray(centers, dtype=np.float64).reshape(-1, 3)
        radii = np.ascontiguousarray(radii, dtype=np.float64).reshape(-1)
        if len(centers) != len(radii):
            raise ValueError("centers and radii differ in length")
        self._check(lib().ppf_init_spheres(self._h, len(radii), _ptr(centers), _ptr(radii),
                                           float(c_liq), float(c_span)))

    def step(self, nsteps=1):
        self._check(lib().ppf_step(self._h, int(nsteps)))

    def step_c_only(self, nsteps=1):
        self._check(lib().ppf_step_c_only(self._h, int(nsteps)))

    def step_report(self, c_only=False):
        r = IncSums()
        self._check(lib().ppf_step_report(self._h, int(bool(c_only)), ctypes.byref(r)))
        return r.sum_dc2, r.sum_ds2, r.sum_dl2

    def reduce(self):
        st = Stats()
        self._check(lib().ppf_reduce(self._h, ctypes.byref(st)))
        return st.as_dict()

    def sync(self):
        self._check(lib().ppf_sync(self._h))

    def set_tiling(self, tx, ty, tz, order=0):
        self._check(lib().ppf_set_tiling(self._h, int(tx), int(ty), int(tz), int(order)))

    def autotune(self, reps=3):
        """Pick the fastest tiling for this grid on this device; the state is
        not advanced.  Returns (tiling, ms per sweep)."""
        ms = ctypes.c_double()
        self._check(lib().ppf_autotune(self._h, int(reps), ctypes.byref(ms)))
        return self.get_tiling(), ms.value

    def get_tiling(self):
        """(tx, ty, tz, order)"""
        v = [ctypes.c_int() for _ in range(4)]
        self._check(lib().ppf_get_tiling(self._h, *(ctypes.byref(a) for a in v)))
        return tuple(a.value for a in v)

    def set_fusion(self, steps_per_launch, tx=0, ty=0, tz=0, order=-1):
        """Two steps per launch (bitwise equal to single steps; on by
        default): steps_per_launch 1 or 2, with that kernel's tiling ((0, 0,
        0, -1) keeps the current one)."""
        self._check(lib().ppf_set_fusion(self._h, int(steps_per_launch), int(tx), int(ty),
                                         int(tz), int(order)))

    def get_fusion(self):
        """(steps per launch in effect, tx, ty, tz, order)"""
        v = [ctypes.c_int() for _ in range(5)]
        self._check(lib().ppf_get_fusion(self._h, *(ctypes.byref(a) for a in v)))
        return tuple(a.value for a in v)

    def particles(self, threshold=0.5, max_sizes=1 << 20):
        """Connected solid particles {etaS > threshold}: (stats dict, sizes in
        sites, largest first)."""
        st = ParticleStats()
        sizes = np.zeros(max(int(max_sizes), 1), np.int64)
        self._check(lib().ppf_particles(self._h, float(threshold), ctypes.byref(st),
                                        sizes.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                        int(max_sizes)))
        d = st.as_dict()
        return d, sizes[:min(d["n_particles"], int(max_sizes))]

    def copy_from(self, other):
        """Device-to-device copy of other's state into this solver (converted
        when the precisions differ; an FP32 solver copying an FP64 one takes
        the FP64 total of c as its solute target)."""
        self._check(lib().ppf_copy_state(self._h, other._h))

    def compare(self, other):
        """Per-field differences to other's state, on the device:
        {"max_abs": (etaS, etaL, c), "rms": (etaS, etaL, c)}."""
        d = Diff()
        self._check(lib().ppf_compare(self._h, other._h, ctypes.byref(d)))
        return {"max_abs": tuple(d.max_abs), "rms": tuple(d.rms)}

    def set_mass_correction(self, every):
        """FP32 only: shift c back onto the FP64 solute target every `every`
        steps (0 = never)."""
        self._check(lib().ppf_set_mass_correction(self._h, int(every)))

    def mass_target(self):
        """FP32 only: (FP64 target of sum(c), number of shifts applied)."""
        t, n = ctypes.c_double(), ctypes.c_int64()
        self._check(lib().ppf_mass_target(self._h, ctypes.byref(t), ctypes.byref(n)))
        return t.value, n.value

    def timer_start(self):
        self._check(lib().ppf_timer_start(self._h))

    def timer_stop(self):
        ms = ctypes.c_double()
        self._check(lib().ppf_timer_stop(self._h, ctypes.byref(ms)))
        return ms.value

    @property
    def steps_taken(self):
        v = ctypes.c_int64()
        self._check(lib().ppf_steps_taken(self._h, ctypes.byref(v)))
        return v.value


class Ensemble:
    """Many independent grids stepped together by one persistent launch per
    step.  dims: list of (nz, ny, nx)."""

    def __init__(self, dims, bc=BC_PERIODIC, precision=F32, device=0, params=None):
        L = lib()
        dims = [tuple(int(v) for v in d) for d in dims]
        arr = (ctypes.c_int * (3 * len(dims)))(*[v for d in dims for v in d])
        p = params if params is not None else default_params()
        h = ctypes.c_void_p()
        rc = L.ppf_ens_create(ctypes.byref(h), len(dims), arr, int(bc), int(precision),
                              int(device), ctypes.byref(p))
        if rc:
            raise PpfError(rc, L.ppf_ens_last_error(None).decode(errors="replace"))
        self._h = h
        self.dims = dims
        self.bc = bc
        self.precision = precision

    def _check(self, rc):
        if rc:
            raise PpfError(rc, lib().ppf_ens_last_error(self._h).decode(errors="replace"))

    def close(self):
        if getattr(self, "_h", None):
            lib().ppf_ens_destroy(self._h)
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

    def upload(self, dom, s, l, c):
        a = [None if x is None else np.ascontiguousarray(x, dtype=np.float64) for x in (s, l, c)]
        for x in a:
            if x is not None and x.shape != self.dims[dom]:
                raise ValueError(f"member {dom}: shape {x.shape} != {self.dims[dom]}")
        self._check(lib().ppf_ens_upload(self._h, int(dom), _ptr(a[0]),
                                         _ptr(a[1]) if a[1] is not None else None, _ptr(a[2])))

    def download(self, dom, f32=False):
        if f32:
            out = [np.empty(self.dims[dom], np.float32) for _ in range(3)]
            fp = ctypes.POINTER(ctypes.c_float)
            self._check(lib().ppf_ens_download_f32(self._h, int(dom),
                                                   *(a.ctypes.data_as(fp) for a in out)))
        else:
            out = [np.empty(self.dims[dom], np.float64) for _ in range(3)]
            self._check(lib().ppf_ens_download(self._h, int(dom), *(_ptr(a) for a in out)))
        return tuple(out)

    def init_spheres(self, dom, centers, radii, c_liq=0.1, c_span=0.8):
        centers = np.ascontiguousarray(centers, dtype=np.float64).reshape(-1, 3)
        radii = np.ascontiguousarray(radii, dtype=np.float64).reshape(-1)
        self._check(lib().ppf_ens_init_spheres(self._h, int(dom), len(radii), _ptr(centers),
                                               _ptr(radii), float(c_liq), float(c_span)))

    def step(self, nsteps=1):
        self._check(lib().ppf_ens_step(self._h, int(nsteps)))

    def step_c_only(self, nsteps=1):
        self._check(lib().ppf_ens_step_c_only(self._h, int(nsteps)))

    def set_tiling(self, dom, shape, tz):
        self._check(lib().ppf_ens_set_tiling(self._h, int(dom), int(shape), int(tz)))

    def get_tiling(self, dom):
        a, b = ctypes.c_int(), ctypes.c_int()
        self._check(lib().ppf_ens_get_tiling(self._h, int(dom), ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def autotune(self, reps=3):
        self._check(lib().ppf_ens_autotune(self._h, int(reps)))

    def set_fusion(self, steps_per_launch):
        """Two steps per launch (bitwise equal to single steps): 1 or 2."""
        self._check(lib().ppf_ens_set_fusion(self._h, int(steps_per_launch)))

    def get_fusion(self):
        v = ctypes.c_int()
        self._check(lib().ppf_ens_get_fusion(self._h, ctypes.byref(v)))
        return v.value

    def set_tiling2(self, dom, shape, tz):
        """Two-step member tiling: shape 0 = 32x8, 1 = 16x16."""
        self._check(lib().ppf_ens_set_tiling2(self._h, int(dom), int(shape), int(tz)))

    def get_tiling2(self, dom):
        a, b = ctypes.c_int(), ctypes.c_int()
        self._check(lib().ppf_ens_get_tiling2(self._h, int(dom), ctypes.byref(a),
                                              ctypes.byref(b)))
        return a.value, b.value

    def reduce(self, dom):
        st = Stats()
        self._check(lib().ppf_ens_reduce(self._h, int(dom), ctypes.byref(st)))
        return st.as_dict()

    def set_mass_correction(self, every):
        self._check(lib().ppf_ens_set_mass_correction(self._h, int(every)))

    def mass_target(self, dom):
        t = ctypes.c_double()
        self._check(lib().ppf_ens_mass_target(self._h, int(dom), ctypes.byref(t)))
        return t.value

    def timer_start(self):
        self._check(lib().ppf_ens_timer_start(self._h))

    def timer_stop(self):
        ms = ctypes.c_double()
        self._check(lib().ppf_ens_timer_stop(self._h, ctypes.byref(ms)))
        return ms.value

    def info(self):
        n, it, g = ctypes.c_int64(), ctypes.c_int(), ctypes.c_int()
        self._check(lib().ppf_ens_info(self._h, ctypes.byref(n), ctypes.byref(it), ctypes.byref(g)))
        return {"total_sites": n.value, "items": it.value, "grid": g.value}


def sweep_report_noalloc(s_in, l_in, c_in, s_out, l_out, c_out, shape, bc=BC_PERIODIC,
                         c_only=False, params=None, device=0):
    """ppf_sweep_report_noalloc on device addresses (ints, e.g. a torch
    tensor's data_ptr()): one sweep of an FP64 grid in caller-owned device
    memory with its save-time sums, allocating nothing.  Returns (sum_dc2,
    sum_ds2, sum_dl2)."""
    nz, ny, nx = (int(v) for v in shape)
    p = params if params is not None else default_params()
    r = IncSums()
    L = lib()
    rc = L.ppf_sweep_report_noalloc(s_in, l_in, c_in, s_out, l_out, c_out, nz, ny, nx, int(bc),
                                    int(bool(c_only)), ctypes.byref(p), int(device),
                                    ctypes.byref(r))
    if rc:
        raise PpfError(rc, L.ppf_last_error(None).decode(errors="replace"))
    return r.sum_dc2, r.sum_ds2, r.sum_dl2
