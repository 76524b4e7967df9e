"""FP32 form against the FP64 path on one workload: fields, solid fraction,
total solute and the particle-size distribution.

    python -m phasefield_coarsening_amd.precision_report [--grid 1024]
        [--spheres 12000] [--steps 2000] [--shift-every 50] [--report-at 50,100,...]

Both runs start from the same state (the sphere initial condition built in
FP64, rounded into the FP32 context on the device), step on the device, and are compared
at the steps listed in ``--report-at``.  Particles are the connected components of
{etaS > 0.5}; their equivalent radii (3V/(4 pi))^(1/3) are binned in unit
bins.  One JSON line per report.

(With R ~ U[6, 14], 12000 spheres fill about 5% of a 1024^3 box.)

``--c-init`` picks the initial composition:
- ``config`` (default): c = 0.1 + 0.8 etaS, the configured construction.
  Both phases start far from equilibrium; the mean composition lies below
  the solid's equilibrium value, so the box solidifies (~97% solid) and the
  particles merge into one cluster within a few hundred steps.
- ``equilibrium``: c = 0.999 - 0.523 etaS, the two equilibrium compositions
  of the reference's free energies (AS = AL; the common tangent gives
  solid 0.476, liquid 0.999).  The solid fraction stays put and the
  particles coarsen, so the mean radius and size distribution carry
  information.
"""
import argparse
import json
import time

import numpy as np

from . import solver as S


def _radii(sizes):
    return np.cbrt(3.0 * np.asarray(sizes, np.float64) / (4.0 * np.pi))


def compare(ref, fp32, thr=0.5):
    """Fields (max |d| and RMS per field, on the device), solid fraction,
    total solute, particle count, mean radius and size histogram of two
    solvers of one grid."""
    out = {}
    d = ref.compare(fp32)
    for i, name in enumerate(("etaS", "etaL", "c")):
        out[f"{name}_max_abs"] = d["max_abs"][i]
        out[f"{name}_rms"] = d["rms"][i]
    s64, s32 = ref.reduce(), fp32.reduce()
    f64, f32 = s64["sum_h"] / s64["n_sites"], s32["sum_h"] / s32["n_sites"]
    out["solid_fraction_fp64"] = f64
    out["solid_fraction_fp32"] = f32
    out["solid_fraction_err_rel"] = abs(f32 - f64) / f64 if f64 else None
    out["sum_c_fp64"] = s64["sum_c"]
    out["sum_c_fp32"] = s32["sum_c"]
    out["sum_c_err_rel"] = abs(s32["sum_c"] - s64["sum_c"]) / abs(s64["sum_c"])
    p64, z64 = ref.particles(thr, max_sizes=1 << 22)
    p32, z32 = fp32.particles(thr, max_sizes=1 << 22)
    r64, r32 = _radii(z64), _radii(z32)
    top = int(max(r64.max() if len(r64) else 1, r32.max() if len(r32) else 1)) + 2
    h64 = np.histogram(r64, bins=top, range=(0, top))[0]
    h32 = np.histogram(r32, bins=top, range=(0, top))[0]
    out["particles_fp64"] = int(p64["n_particles"])
    out["particles_fp32"] = int(p32["n_particles"])
    out["mean_radius_fp64"] = p64["mean_radius"]
    out["mean_radius_fp32"] = p32["mean_radius"]
    out["mean_radius_err_rel"] = (abs(p32["mean_radius"] - p64["mean_radius"]) /
                                  p64["mean_radius"] if p64["mean_radius"] else None)
    n64, n32 = max(h64.sum(), 1), max(h32.sum(), 1)
    out["psd_l1"] = float(np.abs(h64 / n64 - h32 / n32).sum())  # 0 same, 2 disjoint
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--grid", type=int, default=1024)
    ap.add_argument("--spheres", type=int, default=12000)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--shift-every", type=int, default=50)
    ap.add_argument("--report-at", default="50,100,200,500,1000,1500,2000",
                    help="comma-separated steps at which to compare")
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--c-init", default="config", choices=("config", "equilibrium"))
    a = ap.parse_args(argv)
    c_liq, c_span = (0.1, 0.8) if a.c_init == "config" else (0.999, -0.523)
    n = a.grid
    shape = (n, n, n)
    cen, rad = S.spheres(a.seed, a.spheres, shape, 6.0, 14.0)
    ref = S.Solver(shape, precision=S.F64)
    ref.init_spheres(cen, rad, c_liq, c_span)
    fp = S.Solver(shape, precision=S.F32)
    fp.set_mass_correction(a.shift_every)
    fp.copy_from(ref)
    t64 = t32 = 0.0
    done = 0
    print(json.dumps({"step": 0, "c_init": a.c_init, "grid": n, "spheres": a.spheres,
                      **compare(ref, fp)}), flush=True)
    marks = sorted({int(s) for s in a.report_at.split(",") if 0 < int(s) <= a.steps} | {a.steps})
    for mark in marks:
        k = mark - done
        ref.sync()
        t = time.perf_counter()
        ref.step(k)
        ref.sync()
        t64 += time.perf_counter() - t
        t = time.perf_counter()
        fp.step(k)
        fp.sync()
        t32 += time.perf_counter() - t
        done += k
        rep = compare(ref, fp)
        rep.update({"step": done, "glups_fp64": n ** 3 * done / t64 / 1e9,
                    "glups_fp32": n ** 3 * done / t32 / 1e9,
                    "fp32_mass_target": fp.mass_target()[0]})
        print(json.dumps(rep), flush=True)
    ref.close()
    fp.close()


if __name__ == "__main__":
    main()
