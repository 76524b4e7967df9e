#!/usr/bin/env python3
"""Throughput of the explicit Moelans phase-field step on one MI355X.

Workload: 512^3 periodic grid, 2000 seeded spheres with R ~ U[6, 14] (overlaps
merged by max etaS), etaL = 1 - etaS, c = 0.1 + 0.8*etaS, the reference's
constants (phasefield.c:76-97).  A step is one sweep of the whole grid
(phasefield.c:186-215).  The state is created on the device (inputs resident
in HBM before timing starts); K steps are timed between two device
synchronisations.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--size N]
                    [--cpu-steps S] [--no-cpu]

Prints one JSON line.  value = FP64 GLUPS = N^3 * K / seconds / 1e9.
roofline: algorithmic bytes per launch (48 B/site in FP64: read etaS, etaL, c
and write them; a two-step launch, the periodic default, does so once for two
steps) over the kernel's average launch duration, from HIP events recorded on
the stream the kernels run on, with the one-step-equivalent rate (48 B per
site update) beside it; f64_single_step times the one-step kernel alone.  cpu_baseline: the CPU oracle (a
restatement of the reference algorithm, bit-exact against it) timed on this
host's cores on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "GLUPS & %HBM peak at 512^3 FP64 vs FP32; surrogate speedup @≤2% R̄ error"
HBM_PEAK_GBS = 8000.0
BYTES_PER_SITE = {"f64": 48, "f32": 24}


def traffic_from_profiles(kernel, workload):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE, scaled by a copy-kernel calibration), if committed."""
    p = os.path.join(HERE, "profiles", "traffic.json")
    if not os.path.exists(p):
        return None
    try:
        d = json.load(open(p))
        return d.get(f"{kernel}@{workload}", {}).get("bytes_per_launch")
    except Exception:
        return None


def copy_ceiling(nbytes=3 << 30, iters=20):
    """Measured device copy rate (read + write of nbytes) with the plain
    streaming kernel of libppfcal.so: the achievable HBM ceiling."""
    import ctypes

    from phasefield_coarsening_amd import _lib
    cal = _lib.load("libppfcal.so")
    cal.ppfcal_copy.argtypes = [ctypes.c_int, ctypes.c_size_t, ctypes.c_int,
                                ctypes.POINTER(ctypes.c_double)]
    ms = ctypes.c_double()
    rc = cal.ppfcal_copy(0, nbytes, iters, ctypes.byref(ms))
    if rc != 0:
        raise RuntimeError(f"ppfcal_copy failed ({rc})")
    return 2 * nbytes / (ms.value * 1e-3) / 1e9


def stream3_rate(n, pattern, tz=32, iters=20):
    """Measured rate (GB/s, 48 B per site) of a plain 3-in/3-out FP64 stream
    over n^3 sites with libppfcal.so's ppfcal_stream3: pattern 0 in flat
    memory order, 1-3 in the z-marching access pattern of the single-step
    kernel (64x8, 128x4, 512x1 column tiles, tz-plane chunks) with no
    stencil: the ceiling of that access pattern."""
    import ctypes

    from phasefield_coarsening_amd import _lib
    cal = _lib.load("libppfcal.so")
    cal.ppfcal_stream3.argtypes = [ctypes.c_int] * 7 + [ctypes.POINTER(ctypes.c_double)]
    ms = ctypes.c_double()
    rc = cal.ppfcal_stream3(0, n, n, n, pattern, tz, iters, ctypes.byref(ms))
    if rc != 0:
        raise RuntimeError(f"ppfcal_stream3 failed ({rc})")
    return 48.0 * n ** 3 / (ms.value * 1e-3) / 1e9


def ceilings(n):
    """Every stream pattern and chunk length: one JSON line."""
    out = {"grid": [n, n, n], "bytes_per_site": 48, "unit": "GB/s",
           "flat": round(stream3_rate(n, 0), 1)}
    for pattern, tile in ((1, "64x8"), (2, "128x4"), (3, "512x1")):
        for tz in (8, 16, 32, 64, 128):
            if n % tz == 0:
                out[f"march_{tile}_tz{tz}"] = round(stream3_rate(n, pattern, tz), 1)
    return out


def run_gpu(S, shape, precision, steps, warmup, seed=2, nsph=2000, rmin=6.0, rmax=14.0,
            fuse=None):
    cen, rad = S.spheres(seed, nsph, shape, rmin, rmax)
    sv = S.Solver(shape, precision=precision)
    sv.init_spheres(cen, rad)
    if fuse is not None:
        sv.set_fusion(fuse)
    st0 = sv.reduce()
    sv.step(warmup)
    sv.sync()
    t0 = time.perf_counter()
    sv.timer_start()
    sv.step(steps)
    ev_ms = sv.timer_stop()
    sv.sync()
    wall = time.perf_counter() - t0
    st1 = sv.reduce()
    kname = sv.kernel_name
    tiling = sv.get_tiling()
    fusion = sv.get_fusion()
    target = sv.mass_target()[0] if precision == S.F32 else None
    sv.close()
    return {"wall_s": wall, "event_ms": ev_ms, "kernel": kname, "stats0": st0, "stats1": st1,
            "tiling": tiling, "fusion": fusion, "c_target": target}


def cpu_baseline(S, shape, nsteps):
    """The oracle on host cores: nsteps sweeps of the same 512^3 workload."""
    from oracle import pyoracle as O
    cen, rad = S.spheres(2, 2000, shape, 6.0, 14.0)
    with S.Solver(shape) as sv:
        sv.init_spheres(cen, rad)
        s, l, c = sv.download()
    sec, threads = O.time_steps(s, l, c, nsteps)
    n = shape[0] * shape[1] * shape[2]
    return {"value": n * nsteps / sec / 1e9, "unit": "GLUPS", "cores": threads, "kind": "port",
            "sample": f"{nsteps} sweeps of the {shape[0]}^3 workload ({sec:.1f} s), oracle "
                      f"(C restatement of phasefield.c:301-403, bit-exact vs the reference), "
                      f"OpenMP {threads} threads"}


def reference_serial_baseline(S, shape, nsteps=3):
    """The reference program itself (oracle/_ref, its serial path compiled
    for 512^3 periodic), one core: wall time of a 3-step run minus a 0-step
    run on the same input file, i.e. the stepping loop without the HDF5
    read.  None when the binaries or HDF5 are not present."""
    import subprocess
    import tempfile
    ref = os.path.join(HERE, "oracle", "_ref")
    exe3, exe0 = (os.path.join(ref, f"ppf_ref_p{shape[0]}_s{k}") for k in (nsteps, 0))
    if not (os.path.exists(exe3) and os.path.exists(exe0)):
        return None
    try:
        from phasefield_coarsening_amd import h5
        cen, rad = S.spheres(2, 2000, shape, 6.0, 14.0)
        with S.Solver(shape) as sv:
            sv.init_spheres(cen, rad)
            s, l, c = sv.download()
        with tempfile.TemporaryDirectory(dir="/tmp") as tmp:
            h5.write(os.path.join(tmp, "in.h5"), s, l, c)
            del s, l, c
            times = []
            for exe in (exe0, exe3):
                t0 = time.perf_counter()
                subprocess.run([exe, "in.h5"], cwd=tmp, check=True, capture_output=True,
                               timeout=600)
                times.append(time.perf_counter() - t0)
    except Exception as e:  # reported, not fatal: the port baseline stands
        return {"error": str(e)[:200]}
    sec = times[1] - times[0]
    n = shape[0] * shape[1] * shape[2]
    return {"value": n * nsteps / sec / 1e9, "unit": "GLUPS", "cores": 1, "kind": "reference",
            "sample": f"{nsteps} sweeps of the {shape[0]}^3 workload by the reference program "
                      f"(phasefield.c serial path, gcc -O2), {sec:.1f} s = run of {nsteps} "
                      f"steps ({times[1]:.1f} s) minus run of 0 steps ({times[0]:.1f} s)"}


def unet_line(grid=256, reps=3):
    """U-Net forward (BF16 MFMA) at grid^3: time from HIP events and the
    MFMA work 2*27*Cin*Cout per site of each 3x3x3 conv."""
    import numpy as np

    from phasefield_coarsening_amd.unet import UNet
    with UNet(seed=1) as net:
        x = np.random.default_rng(0).uniform(0, 1, (1, grid, grid, grid, 3)).astype(np.float32)
        net.forward(x)
        best = None
        for _ in range(reps):
            net.forward(x)
            ms, flop = net.last_forward()
            best = ms if best is None else min(best, ms)
    tf = flop / best / 1e9
    return {"grid": [grid] * 3, "forward_ms": round(best, 3), "TFLOPs": round(tf, 1),
            "mfma_bf16_peak_TFLOPs": 2500.0, "mfma_frac": round(tf / 2500.0, 4)}


def surrogate_line(S, grid=256, t_end=2200, k=200, relax=20):
    """Learned jumps against exact FP64 stepping at grid^3 (500 spheres,
    R ~ U[6, 12]) to t_end, with the committed K = 200 weights: wall times,
    and the errors of the mean particle radius (connected components) and
    the solid fraction at t_end."""
    import numpy as np

    from phasefield_coarsening_amd import surrogate as SG
    from phasefield_coarsening_amd.unet import UNet
    w = os.path.join(HERE, "phasefield_coarsening_amd", "weights", "unet_k200_b32_f16.npy")
    shape = (grid, grid, grid)
    cen, rad = S.spheres(5, 500, shape, 6.0, 12.0)
    with S.Solver(shape) as ref:
        ref.init_spheres(cen, rad)
        s0 = ref.download()
        ref.sync()
        t0 = time.perf_counter()
        ref.step(t_end)
        ref.sync()
        t_exact = time.perf_counter() - t0
        pr, _ = ref.particles(0.5, max_sizes=1)
        st = ref.reduce()
    with S.Solver(shape, precision=S.F32) as sv, UNet(seed=1) as net:
        net.set_params(np.load(w).astype(np.float32))
        sv.upload(*s0)
        del s0
        ctl = SG.Controller(sv, net, k=k, relax=relax)
        summ = ctl.advance(t_end)
        ps, _ = sv.particles(0.5, max_sizes=1)
        ss = sv.reduce()
        ctl.close()
    f_ref, f_sur = st["sum_h"] / st["n_sites"], ss["sum_h"] / ss["n_sites"]
    return {"grid": [grid] * 3, "t_end": t_end, "k": k, "relax": relax,
            "exact_fp64_s": round(t_exact, 4), "surrogate_s": round(summ["seconds"], 4),
            "speedup": round(t_exact / summ["seconds"], 3), "accepted": summ["accepted"],
            "rejected": summ["rejected"], "skipped": summ["skipped"],
            "R_err_rel": abs(ps["mean_radius"] - pr["mean_radius"]) / pr["mean_radius"],
            "frac_err_rel": abs(f_sur - f_ref) / f_ref}


def ensemble_line(S, members=256, steps=1000, seed=4):
    """Batched ensemble (configs[3] shape): members with edges in {48, 64, 96,
    128}, solid fraction U[0.1, 0.4], radii U[3, 10], FP32, autotuned, one
    persistent launch per step."""
    from phasefield_coarsening_amd import surrogate as SG
    mem = SG.ensemble_members(members, seed)
    ens = S.Ensemble([m[0] for m in mem], precision=S.F32)
    try:
        for i, (_, cen, rad) in enumerate(mem):
            ens.init_spheres(i, cen, rad)
        ens.autotune(reps=2)
        ens.step(10)
        ens.timer_start()
        ens.step(steps)
        ms = ens.timer_stop()
        info = ens.info()
    finally:
        ens.close()
    g = info["total_sites"] * steps / (ms * 1e-3) / 1e9
    return {"members": members, "total_sites": info["total_sites"], "steps": steps,
            "GLUPS": round(g, 2), "roofline_achieved_GBs": round(g * 24, 1),
            "roofline_frac": round(g * 24 / HBM_PEAK_GBS, 4),
            "work_items": info["items"], "persistent_workgroups": info["grid"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--cpu-steps", type=int, default=60)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-f32", action="store_true")
    ap.add_argument("--no-copy", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the ensemble and U-Net lines")
    ap.add_argument("--ceilings", action="store_true",
                    help="only measure the 3-in/3-out stream ceilings (flat and z-marching "
                         "patterns, every chunk length) at --size and print them")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              "(multi-GPU sharding is design only)", file=sys.stderr)
        sys.exit(2)

    if a.ceilings:
        print(json.dumps(ceilings(a.size)), flush=True)
        return
    from phasefield_coarsening_amd import solver as S
    n = a.size
    shape = (n, n, n)
    sites = n ** 3
    r = run_gpu(S, shape, S.F64, a.steps, a.warmup)
    glups = sites * a.steps / r["wall_s"] / 1e9
    # steps per launch: 2 = two steps in one z march (periodic default), so a
    # launch reads and writes the state once for two site updates
    spl = r["fusion"][0]
    per_step_s = r["event_ms"] * 1e-3 / a.steps
    avg_launch_s = per_step_s * spl
    achieved = BYTES_PER_SITE["f64"] * sites / avg_launch_s / 1e9
    equiv = BYTES_PER_SITE["f64"] * sites / per_step_s / 1e9
    workload = f"{n}^3 periodic, 2000 spheres R~U[6,14], FP64"
    out = {
        "metric": METRIC,
        "value": round(glups, 3),
        "unit": "GLUPS",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(r["wall_s"] * 1e3 / a.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded sphere initial condition generated on the device)",
        "config": {"workload": workload, "grid": [n, n, n], "bc": "periodic", "spheres": 2000,
                   "radius": [6, 14], "kernel": r["kernel"]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_from_profiles(r["kernel"], f"{n}^3-f64"),
                     "bytes_per_site": BYTES_PER_SITE["f64"],
                     "avg_launch_us": round(avg_launch_s * 1e6, 2),
                     "steps_per_launch": spl,
                     "hbm_bytes_per_site_update": BYTES_PER_SITE["f64"] / spl,
                     "one_step_equivalent_GBs": round(equiv, 1),
                     "one_step_equivalent_frac": round(equiv / HBM_PEAK_GBS, 4),
                     "kernel_limit": ("FP64 VALU latency: two steps per launch move half the "
                                      "bytes of two single steps" if spl == 2 else "HBM")},
        "solute_drift_rel": abs(r["stats1"]["sum_c"] - r["stats0"]["sum_c"]) /
        abs(r["stats0"]["sum_c"]),
    }
    if spl == 2:
        f = r["fusion"]
        out["config"]["tiling"] = {"steps_per_launch": 2, "tx": f[1], "ty": f[2], "tz": f[3],
                                   "order": f[4]}
    else:
        out["config"]["tiling"] = {"tx": r["tiling"][0], "ty": r["tiling"][1],
                                   "tz": r["tiling"][2], "order": r["tiling"][3]}
    if spl == 2:
        r1 = run_gpu(S, shape, S.F64, a.steps, a.warmup, fuse=1)
        per1_s = r1["event_ms"] * 1e-3 / a.steps
        ach1 = BYTES_PER_SITE["f64"] * sites / per1_s / 1e9
        out["f64_single_step"] = {
            "GLUPS": round(sites * a.steps / r1["wall_s"] / 1e9, 3), "kernel": r1["kernel"],
            "tiling": list(r1["tiling"]), "avg_launch_us": round(per1_s * 1e6, 2),
            "achieved_GBs": round(ach1, 1), "achieved_frac": round(ach1 / HBM_PEAK_GBS, 4),
            "traffic": traffic_from_profiles(r1["kernel"], f"{n}^3-f64")}
    if not a.no_f32:
        r32 = run_gpu(S, shape, S.F32, a.steps, a.warmup)
        g32 = sites * a.steps / r32["wall_s"] / 1e9
        spl32 = r32["fusion"][0]  # steps per launch: 2 = both steps in one z march
        per_step32_s = r32["event_ms"] * 1e-3 / a.steps
        # one-step form: 24 B per site update; two steps per launch move the
        # same 24 B per site once for two updates
        equiv32 = BYTES_PER_SITE["f32"] * sites / per_step32_s / 1e9
        actual32 = equiv32 / spl32
        out["f32"] = {"GLUPS": round(g32, 3), "ms_per_step": round(r32["wall_s"] * 1e3 / a.steps, 4),
                      "vs_f64": round(g32 / glups, 3), "kernel": r32["kernel"],
                      "steps_per_launch": spl32,
                      "hbm_bytes_per_site_update": BYTES_PER_SITE["f32"] / spl32,
                      "achieved_GBs": round(actual32, 1),
                      "achieved_frac": round(actual32 / HBM_PEAK_GBS, 4),
                      "one_step_equivalent_GBs": round(equiv32, 1),
                      "one_step_equivalent_frac": round(equiv32 / HBM_PEAK_GBS, 4),
                      "solute_vs_fp64_target_rel":
                          abs(r32["stats1"]["sum_c"] - r32["c_target"]) / abs(r32["c_target"])}
    if not a.no_extras:
        out["ensemble"] = ensemble_line(S)
        out["unet"] = unet_line()
        out["surrogate"] = surrogate_line(S)
    if not a.no_copy:
        cc = copy_ceiling()
        out["roofline"]["copy_ceiling_GBs"] = round(cc, 1)
        out["roofline"]["frac_of_copy_ceiling"] = round(achieved / cc, 4)
        out["roofline"]["one_step_equivalent_frac_of_copy_ceiling"] = round(equiv / cc, 4)
        if n % 64 == 0 and n % 32 == 0:
            # the same bytes as a plain stream: in flat order, and in the
            # single-step kernel's z-marching pattern (64x8 tiles, 32 planes)
            out["roofline"]["stream3_flat_GBs"] = round(stream3_rate(n, 0), 1)
            out["roofline"]["stream3_zmarch_GBs"] = round(stream3_rate(n, 1, 32), 1)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(S, shape, a.cpu_steps)
        ref = reference_serial_baseline(S, shape)
        if ref is not None:
            out["cpu_baseline_reference_serial"] = ref
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
