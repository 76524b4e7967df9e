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

Prints one JSON line.  value = FP64 GLUPS = N^3 * K / seconds / 1e9 with the
fast FP64 kernel (PPF_KERNEL_FAST: fused multiply-adds, one reciprocal per
site, within 1e-10 per field of the reference after 1000 steps); --f64-kernel
bitwise times the kernel that reproduces the reference bit for bit instead,
and the line always carries the other one under f64_bitwise / f64_fast.
roofline: algorithmic bytes per launch over the kernel's average launch
duration (HIP events recorded on the stream the kernels run on), counted as
the roofline defines them: 48 B per site update in FP64 (read etaS, etaL, c
and write them), 24 B in FP32, times the site updates one launch makes (two
per site with two steps per launch, the periodic default).  The bytes a
launch must move (one read and one write of the state per launch) are
reported beside it as moved_GBs / moved_frac, and the HBM traffic rocprofv3
measured per launch as traffic.  f64_single_step times the one-step kernel
alone.  cpu_baseline: the CPU oracle (a restatement of the reference
algorithm, bit-exact against it) timed on this host's cores on a bounded
sample of the same workload.
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
            fuse=None, kernel=0):
    cen, rad = S.spheres(seed, nsph, shape, rmin, rmax)
    sv = S.Solver(shape, precision=precision, kernel=kernel)
    sv.init_spheres(cen, rad)
    if fuse is not None:
        sv.set_fusion(fuse)
    st0 = sv.reduce()
    sv.step(warmup)
    sv.sync()
    nsh0 = sv.mass_target()[1] if precision == S.F32 else 0
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
    target, nsh1 = sv.mass_target() if precision == S.F32 else (None, 0)
    sv.close()
    return {"wall_s": wall, "event_ms": ev_ms, "kernel": kname, "stats0": st0, "stats1": st1,
            "tiling": tiling, "fusion": fusion, "c_target": target, "n_shifts0": nsh0,
            "n_shifts1": nsh1}


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


def unet_line(grid=256, reps=3, apply_reps=10):
    """U-Net forward (BF16 MFMA) at grid^3: time from HIP events and the
    MFMA work 2*27*Cin*Cout per site of each 3x3x3 conv.  forward_ms: the
    host-side forward (random input in [0, 1], random weights; each call
    copies 200 MB in and out through pageable host memory, so the GPU idles
    ~28 ms between forwards); apply_ms: ppf_unet_apply on an FP32 context of
    the surrogate's workload (500 spheres, equilibrium composition, trained
    weights), the surrogate's path, best of apply_reps back-to-back calls
    (the first calls after an idle GPU run ~10% slower while its clock
    ramps up: 5.85, 5.51, 5.37 ms for the first three), with the first
    call's time reported beside it."""
    import numpy as np

    from phasefield_coarsening_amd import solver as S
    from phasefield_coarsening_amd.unet import UNet
    with UNet(seed=1) as net:
        x = np.random.default_rng(0).uniform(0, 1, (1, grid, grid, grid, 3)).astype(np.float32)
        net.forward(x)
        best = None
        for _ in range(reps):
            net.forward(x)
            ms, flop = net.last_forward()
            best = ms if best is None else min(best, ms)
    shape = (grid,) * 3
    cen, rad = S.spheres(5, 500, shape, 6.0, 12.0)
    w = os.path.join(HERE, "phasefield_coarsening_amd", "weights", "unet_k200_b32_eq_f16.npy")
    with S.Solver(shape, precision=S.F32) as sv, UNet(seed=1) as net:
        net.set_params(np.load(w).astype(np.float32))
        sv.init_spheres(cen, rad, 0.999, -0.523)
        times = []
        for _ in range(apply_reps):
            net.apply(sv)
            sv.sync()
            times.append(net.last_forward()[0])
    best_a = min(times)
    tf = flop / best / 1e9
    return {"grid": [grid] * 3, "forward_ms": round(best, 3), "TFLOPs": round(tf, 1),
            "mfma_bf16_peak_TFLOPs": 2500.0, "mfma_frac": round(tf / 2500.0, 4),
            "apply_ms": round(best_a, 3), "apply_TFLOPs": round(flop / best_a / 1e9, 1),
            "apply_mfma_frac": round(flop / best_a / 1e9 / 2500.0, 4),
            "apply_ms_first": round(times[0], 3), "apply_calls": apply_reps}


def surrogate_line(S, grid=256, t_end=2200, k=200, relax=20, c_init="equilibrium", no_op=True,
                   first_jump=False):
    """Learned jumps against exact stepping at grid^3 (500 spheres,
    R ~ U[6, 12]) to t_end: wall times of the surrogate (jumps of k steps
    with the weights trained for k, 20 FP32 relax steps, judge, exact re-run
    on rejection), of plain FP64
    stepping (the fast FP64 kernel) and of plain FP32 stepping, and the
    errors of the mean particle radius (connected components, over all
    particles) and of the solid fraction at t_end against FP64.  c_init
    "equilibrium" (c = 0.999 - 0.523 etaS, phasefield.c:82-83): the
    particles' interfaces relax and then settle, with the weights trained on
    that regime; "config" (c = 0.1 + 0.8 etaS): the box solidifies into one
    cluster, with the weights trained there.  no_op: also run the same
    controller with jumps that change nothing (the network's head zeroed) and
    report its errors: what the trained jumps gain in accuracy at the same
    cost.  first_jump: the controller tries a jump on the first cycle too."""
    import numpy as np

    from phasefield_coarsening_amd import surrogate as SG
    from phasefield_coarsening_amd.unet import UNet, layer_shapes
    wname = "unet_k%d_b32%s_f16.npy" % (k, "_eq" if c_init == "equilibrium" else "")
    w = np.load(os.path.join(HERE, "phasefield_coarsening_amd", "weights", wname)).astype(np.float32)
    shape = (grid, grid, grid)
    cen, rad = S.spheres(5, 500, shape, 6.0, 12.0)
    with S.Solver(shape, kernel=S.KERNEL_FAST) as ref:
        ref.init_spheres(cen, rad, *SG.COMPOSITIONS[c_init])
        s0 = ref.download()
        ref.sync()
        t0 = time.perf_counter()
        ref.step(t_end)
        ref.sync()
        t_exact = time.perf_counter() - t0
        pr, _ = ref.particles(0.5, max_sizes=1)
        st = ref.reduce()
    # plain FP32 stepping to the same t_end: what the network gains over the
    # stepper it relaxes and falls back with
    with S.Solver(shape, precision=S.F32) as f32:
        f32.upload(*s0)
        f32.sync()
        t0 = time.perf_counter()
        f32.step(t_end)
        f32.sync()
        t_exact32 = time.perf_counter() - t0
    f_ref = st["sum_h"] / st["n_sites"]

    def controlled(params):
        with S.Solver(shape, precision=S.F32) as sv, UNet(seed=1) as net:
            net.set_params(params)
            sv.upload(*s0)
            ctl = SG.Controller(sv, net, k=k, relax=relax, first_jump=first_jump)
            summ = ctl.advance(t_end)
            ps, _ = sv.particles(0.5, max_sizes=1)
            ss = sv.reduce()
            ctl.close()
        return (summ, ps, abs(ps["mean_radius"] - pr["mean_radius"]) / pr["mean_radius"],
                abs(ss["sum_h"] / ss["n_sites"] - f_ref) / f_ref)

    summ, ps, r_err, f_err = controlled(w)
    out = {"grid": [grid] * 3, "t_end": t_end, "k": k, "relax": relax, "c_init": c_init,
           "weights": wname, "particles_fp64": pr["n_particles"],
           "particles_surrogate": ps["n_particles"],
           "exact_fp64_s": round(t_exact, 4), "exact_fp32_s": round(t_exact32, 4),
           "surrogate_s": round(summ["seconds"], 4),
           "speedup_vs_exact_fp64": round(t_exact / summ["seconds"], 3),
           "speedup_vs_exact_fp32": round(t_exact32 / summ["seconds"], 3),
           "accepted": summ["accepted"], "rejected": summ["rejected"],
           "skipped": summ["skipped"], "R_err_rel": r_err, "frac_err_rel": f_err,
           "first_jump": first_jump}
    if no_op:
        wz = w.copy()
        nh = sum(int(np.prod(shp)) for name, shp in layer_shapes(32) if name.startswith("head"))
        wz[-nh:] = 0.0
        _, _, r0, f0 = controlled(wz)
        out["no_op_jumps"] = {"R_err_rel": r0, "frac_err_rel": f0}
    return out


def ensemble_line(S, members=256, steps=1000, seed=4):
    """Batched ensemble of 256 uneven periodic domains: members with edges in {48, 64, 96,
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
        spl = ens.get_fusion()
    finally:
        ens.close()
    g = info["total_sites"] * steps / (ms * 1e-3) / 1e9
    # 24 B per site update (FP32); a two-step launch reads and writes the
    # state once for two updates, so it moves half of that per update
    return {"members": members, "total_sites": info["total_sites"], "steps": steps,
            "GLUPS": round(g, 2), "steps_per_launch": spl,
            "roofline_achieved_GBs": round(g * 24, 1),
            "roofline_frac": round(g * 24 / HBM_PEAK_GBS, 4),
            "moved_GBs": round(g * 24 / spl, 1),
            "moved_frac": round(g * 24 / spl / HBM_PEAK_GBS, 4),
            "work_items": info["items"], "persistent_workgroups": info["grid"]}


def tiling_of(r):
    f = r["fusion"]
    if f[0] == 2:
        return {"steps_per_launch": 2, "tx": f[1], "ty": f[2], "tz": f[3], "order": f[4]}
    t = r["tiling"]
    return {"steps_per_launch": 1, "tx": t[0], "ty": t[1], "tz": t[2], "order": t[3]}


def roofline(r, sites, steps, prec, n):
    """HBM roofline of one timed run.  achieved = B per site update (48 FP64,
    24 FP32) x site updates per launch / average launch time; moved = the
    bytes one launch must read and write (the state once, whatever the
    steps per launch) / the same time; traffic = rocprofv3-measured HBM
    bytes per launch, when profiles/traffic.json holds them."""
    spl = r["fusion"][0]
    b = BYTES_PER_SITE[prec]
    launch_s = r["event_ms"] * 1e-3 / steps * spl
    achieved = b * sites * spl / launch_s / 1e9
    moved = b * sites / launch_s / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic_from_profiles(r["kernel"], f"{n}^3-{prec}"),
            "bytes_per_site_update": b, "steps_per_launch": spl,
            "avg_launch_us": round(launch_s * 1e6, 2),
            "moved_GBs": round(moved, 1), "moved_frac": round(moved / HBM_PEAK_GBS, 4)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--cpu-steps", type=int, default=60)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--f64-kernel", choices=("fast", "bitwise"), default="fast",
                    help="FP64 kernel of the headline value (the other is reported beside it)")
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
    kernels = {"fast": S.KERNEL_FAST, "bitwise": S.KERNEL_TILED}
    head = a.f64_kernel
    other = "bitwise" if head == "fast" else "fast"
    workload = f"{n}^3 periodic, 2000 spheres R~U[6,14], FP64"
    # side measurements first: they also bring the GPU to its working clocks
    # before the timed region of the headline
    side = {}
    if not a.no_copy:
        side["copy"] = copy_ceiling()
        if n % 64 == 0 and n % 32 == 0:
            # the same bytes as a plain stream: in flat order, and in the
            # single-step kernel's z-marching pattern (64x8 tiles, 32 planes)
            side["flat"] = stream3_rate(n, 0)
            side["zmarch"] = stream3_rate(n, 1, 32)
    r_other = run_gpu(S, shape, S.F64, a.steps, a.warmup, kernel=kernels[other])
    r = run_gpu(S, shape, S.F64, a.steps, a.warmup, kernel=kernels[head])
    glups = sites * a.steps / r["wall_s"] / 1e9
    rf = roofline(r, sites, a.steps, "f64", n)
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
                   "radius": [6, 14], "kernel": r["kernel"],
                   "f64_form": ("fast: fused multiply-adds and the collapsed AS == AL "
                                "chemistry, within 1e-10 per field of the reference after "
                                "1000 steps" if head == "fast" else
                                "bitwise: the reference's operand order, bit for bit"),
                   "tiling": tiling_of(r)},
        "roofline": rf,
        "solute_drift_rel": abs(r["stats1"]["sum_c"] - r["stats0"]["sum_c"]) /
        abs(r["stats0"]["sum_c"]),
    }
    g_other = sites * a.steps / r_other["wall_s"] / 1e9
    out["f64_" + other] = {"GLUPS": round(g_other, 3), "kernel": r_other["kernel"],
                           "tiling": tiling_of(r_other),
                           "roofline": roofline(r_other, sites, a.steps, "f64", n),
                           "solute_drift_rel": abs(r_other["stats1"]["sum_c"] -
                                                   r_other["stats0"]["sum_c"]) /
                           abs(r_other["stats0"]["sum_c"])}
    r1 = run_gpu(S, shape, S.F64, a.steps, a.warmup, fuse=1, kernel=S.KERNEL_TILED)
    out["f64_single_step"] = {"GLUPS": round(sites * a.steps / r1["wall_s"] / 1e9, 3),
                              "kernel": r1["kernel"], "tiling": list(r1["tiling"]),
                              "roofline": roofline(r1, sites, a.steps, "f64", n)}
    if not a.no_f32:
        # at least 100 steps, so that the FP32 form's FP64 mass shifts (every
        # 50 steps) fall inside the timed window
        k32 = max(a.steps, 100)
        r32 = run_gpu(S, shape, S.F32, k32, a.warmup)
        g32 = sites * k32 / r32["wall_s"] / 1e9
        out["f32"] = {"GLUPS": round(g32, 3), "steps": k32,
                      "ms_per_step": round(r32["wall_s"] * 1e3 / k32, 4),
                      "vs_f64": round(g32 / glups, 3),
                      "vs_f64_bitwise": round(g32 / (glups if head == "bitwise" else g_other), 3),
                      "kernel": r32["kernel"], "tiling": tiling_of(r32),
                      "mass_shifts_in_window": r32["n_shifts1"] - r32["n_shifts0"],
                      "roofline": roofline(r32, sites, k32, "f32", n),
                      "solute_vs_fp64_target_rel":
                          abs(r32["stats1"]["sum_c"] - r32["c_target"]) / abs(r32["c_target"])}
    if not a.no_extras:
        out["ensemble"] = ensemble_line(S)
        out["unet"] = unet_line()
        out["surrogate"] = surrogate_line(S)
        out["surrogate_long"] = surrogate_line(S, t_end=19800, no_op=False)
        out["surrogate_solidifying"] = surrogate_line(S, c_init="config", no_op=False)
        out["surrogate_first_jump"] = surrogate_line(S, no_op=False, first_jump=True)
        out["surrogate_k400"] = surrogate_line(S, t_end=4400, k=400)
        out["surrogate_k400_long"] = surrogate_line(S, t_end=19800, k=400, no_op=False)
        out["surrogate_k400_first_jump"] = surrogate_line(S, t_end=4400, k=400, no_op=False,
                                                          first_jump=True)
        out["surrogate_k800_long"] = surrogate_line(S, t_end=19800, k=800)
        out["surrogate_k400_solidifying"] = surrogate_line(S, t_end=4400, k=400, c_init="config",
                                                           no_op=False)
    if "copy" in side:
        cc = side["copy"]
        out["roofline"]["copy_ceiling_GBs"] = round(cc, 1)
        out["roofline"]["frac_of_copy_ceiling"] = round(rf["achieved"] / cc, 4)
        if "flat" in side:
            out["roofline"]["stream3_flat_GBs"] = round(side["flat"], 1)
            out["roofline"]["stream3_zmarch_GBs"] = round(side["zmarch"], 1)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(S, shape, a.cpu_steps)
        ref = reference_serial_baseline(S, shape)
        if ref is not None:
            out["cpu_baseline_reference_serial"] = ref
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
