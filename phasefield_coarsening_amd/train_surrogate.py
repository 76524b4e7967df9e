"""Train the K-step U-Net on ensemble trajectories and measure the surrogate.

    python -m phasefield_coarsening_amd.train_surrogate [options]

1. Ensemble (batched, one persistent launch per step): ``--members``
   periodic domains, edges drawn from ``--edges``, solid fraction
   U[0.1, 0.4], radii U[3, 10], FP32, autotuned tiles; ``--ens-steps``
   steps with snapshots every K = ``--k`` steps of the members whose edge is
   ``--train-edge``.  Reports the ensemble GLUPS.
2. Training: ``--train-steps`` Adam steps (batch ``--batch``) on snapshot
   pairs K steps apart; one member in eight is held out.
3. Held-out evaluation: mean squared error of the network's jump against
   "no change" (persistence) on the held-out pairs.
4. Surrogate at ``--grid``^3 (``--spheres`` spheres, R ~ U[6, 12]): the
   exact FP64 run (fast FP64 kernel) to ``--t-end`` steps against the
   controller (jump K, 20 FP32 relax steps, judge ``--judge``, exact re-run
   on rejection), compared on mean particle radius (connected components)
   and solid fraction at t_end; wall times give the speedup against plain
   FP64 stepping and against plain FP32 stepping (what the network gains).
``--c-init`` picks the initial composition of the ensemble and of the
surrogate grid (surrogate.COMPOSITIONS).

Weights go to ``--save`` (float16 .npy) when given; ``--load`` skips
training.  One JSON line per stage is printed.
"""
import argparse
import json
import os
import time

import numpy as np

from . import solver as S
from . import surrogate as SG
from .unet import UNet


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--members", type=int, default=256)
    ap.add_argument("--edges", default="48,64,96,128")
    ap.add_argument("--ens-steps", type=int, default=4000)
    ap.add_argument("--k", type=int, default=200)
    ap.add_argument("--train-edge", type=int, default=64)
    ap.add_argument("--max-train-members", type=int, default=48)
    ap.add_argument("--train-steps", type=int, default=600)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--spheres", type=int, default=500)
    ap.add_argument("--t-end", type=int, default=2200)
    ap.add_argument("--relax", type=int, default=20)
    ap.add_argument("--judge", default="area", choices=("area", "particles"),
                    help="area: volume-to-area radius from one reduction; particles: "
                         "connected-component mean radius")
    ap.add_argument("--timing", action="store_true",
                    help="synchronise around each controller phase and report its seconds "
                         "(adds a little time)")
    ap.add_argument("--early", default="0,1",
                    help="n,rep: training pairs from the first n snapshots of a trajectory "
                         "count rep times")
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--c-init", default="config", choices=sorted(SG.COMPOSITIONS),
                    help="initial composition of the ensemble and of the surrogate grid: "
                         "config (c = 0.1 + 0.8 etaS) or equilibrium (c = 0.999 - 0.523 etaS)")
    ap.add_argument("--save", default="")
    ap.add_argument("--load", default="",
                    help="weights .npy (float16 or float32); skips the ensemble and training")
    a = ap.parse_args(argv)

    net = UNet(seed=a.seed)
    if a.load:
        net.set_params(np.load(a.load).astype(np.float32))
    else:
        edges = tuple(int(e) for e in a.edges.split(","))
        members = SG.ensemble_members(a.members, a.seed, edges=edges)
        keep = [i for i, m in enumerate(members) if m[0][0] == a.train_edge][:a.max_train_members]
        snaps, ens = SG.trajectories(members, a.ens_steps, a.k, keep=keep, c_init=a.c_init)
        print(json.dumps({"stage": "ensemble", **ens, "train_members": len(keep)}), flush=True)
        held = {i: v for j, (i, v) in enumerate(sorted(snaps.items())) if j % 8 == 7}
        train_snaps = {i: v for i, v in snaps.items() if i not in held}
        early = tuple(int(v) for v in a.early.split(","))
        data, test = SG.pairs(train_snaps, early=early), SG.pairs(held)
        t0 = time.perf_counter()
        losses = SG.train(net, data, a.train_steps, lr=a.lr, batch=a.batch, seed=a.seed,
                          log_every=max(1, a.train_steps // 10))
        dt = time.perf_counter() - t0
        print(json.dumps({"stage": "train", "steps": a.train_steps, "seconds": round(dt, 1),
                          "loss_first": losses[0], "loss_last": float(np.mean(losses[-20:])),
                          "pairs": sum(len(v) for v in data.values())}), flush=True)
        ln, lp = SG.eval_loss(net, test, a.batch)
        print(json.dumps({"stage": "heldout", "mse_net": ln, "mse_persistence": lp,
                          "gain": lp / ln if ln > 0 else None}), flush=True)
        if a.save:
            np.save(a.save, net.get_params().astype(np.float16))

    # surrogate vs exact FP64 at grid^3
    n = a.grid
    shape = (n, n, n)
    cen, rad = S.spheres(5, a.spheres, shape, 6.0, 12.0)
    with S.Solver(shape, kernel=S.KERNEL_FAST) as ref:
        ref.init_spheres(cen, rad, *SG.COMPOSITIONS[a.c_init])
        s0 = ref.download()
        ref.sync()
        t0 = time.perf_counter()
        ref.step(a.t_end)
        ref.sync()
        t_exact = time.perf_counter() - t0
        pr, _ = ref.particles(0.5, max_sizes=1)
        st = ref.reduce()
        r_exact, f_exact = pr["mean_radius"], st["sum_h"] / st["n_sites"]
        n_exact = pr["n_particles"]
    with S.Solver(shape, precision=S.F32) as f32:
        f32.upload(*s0)
        f32.sync()
        t0 = time.perf_counter()
        f32.step(a.t_end)
        f32.sync()
        t_exact32 = time.perf_counter() - t0
    with S.Solver(shape, precision=S.F32) as sv:
        sv.upload(*s0)
        del s0
        ctl = SG.Controller(sv, net, k=a.k, relax=a.relax, judge=a.judge, timing=a.timing)
        summ = ctl.advance(a.t_end)
        pr, _ = sv.particles(0.5, max_sizes=1)
        st = sv.reduce()
        r_sur, f_sur = pr["mean_radius"], st["sum_h"] / st["n_sites"]
        ms, fl = net.last_forward()
        ctl.close()
    print(json.dumps({"stage": "surrogate", "grid": n, "t_end": a.t_end, "k": a.k,
                      "c_init": a.c_init, "particles_exact": n_exact,
                      "relax": a.relax, "exact_fp64_seconds": round(t_exact, 4),
                      "exact_fp32_seconds": round(t_exact32, 4),
                      "surrogate_seconds": round(summ["seconds"], 4),
                      "speedup_vs_exact_fp64": round(t_exact / summ["seconds"], 3),
                      "speedup_vs_exact_fp32": round(t_exact32 / summ["seconds"], 3),
                      "judge": a.judge, "accepted": summ["accepted"],
                      "rejected": summ["rejected"], "skipped": summ["skipped"],
                      "phase_seconds": summ.get("phase_seconds"),
                      "exact_steps_in_surrogate": summ["exact_steps"],
                      "R_exact": r_exact, "R_surrogate": r_sur,
                      "R_err_rel": abs(r_sur - r_exact) / r_exact,
                      "frac_exact": f_exact, "frac_surrogate": f_sur,
                      "frac_err_rel": abs(f_sur - f_exact) / f_exact,
                      "unet_forward_ms": ms, "unet_forward_TFLOPs": fl / ms / 1e9 if ms else None,
                      "jumps": ctl.log}), flush=True)
    net.close()


if __name__ == "__main__":
    main()
