"""numpy float32 restatement of the FP32 step form (PhysF32 in
phasefield_coarsening_amd/csrc/ppf_physics.h), for its parity tests.

Same dimensionless constants (computed in float64, rounded once to
float32) and the same operation order as the kernel, each operation
rounded to float32 (numpy float32 arithmetic is IEEE single), so one step
agrees with the GPU bit for bit.  Periodic boundaries.
"""
import numpy as np

f32 = np.float32


def constants(p):
    tmp1 = -1.0 * p.L * p.dt
    tmp3 = p.dt * 0.5 * p.h2
    return dict(cL0=f32(p.cL0), tmp6=f32(p.cL0 - p.cS0), rM=f32((p.MS - p.ML) / p.ML),
                a1=f32(tmp1 * p.m), a2=f32(-tmp1 * p.kappa * p.h2), g2=f32(2.0 * p.gamma),
                b0=f32(tmp1 * (p.CS - p.CL)), b1=f32(tmp1 * p.AL * (p.cL0 - p.cS0)),
                d0=f32(tmp3 * p.ML * p.AL))


def _nb(a):
    """six neighbours in the order z+1, z-1, y+1, y-1, x+1, x-1 (periodic)"""
    return [np.roll(a, -1, 0), np.roll(a, 1, 0), np.roll(a, -1, 1), np.roll(a, 1, 1),
            np.roll(a, -1, 2), np.roll(a, 1, 2)]


def step(s, l, c, k):
    s, l, c = (np.asarray(a, dtype=f32) for a in (s, l, c))
    ss = s * s
    h = ss / (ss + l * l)
    u = (c - k["cL0"]) + k["tmp6"] * h
    mm = f32(1.0) + k["rM"] * h
    sn, ln, un, mn = _nb(s), _nb(l), _nb(u), _nb(mm)
    ll = l * l
    den = ss + ll
    D = (f32(2.0) * s * l) / (den * den) * (k["b0"] + k["b1"] * u)
    lap_s = ((sn[0] + sn[1]) + (sn[2] + sn[3])) + (sn[4] + sn[5]) - f32(6.0) * s
    lap_l = ((ln[0] + ln[1]) + (ln[2] + ln[3])) + (ln[4] + ln[5]) - f32(6.0) * l
    ds = k["a1"] * s * (ss - f32(1.0) + k["g2"] * ll) + k["a2"] * lap_s + l * D
    dl = k["a1"] * l * (ll - f32(1.0) + k["g2"] * ss) + k["a2"] * lap_l - s * D
    fl = (((mn[0] + mm) * (un[0] - u) + (mn[1] + mm) * (un[1] - u)) +
          ((mn[2] + mm) * (un[2] - u) + (mn[3] + mm) * (un[3] - u)) +
          ((mn[4] + mm) * (un[4] - u) + (mn[5] + mm) * (un[5] - u)))
    dc = k["d0"] * fl
    return s + ds, l + dl, c + dc
