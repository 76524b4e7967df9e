"""Tiling autotuner front end with a per-device cache.

``tuned(solver)`` returns the tiling for the solver's grid, precision and
boundary condition, and whether two steps run per launch: from the cache
file when present, otherwise by running ``ppf_autotune`` on the device
(which does not advance the state) and recording the winner.  The cache is a JSON table keyed by
``"<nz>x<ny>x<nx>/<f64|f32>/<bc>"``; its path is PPF_TUNING_CACHE or
~/.cache/ppf_tuning.json.
"""
import json
import os

from . import solver as S


def cache_path():
    return os.environ.get("PPF_TUNING_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "ppf_tuning.json"))


def key(shape, precision, bc):
    return "%dx%dx%d/%s/%s" % (tuple(shape) + ("f32" if precision == S.F32 else "f64",
                                               "noflux" if bc == S.BC_NOFLUX else "periodic"))


def load():
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def save(table):
    p = cache_path()
    os.makedirs(os.path.dirname(p), exist_ok=True)
    tmp = p + ".tmp"
    with open(tmp, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    os.replace(tmp, p)


def tuned(sv, reps=3, refresh=False):
    """Apply the cached or freshly tuned tiling to ``sv``; return it."""
    table = load()
    k = key(sv.shape, sv.precision, sv.bc)
    if not refresh and k in table and "fusion" in table[k]:
        t = table[k]
        sv.set_tiling(t["tx"], t["ty"], t["tz"], t["order"])
        f = t["fusion"]
        if f[0] == 2:
            sv.set_fusion(2, f[1], f[2], f[3], f[4])
        elif sv.bc == S.BC_PERIODIC:
            sv.set_fusion(1)
        return t
    (tx, ty, tz, order), ms = sv.autotune(reps)
    t = {"tx": tx, "ty": ty, "tz": tz, "order": order, "ms_per_sweep": ms,
         "fusion": list(sv.get_fusion())}
    table[k] = t
    save(table)
    return t
