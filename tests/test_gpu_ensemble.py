"""Batched ensembles: one persistent launch per step over members of
uneven size.  Each member must evolve exactly as it would alone (FP64:
bitwise equal to the oracle; FP32: bitwise equal to a single-grid FP32
context), whatever tile shape and z chunk it is given."""
import numpy as np
import pytest

from oracle import pyoracle as O
from phasefield_coarsening_amd import solver as S

pytestmark = pytest.mark.gpu

DIMS = [(48, 48, 48), (20, 16, 12), (33, 17, 65), (64, 64, 64), (9, 40, 96), (48, 48, 48)]


def _member_state(shape, seed):
    cen, rad = S.spheres(seed, 6, shape, 2.5, 6.0)
    with S.Solver(shape) as sv:
        sv.init_spheres(cen, rad)
        return sv.download()


@pytest.mark.parametrize("bc", [S.BC_PERIODIC, S.BC_NOFLUX])
def test_ensemble_fp64_members_bitwise_vs_oracle(bc):
    states = [_member_state(d, 10 + i) for i, d in enumerate(DIMS)]
    with S.Ensemble(DIMS, bc=bc, precision=S.F64) as ens:
        for i, st in enumerate(states):
            ens.upload(i, *st)
        # mixed tilings: every shape and several z chunks in one launch
        for i in range(len(DIMS)):
            ens.set_tiling(i, i % 3, (8, 16, 32)[i % 3])
        ens.step(5)
        got = [ens.download(i) for i in range(len(DIMS))]
    for i, st in enumerate(states):
        want = O.run(*st, 5, bc=bc)
        for g, w in zip(got[i], want[:3]):
            np.testing.assert_array_equal(g, w, err_msg=f"member {i} {DIMS[i]}")


@pytest.mark.parametrize("bc", [S.BC_PERIODIC, S.BC_NOFLUX])
def test_ensemble_general_constants_bitwise_vs_oracle(bc):
    """AS != AL (the general FP64 kernels): members of uneven size, one and
    two steps per launch, bitwise the oracle run with the same constants."""
    dims = [(20, 16, 12), (33, 17, 65), (16, 32, 48)]
    states = [_member_state(d, 30 + i) for i, d in enumerate(dims)]
    po = O.default_params()
    po.AL = 1.25e9
    pg = S.default_params()
    for f, _ in S.Params._fields_:
        setattr(pg, f, getattr(po, f))
    wants = [O.run(*st, 5, bc=bc, params=po) for st in states]
    for fuse in (1, 2):
        print(f"bc={bc} fuse={fuse}", flush=True)
        with S.Ensemble(dims, bc=bc, precision=S.F64, params=pg) as ens:
            for i, st in enumerate(states):
                ens.upload(i, *st)
            ens.set_fusion(fuse)
            ens.step(5)
            for i in range(len(dims)):
                for g, w in zip(ens.download(i), wants[i][:3]):
                    np.testing.assert_array_equal(g, w, err_msg=f"member {i} fuse={fuse}")


def test_ensemble_fp32_members_bitwise_vs_single_grid():
    states = [_member_state(d, 20 + i) for i, d in enumerate(DIMS)]
    with S.Ensemble(DIMS, precision=S.F32) as ens:
        ens.set_mass_correction(0)
        for i, st in enumerate(states):
            ens.upload(i, *st)
        ens.autotune(reps=1)
        after = [ens.download(i) for i in range(len(DIMS))]
        for i, st in enumerate(states):   # autotuning did not advance the state
            for a, b in zip(after[i], st):
                np.testing.assert_array_equal(a, b.astype(np.float32).astype(np.float64))
        ens.step(7)
        got = [ens.download(i) for i in range(len(DIMS))]
    for i, st in enumerate(states):
        with S.Solver(DIMS[i], precision=S.F32) as sv:
            sv.set_mass_correction(0)
            sv.upload(*st)
            sv.step(7)
            want = sv.download()
        for g, w in zip(got[i], want):
            np.testing.assert_array_equal(g, w, err_msg=f"member {i} {DIMS[i]}")


def test_ensemble_fp32_mass_targets():
    dims = [(48, 48, 48), (64, 64, 64), (32, 32, 32)]
    with S.Ensemble(dims, precision=S.F32) as ens:
        for i, d in enumerate(dims):
            cen, rad = S.spheres(30 + i, 8, d, 3.0, 7.0)
            ens.init_spheres(i, cen, rad)
        targets = [ens.mass_target(i) for i in range(len(dims))]
        ens.step(150)
        for i in range(len(dims)):
            st = ens.reduce(i)
            assert abs(st["sum_c"] - targets[i]) / targets[i] < 1e-9
            assert st["n_nonfinite"] == 0


@pytest.mark.parametrize("bc", [S.BC_PERIODIC, S.BC_NOFLUX])
@pytest.mark.parametrize("prec", [S.F64, S.F32])
def test_ensemble_two_step_equals_single_steps(prec, bc):
    """Two steps per launch (periodic or no-flux members, both two-step
    shapes, several z chunks, FP32 mass shifts every 5 steps so pairs split
    around them): every member bitwise equal to the one-step-per-launch
    ensemble after an odd number of steps."""
    print(f"prec={prec} bc={bc}", flush=True)
    dims = DIMS + [(1, 1, 1), (3, 5, 2)]
    states = [_member_state(d, 40 + i) for i, d in enumerate(dims)]

    def run(fuse):
        ens = S.Ensemble(dims, bc=bc, precision=prec)
        try:
            for i, st in enumerate(states):
                ens.upload(i, *st)
                if fuse:
                    ens.set_tiling2(i, i % 2, (1, 3, 32)[i % 3])
            if prec == S.F32:
                ens.set_mass_correction(5)
            ens.set_fusion(2 if fuse else 1)
            assert ens.get_fusion() == (2 if fuse else 1)
            ens.step(17)
            return [ens.download(i) for i in range(len(dims))]
        finally:
            ens.close()

    a, b = run(False), run(True)
    for i in range(len(dims)):
        for x, y in zip(a[i], b[i]):
            np.testing.assert_array_equal(x, y)


def test_config4_member_list_edges_bitwise_vs_single_grids():
    """The configured ensemble itself (256 members, edges drawn from {48, 64,
    96, 128}, solid fraction U[0.1, 0.4], radii U[3, 10], FP32, seed 4):
    after 6 steps (two per launch), one member of each edge equals a
    single-grid FP32 context stepped from the same initial condition, bit
    for bit."""
    from phasefield_coarsening_amd import surrogate as SG
    mem = SG.ensemble_members(256, 4)
    first = {}
    for i, (dims, _, _) in enumerate(mem):
        first.setdefault(dims[0], i)
    assert sorted(first) == [48, 64, 96, 128]
    with S.Ensemble([m[0] for m in mem], precision=S.F32) as ens:
        for i, (_, cen, rad) in enumerate(mem):
            ens.init_spheres(i, cen, rad)
        ens.step(6)
        assert ens.get_fusion() == 2
        got = {e: ens.download(i) for e, i in first.items()}
    for e, i in first.items():
        dims, cen, rad = mem[i]
        print(f"edge {e}: member {i}, {len(rad)} spheres", flush=True)
        with S.Solver(dims, precision=S.F32) as sv:
            sv.init_spheres(cen, rad)
            sv.step(6)
            want = sv.download()
        for g, w in zip(got[e], want):
            np.testing.assert_array_equal(g, w, err_msg=f"edge {e}")
