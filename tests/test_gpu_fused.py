"""Two steps per launch (k_tiled2): bit for bit the same as single steps, for
every compiled two-step tile, z chunk and register variant, with periodic
and with mirror (no-flux) boundaries, on ragged and tiny periodic grids,
for odd step counts (a single step finishes), in FP64 and in FP32 with mass
shifts that fall between the two steps of a pair; and against the
reference's own outputs."""
import numpy as np
import pytest

from oracle import pyoracle as O
from phasefield_coarsening_amd import solver as S

pytestmark = pytest.mark.gpu

TILES2 = [(64, 8), (32, 8), (32, 16), (64, 16), (16, 16)]
RING_WAVE_TILES = [(16, 16), (32, 8), (32, 16)]
SHAPES = [(1, 1, 1), (2, 3, 5), (7, 9, 70), (20, 16, 12), (33, 17, 65), (9, 130, 31),
          (16, 64, 64)]


def _state(shape, seed):
    rng = np.random.default_rng(seed)
    s = rng.uniform(0.05, 0.95, shape)
    l = np.clip(1 - s + rng.uniform(-0.1, 0.1, shape), 0.0, 1.0)
    c = 0.1 + 0.8 * s + rng.uniform(-0.02, 0.02, shape)
    return s, l, c


def _run(shape, state, nsteps, precision, fuse, tile=(0, 0), tz=0, order=-1, shift=0,
         bc=S.BC_PERIODIC):
    with S.Solver(shape, precision=precision, bc=bc) as sv:
        sv.upload(*state)
        if precision == S.F32:
            sv.set_mass_correction(shift)
        if fuse:
            sv.set_fusion(2, tile[0], tile[1], tz, order)
            assert sv.get_fusion()[0] == 2
        else:
            sv.set_fusion(1)
        sv.step(nsteps)
        return sv.download()


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("tile", TILES2)
def test_fp64_fused_equals_single_steps(shape, tile):
    print(f"shape={shape} tile={tile}", flush=True)
    st = _state(shape, 7)
    want = _run(shape, st, 7, S.F64, False)
    for tz in (1, 3, 32):
        # order bit 2: the three-wave FP64 variant (tiles 32x8 and 16x16);
        # bit 3: the ring in waves of its own (tiles 16x16, 32x8, 32x16)
        for var in (0, 4, 8):
            if var == 8 and tile not in RING_WAVE_TILES:
                continue
            got = _run(shape, st, 7, S.F64, True, tile, tz, order=(tz % 4) | var)
            for a, b in zip(got, want):
                np.testing.assert_array_equal(a, b, err_msg=f"tz={tz} order={(tz % 4) | var}")


@pytest.mark.parametrize("shape", [(20, 16, 12), (33, 17, 65), (16, 64, 64)])
@pytest.mark.parametrize("tile", TILES2)
def test_fp32_fused_equals_single_steps_with_mass_shifts(shape, tile):
    """Mass shift every 5 steps: pairs must split around each shift."""
    print(f"shape={shape} tile={tile}", flush=True)
    st = _state(shape, 11)
    want = _run(shape, st, 23, S.F32, False, shift=5)
    # order bit 3: ring waves; bit 4: loads issued a plane early
    orders = [2] + ([8] if tile in RING_WAVE_TILES else []) + \
        ([18] if tile in [(32, 8), (16, 16), (32, 16), (64, 8)] else [])
    for order in orders:
        got = _run(shape, st, 23, S.F32, True, tile, 8, order=order, shift=5)
        for a, b in zip(got, want):
            np.testing.assert_array_equal(a, b, err_msg=f"order={order}")


NOFLUX_TILES = {S.F64: [(16, 16), (32, 8)], S.F32: [(16, 16), (32, 8), (32, 16)]}


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("tile", NOFLUX_TILES[S.F64])
def test_fp64_noflux_fused_equals_single_steps(shape, tile):
    """Mirror boundaries: ghost ring sites and ghost planes take their
    mirror site's step t+1; z chunks of 1 and 3 planes put chunk ends on the
    grid's first and last planes, and (1, 1, 1) mirrors every face onto the
    one site."""
    print(f"noflux shape={shape} tile={tile}", flush=True)
    st = _state(shape, 17)
    want = _run(shape, st, 7, S.F64, False, bc=S.BC_NOFLUX)
    for tz in (1, 3, 32):
        for order in (0, 2, 1):
            got = _run(shape, st, 7, S.F64, True, tile, tz, order=order, bc=S.BC_NOFLUX)
            for a, b in zip(got, want):
                np.testing.assert_array_equal(a, b, err_msg=f"tz={tz} order={order}")


@pytest.mark.parametrize("shape", [(1, 1, 1), (20, 16, 12), (33, 17, 65), (16, 64, 64)])
@pytest.mark.parametrize("tile", NOFLUX_TILES[S.F32])
def test_fp32_noflux_fused_equals_single_steps_with_mass_shifts(shape, tile):
    print(f"noflux f32 shape={shape} tile={tile}", flush=True)
    st = _state(shape, 19)
    want = _run(shape, st, 23, S.F32, False, shift=5, bc=S.BC_NOFLUX)
    got = _run(shape, st, 23, S.F32, True, tile, 8, order=2, shift=5, bc=S.BC_NOFLUX)
    for a, b in zip(got, want):
        np.testing.assert_array_equal(a, b)


def test_fused_matches_reference_outputs(golden):
    """The golden cases (the reference program's own outputs, periodic and
    no-flux), stepped with pairs."""
    man, arrays = golden
    for case in ("p16_s10", "p16_s100", "p16_s1000", "pnc_s100", "p16compat_s10", "n16_s10",
                 "nnc_s100", "nnc_compat_s10"):
        m, z = man["cases"][case], arrays[case]
        assert m["n_iter_init"] >= m["sweeps_before_output"][-1]
        print(f"case={case}", flush=True)
        bc = S.BC_PERIODIC if m["bc"] == "periodic" else S.BC_NOFLUX
        with S.Solver(z["s0"].shape, bc=bc) as sv:
            sv.set_fusion(2)
            sv.upload(z["s0"], z["l0"] if m["etaL"] == "read" else None, z["c0"])
            done = 0
            for k, nsw in enumerate(m["sweeps_before_output"], start=1):
                sv.step(nsw - done)
                done = nsw
                s, l, c = sv.download()
                np.testing.assert_array_equal(s, z[f"s{k}"])
                np.testing.assert_array_equal(l, z[f"l{k}"])
                np.testing.assert_array_equal(c, z[f"c{k}"])


def test_fused_512_equals_single_steps():
    """The benchmark workload: 512^3, 2000 spheres, 4 steps."""
    shape = (512, 512, 512)
    print("512^3 fused vs single", flush=True)
    cen, rad = S.spheres(2, 2000, shape, 6.0, 14.0)
    with S.Solver(shape) as a, S.Solver(shape) as b:
        a.init_spheres(cen, rad)
        b.init_spheres(cen, rad)
        a.set_fusion(1)
        b.set_fusion(2)
        a.step(4)
        b.step(4)
        for x, y in zip(a.download(), b.download()):
            assert np.array_equal(x, y)


def test_fusion_refused_where_not_offered():
    """The naive kernel has no two-step form; no-flux grids have no 64x8 or
    ring-wave two-step kernel.  A refused call leaves the setting as it was."""
    with S.Solver((8, 8, 8), kernel=S.KERNEL_NAIVE) as sv:
        with pytest.raises(RuntimeError):
            sv.set_fusion(2)
        assert sv.get_fusion()[0] == 1
    with S.Solver((8, 8, 8), bc=S.BC_NOFLUX) as sv:
        before = sv.get_fusion()
        assert before[0] == 2  # two steps per launch by default
        with pytest.raises(RuntimeError):
            sv.set_fusion(2, 64, 8, 0, -1)
        with pytest.raises(RuntimeError):
            sv.set_fusion(2, 32, 16, 0, -1)  # FP64: no mirror kernel for 32x16
        assert sv.get_fusion() == before
