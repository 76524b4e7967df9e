"""GPU parity: the gfx950 kernels, called through the C ABI (libppf.so),
against the reference's own outputs (golden fixtures) and the CPU oracle.

The FP64 kernels keep the reference's operand order with contraction off,
so the bar here is bitwise equality
(max|d| <= 1e-10 per field after 1000 steps) is asserted as well where it
is the stated contract.
"""
import numpy as np
import pytest

from oracle import pyoracle as O
from phasefield_coarsening_amd import solver as S

pytestmark = pytest.mark.gpu

KERNELS = [S.KERNEL_TILED, S.KERNEL_NAIVE]
GOLDEN = ["p16_s2", "p16_s10", "p16_s100", "p16_s1000", "pnc_s10", "pnc_s100", "n16_s10",
          "nnc_s100", "p16compat_s10", "nnc_compat_s10", "pnc_conly"]


def _bc(m):
    return S.BC_PERIODIC if m["bc"] == "periodic" else S.BC_NOFLUX


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("case", GOLDEN)
def test_bitwise_vs_reference_outputs(golden, case, kernel):
    man, arrays = golden
    m, z = man["cases"][case], arrays[case]
    print(f"case={case} kernel={kernel}", flush=True)
    with S.Solver(z["s0"].shape, bc=_bc(m), kernel=kernel) as sv:
        sv.upload(z["s0"], z["l0"] if m["etaL"] == "read" else None, z["c0"])
        done = 0
        for k, nsw in enumerate(m["sweeps_before_output"], start=1):
            while done < nsw:
                if done < m["n_iter_init"]:
                    sv.step(1)
                else:
                    sv.step_c_only(1)
                done += 1
            s, l, c = sv.download()
            np.testing.assert_array_equal(s, z[f"s{k}"])
            np.testing.assert_array_equal(l, z[f"l{k}"])
            np.testing.assert_array_equal(c, z[f"c{k}"])


SHAPES = [(1, 1, 1), (2, 3, 5), (7, 9, 70), (20, 16, 12), (33, 17, 65), (9, 130, 31),
          (5, 8, 128), (16, 64, 64)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("bc", [S.BC_PERIODIC, S.BC_NOFLUX])
def test_bitwise_vs_oracle_ragged_shapes(shape, bc):
    """Sizes that are not multiples of any tile, down to 1x1x1."""
    print(f"shape={shape} bc={bc}", flush=True)
    rng = np.random.default_rng(hash(shape) % 2**32)
    s = rng.uniform(0.05, 0.95, shape)
    l = np.clip(1 - s + rng.uniform(-0.1, 0.1, shape), 0.0, 1.0)
    c = 0.1 + 0.8 * s + rng.uniform(-0.02, 0.02, shape)
    want = O.run(s, l, c, 3, bc=bc)
    for kernel in KERNELS:
        with S.Solver(shape, bc=bc, kernel=kernel) as sv:
            sv.upload(s, l, c)
            sv.step(3)
            got = sv.download()
        for g, w in zip(got, want[:3]):
            np.testing.assert_array_equal(g, w)


@pytest.mark.parametrize("c_only", [False, True])
def test_report_sums_match_oracle(golden, c_only):
    """ppf_step_report returns the sums the reference prints at a save
    (phasefield.c:228-240), summed in a different (tree) order."""
    _, arrays = golden
    z = arrays["pnc_s10"]
    ds, dl, dc = O.increments(z["s0"], z["l0"], z["c0"], c_only=c_only)
    with S.Solver(z["s0"].shape) as sv:
        sv.upload(z["s0"], z["l0"], z["c0"])
        got = sv.step_report(c_only=c_only)
    np.testing.assert_allclose(got[0], (dc * dc).sum(), rtol=1e-13)
    if not c_only:
        np.testing.assert_allclose(got[1], (ds * ds).sum(), rtol=1e-13)
        np.testing.assert_allclose(got[2], (dl * dl).sum(), rtol=1e-13)


def _spheres_state(shape, n, rmin, rmax, seed):
    cen, rad = S.spheres(seed, n, shape, rmin, rmax)
    with S.Solver(shape) as sv:
        sv.init_spheres(cen, rad)
        return sv.download(), cen, rad


def test_sphere_initial_condition_matches_numpy():
    shape = (24, 20, 28)
    (s, l, c), cen, rad = _spheres_state(shape, 6, 2.0, 5.0, 3)
    z, y, x = np.meshgrid(*(np.arange(n) for n in shape), indexing="ij")
    d = np.full(shape, 1e30)
    for (cz, cy, cx), R in zip(cen, rad):
        dz = np.abs(z - cz); dz = np.minimum(dz, shape[0] - dz)
        dy = np.abs(y - cy); dy = np.minimum(dy, shape[1] - dy)
        dx = np.abs(x - cx); dx = np.minimum(dx, shape[2] - dx)
        d = np.minimum(d, np.sqrt(dz * dz + dy * dy + dx * dx) - R)
    want = 0.5 * (1 - np.tanh(d / 2))
    np.testing.assert_allclose(s, want, rtol=0, atol=1e-15)
    np.testing.assert_allclose(l, 1 - want, rtol=0, atol=1e-15)
    np.testing.assert_allclose(c, 0.1 + 0.8 * want, rtol=0, atol=1e-15)


def test_sphere_initial_condition_culled_blocks():
    """A grid long enough in z (and x) that blocks of 16^3 sites far from
    every sphere get no candidates: the values still match the full
    evaluation, and sites farther than 60 cells from every surface are 0
    exactly (tanh(d/2) rounds to 1 there)."""
    shape = (300, 24, 170)
    cen = np.array([[20.5, 10.0, 30.0], [150.0, 3.5, 160.0], [290.0, 20.0, 85.0]])
    rad = np.array([5.0, 7.5, 4.0])
    with S.Solver(shape) as sv:
        sv.init_spheres(cen, rad)
        s, l, c = sv.download()
    z, y, x = np.meshgrid(*(np.arange(n) for n in shape), indexing="ij")
    d = np.full(shape, 1e30)
    for (cz, cy, cx), R in zip(cen, rad):
        dz = np.abs(z - cz); dz = np.minimum(dz, shape[0] - dz)
        dy = np.abs(y - cy); dy = np.minimum(dy, shape[1] - dy)
        dx = np.abs(x - cx); dx = np.minimum(dx, shape[2] - dx)
        d = np.minimum(d, np.sqrt(dz * dz + dy * dy + dx * dx) - R)
    want = 0.5 * (1 - np.tanh(d / 2))
    np.testing.assert_allclose(s, want, rtol=0, atol=1e-15)
    np.testing.assert_allclose(c, 0.1 + 0.8 * want, rtol=0, atol=1e-15)
    far = d > 60.0
    assert far.sum() > 0.3 * far.size
    assert (s[far] == 0.0).all() and (l[far] == 1.0).all()


def test_config1_64cubed_1000_steps_vs_oracle():
    """64^3 periodic, 40 spheres with R ~ U[4, 8], 1000 explicit steps: the
    FP64 kernel against the oracle, max|d| <= 1e-10 per field (observed:
    bitwise), and total solute conserved to 1e-12 relative."""
    shape = (64, 64, 64)
    (s0, l0, c0), _, _ = _spheres_state(shape, 40, 4.0, 8.0, 1)
    want = O.run(s0, l0, c0, 1000)
    with S.Solver(shape) as sv:
        sv.upload(s0, l0, c0)
        st0 = sv.reduce()
        sv.step(1000)
        st1 = sv.reduce()
        got = sv.download()
    for g, w in zip(got, want[:3]):
        assert np.abs(g - w).max() <= 1e-10
        np.testing.assert_array_equal(g, w)
    assert abs(st1["sum_c"] - st0["sum_c"]) / abs(st0["sum_c"]) <= 1e-12
    np.testing.assert_allclose(st1["sum_c"], got[2].sum(), rtol=1e-13)


def test_stats_match_numpy():
    shape = (20, 24, 36)
    (s, l, c), _, _ = _spheres_state(shape, 5, 3.0, 6.0, 9)
    with S.Solver(shape) as sv:
        sv.upload(s, l, c)
        sv.step(5)
        st = sv.reduce()
        s, l, c = sv.download()
    h = s * s / (s * s + l * l)
    np.testing.assert_allclose(st["sum_c"], np.sum(c), rtol=1e-14)
    np.testing.assert_allclose(st["sum_h"], h.sum(), rtol=1e-12)
    np.testing.assert_allclose(st["sum_h1h"], (h * (1.0 - h)).sum(), rtol=1e-12)
    np.testing.assert_allclose(st["sum_s"], s.sum(), rtol=1e-12)
    assert st["n_solid"] == (s > 0.5).sum()
    assert st["min_c"] == c.min() and st["max_c"] == c.max()
    assert st["n_sites"] == s.size and st["n_nonfinite"] == 0


def test_full_size_512_properties():
    """At the 512^3 workload size, where the oracle is too slow to compare
    every step: the tiled kernel equals the independent naive kernel bit for
    bit, total solute is conserved to 1e-12, and a periodic shift of the
    input shifts the output (bitwise)."""
    shape = (512, 512, 512)
    cen, rad = S.spheres(2, 2000, shape, 6.0, 14.0)
    with S.Solver(shape, kernel=S.KERNEL_TILED) as a:
        a.init_spheres(cen, rad)
        s0, l0, c0 = a.download()
        st0 = a.reduce()
        a.step(20)
        st1 = a.reduce()
        ga = a.download()
    assert abs(st1["sum_c"] - st0["sum_c"]) / st0["sum_c"] <= 1e-12
    with S.Solver(shape, kernel=S.KERNEL_NAIVE) as b:
        b.upload(s0, l0, c0)
        b.step(20)
        gb = b.download()
    for x, y in zip(ga, gb):
        np.testing.assert_array_equal(x, y)
    del gb
    shift = (37, 200, 101)
    with S.Solver(shape) as r:
        r.upload(*(np.roll(f, shift, axis=(0, 1, 2)) for f in (s0, l0, c0)))
        r.step(20)
        gr = r.download()
    for x, y in zip(ga, gr):
        np.testing.assert_array_equal(np.roll(x, shift, axis=(0, 1, 2)), y)


@pytest.mark.parametrize("bc", [S.BC_PERIODIC, S.BC_NOFLUX])
@pytest.mark.parametrize("shape", [(40, 36, 44), (17, 64, 9)])
def test_particles_match_cpu_labelling(shape, bc):
    from tests.ccl_ref import component_sizes, mean_radius
    print(f"particles shape={shape} bc={bc}", flush=True)
    cen, rad = S.spheres(11, 25, shape, 1.5, 4.5)
    with S.Solver(shape, bc=bc) as sv:
        sv.init_spheres(cen, rad)
        sv.step(30)
        s, _, _ = sv.download()
        st, sizes = sv.particles(0.5)
    want = component_sizes(s > 0.5, periodic=(bc == S.BC_PERIODIC))
    assert st["n_particles"] == len(want)
    np.testing.assert_array_equal(sizes, want)
    assert st["solid_sites"] == (s > 0.5).sum()
    assert abs(st["mean_radius"] - mean_radius(want)) < 1e-12


@pytest.mark.parametrize("precision", [S.F64, S.F32])
def test_autotune_keeps_state_and_results(precision, tmp_path, monkeypatch):
    """ppf_autotune times trial sweeps without advancing the state, and the
    tiling it picks gives the same result as the default (FP64: bitwise,
    also against the oracle)."""
    from phasefield_coarsening_amd import tuning
    monkeypatch.setenv("PPF_TUNING_CACHE", str(tmp_path / "t.json"))
    shape = (40, 48, 72)
    (s0, l0, c0), _, _ = _spheres_state(shape, 6, 3.0, 7.0, 8)
    with S.Solver(shape, precision=precision) as a:
        a.upload(s0, l0, c0)
        before = a.download()
        t = tuning.tuned(a, reps=1)
        after = a.download()
        for x, y in zip(before, after):
            np.testing.assert_array_equal(x, y)
        a.step(4)
        got = a.download()
    assert tuning.load()[tuning.key(shape, precision, S.BC_PERIODIC)]["tx"] == t["tx"]
    with S.Solver(shape, precision=precision) as b:
        b.upload(s0, l0, c0)
        b.step(4)
        ref = b.download()
    for x, y in zip(got, ref):
        np.testing.assert_array_equal(x, y)
    if precision == S.F64:
        want = O.run(s0, l0, c0, 4)
        for x, y in zip(got, want[:3]):
            np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("bc", [S.BC_PERIODIC, S.BC_NOFLUX])
def test_tilings_and_march_variants_bitwise(bc):
    """Every march variant of the FP64 tiled kernel (order bits: 4 the
    unpipelined march, 8 the column loaded three planes ahead, 2 the XCD
    order) on z chunks of 1, 2, 3 and 5 planes (the ends of the load
    pipeline) and a ragged grid: bitwise the oracle."""
    shape = (13, 20, 70)
    print(f"tilings bc={bc}", flush=True)
    rng = np.random.default_rng(5)
    s = rng.uniform(0.05, 0.95, shape)
    l = np.clip(1 - s + rng.uniform(-0.1, 0.1, shape), 0.0, 1.0)
    c = 0.1 + 0.8 * s + rng.uniform(-0.02, 0.02, shape)
    want = O.run(s, l, c, 3, bc=bc)
    with S.Solver(shape, bc=bc) as sv:
        sv.set_fusion(1)  # single steps only: periodic grids pair steps by default
        for tile in [(64, 8), (16, 16)]:
            for tz in (1, 2, 3, 5):
                for order in (0, 4, 8, 10):
                    sv.upload(s, l, c)
                    sv.set_tiling(tile[0], tile[1], tz, order)
                    sv.step(3)
                    got = sv.download()
                    for g, w in zip(got, want[:3]):
                        np.testing.assert_array_equal(g, w, err_msg=f"{tile} tz={tz} order={order}")


@pytest.mark.parametrize("bc", [S.BC_PERIODIC, S.BC_NOFLUX])
def test_general_constants_bitwise_vs_oracle(bc):
    """Constants with AS != AL select the general FP64 kernels (the two
    divides by AS+(AL-AS)*h of phasefield.c:331 and :351 that the
    reference's own constants, AS == AL, fold away): naive, single steps
    under every compiled march variant and two steps per launch under every
    two-step variant, then the c-only sweep and the reported sums, on a
    ragged grid, bitwise the oracle run with the same constants (AL changed,
    every other constant the reference's)."""
    shape = (13, 20, 37)
    print(f"AS != AL bc={bc}", flush=True)
    rng = np.random.default_rng(23)
    s = rng.uniform(0.05, 0.95, shape)
    l = np.clip(1 - s + rng.uniform(-0.1, 0.1, shape), 0.0, 1.0)
    c = 0.1 + 0.8 * s + rng.uniform(-0.02, 0.02, shape)
    po = O.default_params()
    po.AL = 1.25e9
    want = O.run(s, l, c, 7, bc=bc, params=po)
    pg = S.default_params()
    for f, _ in S.Params._fields_:
        setattr(pg, f, getattr(po, f))
    runs = [(S.KERNEL_NAIVE, 1, None, None)]
    for tile in [(64, 8), (16, 16), (32, 16)]:
        for tz, order in ((1, 0), (3, 4), (5, 8), (32, 10)):
            runs.append((S.KERNEL_TILED, 1, (tile[0], tile[1], tz, order), None))
    tiles2 = [(16, 16), (32, 8)] if bc == S.BC_NOFLUX else [(16, 16), (32, 8), (64, 8), (32, 16)]
    for tile in tiles2:
        for order in (0, 2, 4, 6, 8, 10):
            runs.append((S.KERNEL_TILED, 2, None, (tile[0], tile[1], 3, order)))
    for kernel, fuse, tiling, tiling2 in runs:
        with S.Solver(shape, bc=bc, kernel=kernel, params=pg) as sv:
            if kernel == S.KERNEL_TILED:
                if fuse == 1:
                    sv.set_fusion(1)
                    sv.set_tiling(*tiling)
                else:
                    try:
                        sv.set_fusion(2, *tiling2)
                    except S.PpfError:
                        continue  # no variant of this kind for this tile
            sv.upload(s, l, c)
            sv.step(7)
            got = sv.download()
        for g, w in zip(got, want[:3]):
            np.testing.assert_array_equal(g, w, err_msg=f"{kernel} {fuse} {tiling} {tiling2}")
    # c-only sweeps after four full ones (phasefield.c:214-218), and the sums
    # of one reported sweep of each kind
    want_c = O.run(s, l, c, 7, bc=bc, n_iter_init=4, params=po)
    want_r = O.run(s, l, c, 1, bc=bc, params=po)[3]
    want_rc = O.run(s, l, c, 1, bc=bc, n_iter_init=0, params=po)[3]
    for kernel in (S.KERNEL_NAIVE, S.KERNEL_TILED):
        with S.Solver(shape, bc=bc, kernel=kernel, params=pg) as sv:
            sv.upload(s, l, c)
            sv.step(4)
            sv.step_c_only(3)
            got = sv.download()
            for g, w in zip(got, want_c[:3]):
                np.testing.assert_array_equal(g, w, err_msg=f"c-only kernel={kernel}")
            sv.upload(s, l, c)
            r = sv.step_report()
            np.testing.assert_allclose(r, want_r, rtol=1e-13)
            sv.upload(s, l, c)
            rc = sv.step_report(c_only=True)
            np.testing.assert_allclose(rc[0], want_rc[0], rtol=1e-13)


def test_oversize_grid_is_out_of_memory():
    """A grid whose first field alone exceeds the device's 288 GB (3400^3
    FP64 = 314 GB): ppf_create returns PPF_ENOMEM with a message, and the
    device stays usable."""
    with pytest.raises(S.PpfError) as e:
        S.Solver((3400, 3400, 3400))
    assert e.value.code == -2 and "hipMalloc" in str(e.value)
    with S.Solver((8, 8, 8)) as sv:
        sv.upload(*(np.full((8, 8, 8), v) for v in (0.5, 0.5, 0.4)))
        sv.step(2)
        assert np.isfinite(sv.download()[0]).all()


def test_compare_and_converting_copy():
    """ppf_compare against numpy, and ppf_copy_state between precisions:
    FP64 -> FP32 rounds to nearest and takes the FP64 total of c as the
    solute target; FP32 -> FP64 widens exactly."""
    import math
    shape = (20, 16, 12)
    rng = np.random.default_rng(4)
    s = rng.uniform(0.0, 1.0, shape)
    l = 1.0 - s
    c = rng.uniform(0.1, 0.9, shape)
    with S.Solver(shape) as a, S.Solver(shape, precision=S.F32) as b, S.Solver(shape) as w:
        a.upload(s, l, c)
        b.copy_from(a)
        got = b.download()
        for g, x in zip(got, (s, l, c)):
            np.testing.assert_array_equal(g, x.astype(np.float32).astype(np.float64))
        target, _ = b.mass_target()
        assert abs(target - math.fsum(c.ravel())) <= 1e-15 * abs(target)
        d = a.compare(b)
        for i, x in enumerate((s, l, c)):
            e = x - x.astype(np.float32).astype(np.float64)
            assert d["max_abs"][i] == np.abs(e).max()
            np.testing.assert_allclose(d["rms"][i], np.sqrt(np.mean(e * e)), rtol=1e-12)
        w.copy_from(b)
        for g, x in zip(w.download(), got):
            np.testing.assert_array_equal(g, x)
        assert w.compare(b)["max_abs"] == (0.0, 0.0, 0.0)
