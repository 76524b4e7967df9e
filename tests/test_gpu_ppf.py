"""The ppf program end to end: same invocation, log lines, file numbering
and file contents as the reference program (phasefield.c:73-269), checked
against the reference's own runs stored in tests/golden (inputs, outputs,
stdout and stderr)."""
import os
import subprocess

import numpy as np
import pytest

from phasefield_coarsening_amd import _lib, h5

pytestmark = pytest.mark.gpu
PPF = os.path.join(_lib.REPO_DIR, "ppf")
CASES = ["p16_s2", "p16_s10", "p16_s100", "pnc_s10", "n16_s10", "nnc_s100", "p16compat_s10",
         "nnc_compat_s10", "pnc_conly"]


@pytest.mark.parametrize("case", CASES)
def test_ppf_matches_reference_run(golden, case, tmp_path):
    man, arrays = golden
    m, z = man["cases"][case], arrays[case]
    print(f"case={case}", flush=True)
    h5.write(tmp_path / "in.h5", z["s0"], z["l0"], z["c0"])
    args = [PPF, f"--bc={m['bc']}", f"--n-iter={m['n_iter']}", f"--n-save={m['n_save']}",
            f"--n-iter-init={m['n_iter_init']}"]
    if m["etaL"] != "read":
        args.append("--etal-compat")
    r = subprocess.run(args + ["in.h5"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.splitlines() == m["stdout"]
    assert r.stderr.splitlines() == m["stderr"]
    for k, name in enumerate(m["outputs"], start=1):
        s, l, c = h5.read(tmp_path / name)
        np.testing.assert_array_equal(s, z[f"s{k}"])
        np.testing.assert_array_equal(l, z[f"l{k}"])
        np.testing.assert_array_equal(c, z[f"c{k}"])


def test_ppf_output_is_a_restart_input(golden, tmp_path):
    """An output file is a valid input; 4 + 6 sweeps through a restart equal
    10 sweeps in one run (periodic, /etaL read)."""
    _, arrays = golden
    z = arrays["p16_s10"]
    h5.write(tmp_path / "a.h5", z["s0"], z["l0"], z["c0"])
    base = [PPF, "--bc=periodic"]
    subprocess.run(base + ["--n-iter=4", "--n-save=3", "a.h5"], cwd=tmp_path, check=True,
                   capture_output=True, timeout=120)
    os.rename(tmp_path / "a_1.h5", tmp_path / "b.h5")
    subprocess.run(base + ["--n-iter=6", "--n-save=5", "b.h5"], cwd=tmp_path, check=True,
                   capture_output=True, timeout=120)
    s, l, c = h5.read(tmp_path / "b_1.h5")
    np.testing.assert_array_equal(s, z["s1"])
    np.testing.assert_array_equal(c, z["c1"])


@pytest.mark.parametrize("opt,chunk", [("--blocks=2", (8, 8, 8)), ("--chunk=4,16,8", (4, 16, 8))])
def test_ppf_chunked_output(golden, tmp_path, opt, chunk):
    """--blocks=Q writes the file an `mpirun -np Q^3` run of the reference
    writes (datasets chunked by the local block, phasefield.c:576-582);
    --chunk picks any chunk.  Values equal the contiguous run's."""
    _, arrays = golden
    z = arrays["p16_s10"]
    h5.write(tmp_path / "in.h5", z["s0"], z["l0"], z["c0"])
    base = [PPF, "--bc=periodic", "--n-iter=10", "--n-save=9"]
    subprocess.run(base + [opt, "in.h5"], cwd=tmp_path, check=True, capture_output=True,
                   timeout=120)
    assert h5.layout(tmp_path / "in_1.h5") == chunk
    for a, b in zip(h5.read(tmp_path / "in_1.h5"), (z["s1"], z["l1"], z["c1"])):
        np.testing.assert_array_equal(a, b)


def test_ppf_rejects_blocks_that_do_not_divide(golden, tmp_path):
    _, arrays = golden
    z = arrays["pnc_s10"]  # 20 x 16 x 12
    h5.write(tmp_path / "in.h5", z["s0"], z["l0"], z["c0"])
    r = subprocess.run([PPF, "--bc=periodic", "--n-iter=2", "--blocks=3", "in.h5"],
                       cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "--blocks" in r.stderr


def test_ppf_stops_on_threshold(tmp_path):
    """A uniform pure-liquid state has dc = 0, so err = 0 < 1e-12 at the
    first save and the loop ends there (phasefield.c:179, 247)."""
    shape = (4, 5, 6)
    h5.write(tmp_path / "u.h5", np.zeros(shape), np.ones(shape), np.full(shape, 0.3))
    r = subprocess.run([PPF, "--n-save=5", "u.h5"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0
    lines = r.stderr.splitlines()
    assert lines[1] == "5: c=0.0000e+00; S=0.0000e+00; L=0.0000e+00\titer: 5; err: 0.0000e+00"
    assert lines[-1] == "Total steps: 6; error: 0.000000e+00"
    assert r.stdout.splitlines() == ["u_1.h5"]


def test_ppf_unwritable_output_ends_the_run(golden, tmp_path):
    """The first output path is taken by a directory, so the HDF5 write
    fails: ppf names the file and exits with status 1 after printing the
    name (the reference ignores every HDF5 status, phasefield.c:1052-1109,
    and carries on)."""
    _, arrays = golden
    z = arrays["p16_s10"]
    h5.write(tmp_path / "w.h5", z["s0"], z["l0"], z["c0"])
    os.mkdir(tmp_path / "w_1.h5")
    r = subprocess.run([PPF, "--bc=periodic", "--n-iter=4", "--n-save=2", "w.h5"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 1
    assert r.stdout.splitlines() == ["w_1.h5"]
    assert "ppf: w_1.h5:" in r.stderr
    assert not (tmp_path / "w_2.h5").exists()


def test_ppf_fast_kernel_within_1e10(golden, tmp_path):
    """--kernel=fast: the same log lines' format and files, fields within
    1e-10 of the reference's own run (p16_s1000: 1000 periodic sweeps)."""
    man, arrays = golden
    m, z = man["cases"]["p16_s1000"], arrays["p16_s1000"]
    h5.write(tmp_path / "in.h5", z["s0"], z["l0"], z["c0"])
    r = subprocess.run([PPF, "--kernel=fast", f"--bc={m['bc']}", f"--n-iter={m['n_iter']}",
                        f"--n-save={m['n_save']}", "in.h5"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.splitlines() == m["stdout"]
    for k, name in enumerate(m["outputs"], start=1):
        for got, key in zip(h5.read(tmp_path / name), "slc"):
            assert np.abs(got - z[f"{key}{k}"]).max() <= 1e-10


def test_config1_pinned_to_the_reference_run(golden, tmp_path):
    """Config 1 (64^3 periodic, 40 spheres R ~ U[4, 8], 1000 steps) from the
    machine-independent input of tests/golden/det_ic.py: the ppf program's
    output file and the library's FP64 state hash as the reference
    program's own run on the same input (manifest config1_reference_run,
    phasefield.c:178-258, 1052-1109), with the same log lines; the fast
    FP64 kernel lies within 1e-10 of it."""
    from phasefield_coarsening_amd import solver as S
    from tests.golden import det_ic
    man, _ = golden
    m = man["config1_reference_run"]
    shape = tuple(m["shape_zyx"])
    cen, rad = S.spheres(m["sphere_seed"], m["spheres"], shape, *m["radius"])
    s0, l0, c0 = det_ic.sphere_state(shape, cen, rad)
    assert {"s0": det_ic.sha256(s0), "l0": det_ic.sha256(l0),
            "c0": det_ic.sha256(c0)} == m["sha256_input"]
    want = m["sha256_output"]
    h5.write(tmp_path / "in.h5", s0, l0, c0)
    r = subprocess.run([PPF, "--bc=periodic", f"--n-iter={m['n_iter']}",
                        f"--n-save={m['n_save']}", "in.h5"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.splitlines() == m["stdout"]
    assert r.stderr.splitlines() == m["stderr"]
    s1, l1, c1 = h5.read(tmp_path / m["output"])
    assert {"s1": det_ic.sha256(s1), "l1": det_ic.sha256(l1), "c1": det_ic.sha256(c1)} == want
    with S.Solver(shape) as a, S.Solver(shape, kernel=S.KERNEL_FAST) as b:
        a.upload(s0, l0, c0)
        b.upload(s0, l0, c0)
        a.step(m["sweeps"])
        b.step(m["sweeps"])
        ga, gb = a.download(), b.download()
    assert {"s1": det_ic.sha256(ga[0]), "l1": det_ic.sha256(ga[1]),
            "c1": det_ic.sha256(ga[2])} == want
    errs = [float(np.abs(x - y).max()) for x, y in zip(gb, ga)]
    print(f"config 1 vs the reference run: fast FP64 max|d| = {errs}", flush=True)
    assert max(errs) <= 1e-10
    assert abs(gb[2].sum() - m["sum_c_input"]) / m["sum_c_input"] <= 1e-12
