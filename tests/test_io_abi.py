"""Host-side tests: the HDF5 file contract, output naming, and that the
native libraries load and export every symbol their headers declare.  No GPU
is needed: nothing here launches a kernel."""
import ctypes
import os
import re

import numpy as np
import pytest

from phasefield_coarsening_amd import _lib, h5, solver

REPO = _lib.REPO_DIR


def _declared(header):
    text = open(os.path.join(REPO, "include", header)).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(ppfi?o?_\w+)\s*\(", text, re.M)))


def test_libppf_exports_every_declared_symbol():
    names = _declared("ppf.h")
    assert "ppf_step" in names and "ppf_create" in names
    lib = solver.lib()
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(solver.SIGNATURES)


def test_libppfio_exports_every_declared_symbol():
    names = _declared("ppf_io.h")
    lib = ctypes.CDLL(os.path.join(_lib.PKG_DIR, "libppfio.so"))
    for n in names:
        assert hasattr(lib, n), n
    assert len(names) == 8


def test_default_params_match_oracle():
    from oracle import pyoracle as O
    a, b = solver.default_params(), O.default_params()
    for f, _ in a._fields_:
        assert getattr(a, f) == getattr(b, f), f


def test_sphere_generator_is_seeded_and_in_range():
    c1, r1 = solver.spheres(42, 50, (30, 20, 10), 3.0, 6.0)
    c2, r2 = solver.spheres(42, 50, (30, 20, 10), 3.0, 6.0)
    c3, _ = solver.spheres(43, 50, (30, 20, 10), 3.0, 6.0)
    np.testing.assert_array_equal(c1, c2)
    np.testing.assert_array_equal(r1, r2)
    assert not np.array_equal(c1, c3)
    assert (c1 >= 0).all() and (c1 < np.array([30, 20, 10])).all()
    assert (r1 >= 3).all() and (r1 < 6).all()


def test_create_rejects_bad_arguments():
    with pytest.raises(solver.PpfError) as e:
        solver.Solver((0, 4, 4))
    assert e.value.code == -1
    with pytest.raises(solver.PpfError) as e:
        solver.Solver((4, 4, 4), bc=7)
    assert e.value.code == -1
    # the fast FP64 form and the FP32 form need the collapsed chemistry of
    # AS == AL (phasefield.c:76-77) and a positive liquid mobility
    for kw in (dict(kernel=solver.KERNEL_FAST), dict(precision=solver.F32)):
        p = solver.default_params()
        p.AL = 1.25e9
        with pytest.raises(solver.PpfError) as e:
            solver.Solver((4, 4, 4), params=p, **kw)
        assert e.value.code == -1 and "AS == AL" in str(e.value)
        p = solver.default_params()
        p.ML = 0.0
        with pytest.raises(solver.PpfError) as e:
            solver.Solver((4, 4, 4), params=p, **kw)
        assert e.value.code == -1 and "ML > 0" in str(e.value)
    with pytest.raises(solver.PpfError) as e:
        solver.Solver((4, 4, 4), precision=solver.F32, kernel=solver.KERNEL_FAST)
    assert e.value.code == -1
    with pytest.raises(solver.PpfError) as e:
        solver.Solver((4, 4, 4), kernel=4)
    assert e.value.code == -1
    # the step kernels address a plane with 32-bit byte offsets: an FP64
    # plane of 23200^2 sites (4.3 GB) is refused before any device call
    with pytest.raises(solver.PpfError) as e:
        solver.Solver((1, 23200, 23200))
    assert e.value.code == -1 and "4 GiB" in str(e.value)


def test_h5_roundtrip_and_compat(tmp_path):
    rng = np.random.default_rng(0)
    s, l, c = (rng.random((5, 4, 3)) for _ in range(3))
    p = tmp_path / "x.h5"
    h5.write(p, s, l, c)
    s2, l2, c2 = h5.read(p)
    np.testing.assert_array_equal(s, s2)
    np.testing.assert_array_equal(l, l2)
    np.testing.assert_array_equal(c, c2)
    s3, l3, c3 = h5.read(p, etal_compat=True)
    np.testing.assert_array_equal(s, s3)
    np.testing.assert_array_equal(l3, np.ones_like(l))
    np.testing.assert_array_equal(c, c3)


def test_h5_chunked_layout(tmp_path):
    """The reference's parallel writer chunks each dataset by the local
    block of its decomposition (phasefield.c:576-582): same values, chunked
    storage; the default stays contiguous."""
    rng = np.random.default_rng(1)
    s, l, c = (rng.random((8, 6, 10)) for _ in range(3))
    h5.write(tmp_path / "plain.h5", s, l, c)
    assert h5.layout(tmp_path / "plain.h5") == (0, 0, 0)
    h5.write(tmp_path / "chunked.h5", s, l, c, chunk=(4, 3, 5))
    for name in ("etaS", "etaL", "c"):
        assert h5.layout(tmp_path / "chunked.h5", name) == (4, 3, 5)
    for a, b in zip(h5.read(tmp_path / "chunked.h5"), (s, l, c)):
        np.testing.assert_array_equal(a, b)
    # a chunk that does not divide the grid is legal HDF5 (edge chunks are partial)
    h5.write(tmp_path / "ragged.h5", s, l, c, chunk=(3, 6, 7))
    assert h5.layout(tmp_path / "ragged.h5") == (3, 6, 7)
    np.testing.assert_array_equal(h5.read(tmp_path / "ragged.h5")[2], c)
    with pytest.raises(h5.PpfIOError):
        h5.write(tmp_path / "bad.h5", s, l, c, chunk=(9, 6, 10))


def test_h5_errors(tmp_path):
    with pytest.raises(h5.PpfIOError) as e:
        h5.read(tmp_path / "missing.h5")
    assert e.value.code == -2
    with pytest.raises(h5.PpfIOError):
        h5.write(tmp_path / "no" / "dir" / "x.h5", *(np.zeros((2, 2, 2)),) * 3)


def test_job_stem_and_output_names():
    # phasefield.c:117-120: last three characters -> one '_'
    assert h5.job_stem("run.h5") == "run_"
    assert h5.job_stem("dir/in64.h5") == "dir/in64_"
    assert h5.output_name("run_", 1) == "run_1.h5"
    assert h5.output_name("run_", 120) == "run_120.h5"
    long = "a" * 300 + ".h5"  # the reference overflows char[50] here
    assert h5.job_stem(long) == "a" * 300 + "_"
    with pytest.raises(h5.PpfIOError):
        h5.job_stem("ab")


def test_reader_dims_come_from_etaS(tmp_path):
    s = np.arange(24.0).reshape(2, 3, 4)
    p = tmp_path / "x.h5"
    h5.write(p, s, s + 1, s + 2)
    s2, l2, c2 = h5.read(p)
    assert s2.shape == (2, 3, 4)
    np.testing.assert_array_equal(l2, s + 1)
    np.testing.assert_array_equal(c2, s + 2)


def test_noalloc_sweep_refuses_before_touching_the_device():
    """Argument and size checks of ppf_sweep_report_noalloc come first: a
    plane that needs more column tiles than its scratch holds partial sums
    for, a bad boundary condition or a null pointer is refused with
    PPF_EARG (the dummy addresses below are never dereferenced)."""
    dummy = 4096
    for shape, bc in (((1, 8192, 8192), 0), ((4, 4, 4), 5), ((0, 4, 4), 0)):
        with pytest.raises(solver.PpfError) as e:
            solver.sweep_report_noalloc(dummy, dummy, dummy, dummy, dummy, dummy, shape, bc=bc)
        assert e.value.code == -1
    with pytest.raises(solver.PpfError) as e:
        solver.sweep_report_noalloc(None, dummy, dummy, dummy, dummy, dummy, (4, 4, 4))
    assert e.value.code == -1
