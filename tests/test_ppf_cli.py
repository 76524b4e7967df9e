"""The ppf program's argument handling (no GPU needed: these paths end
before a device is opened)."""
import os
import subprocess

import numpy as np

from phasefield_coarsening_amd import _lib, h5

PPF = os.path.join(_lib.REPO_DIR, "ppf")


def test_usage_without_filename():
    r = subprocess.run([PPF], capture_output=True, text=True)
    # phasefield.c:122-124: usage on stderr, return -1
    assert r.returncode == 255
    assert r.stderr == "Usage:\nppf filename\n"


def test_usage_with_two_filenames():
    r = subprocess.run([PPF, "a.h5", "b.h5"], capture_output=True, text=True)
    assert r.returncode == 255 and r.stderr.startswith("Usage:")


def test_missing_input_is_an_error(tmp_path):
    r = subprocess.run([PPF, "nope.h5"], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 1
    assert r.stderr.splitlines()[0] == "The job name is nope_"
    assert "cannot open" in r.stderr


def test_bad_option_is_rejected(tmp_path):
    h5.write(tmp_path / "x.h5", *(np.zeros((2, 2, 2)),) * 3)
    r = subprocess.run([PPF, "--bc=sideways", "x.h5"], cwd=tmp_path, capture_output=True,
                       text=True)
    assert r.returncode == 255 and "bad option" in r.stderr
