import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def golden():
    """Manifest and arrays of the fixtures made from the compiled reference."""
    import json

    import numpy as np
    d = os.path.join(REPO, "tests", "golden")
    man = json.load(open(os.path.join(d, "manifest.json")))
    arrays = {name: dict(np.load(os.path.join(d, f"{name}.npz"))) for name in man["cases"]}
    return man, arrays
