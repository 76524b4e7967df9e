"""bench.py keeps its contract: one JSON line with the metric, the value and
its context, a roofline object whose fraction is achieved / peak, and a
refusal (exit status 2, with a message) for more than one GPU."""
import json
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(REPO, "bench.py")


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, BENCH, "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


@pytest.mark.gpu
def test_json_line_keys_and_roofline():
    print("bench 64^3, 4 steps", flush=True)
    r = subprocess.run([sys.executable, "-u", BENCH, "--size", "64", "--steps", "4", "--warmup", "2",
                        "--no-cpu", "--no-extras", "--no-copy"], capture_output=True, text=True,
                       timeout=300, cwd=REPO)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [l for l in r.stdout.splitlines() if l.strip()]
    assert len(lines) == 1
    out = json.loads(lines[0])
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
                "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config",
                "roofline"):
        assert key in out, key
    assert out["n_gpus"] == 1 and out["steps"] == 4 and out["warmup"] == 2
    assert out["value"] > 0 and out["unit"] == "GLUPS" and out["dtype"] == "f64"
    assert out["higher_is_better"] is True and out["vs_baseline"] is None
    assert "workload" in out["config"]
    rf = out["roofline"]
    for key in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert key in rf, key
    assert rf["bound"] in ("hbm", "mfma") and rf["unit"] == "GB/s" and rf["peak"] == 8000.0
    assert abs(rf["frac"] - rf["achieved"] / rf["peak"]) < 1e-3
    # achieved counts 48 B per site update (FP64) times the updates of one
    # launch; moved counts the state read and written once per launch, so a
    # two-step launch moves half the algorithmic bytes of its two updates
    spl = rf["steps_per_launch"]
    assert rf["bytes_per_site_update"] == 48 and spl in (1, 2)
    assert abs(rf["moved_GBs"] * spl - rf["achieved"]) < 0.01 * rf["achieved"] + 1
    # kernel time (HIP events) is within the wall time of the same steps
    assert rf["achieved"] >= 0.95 * out["value"] * 48
    assert "fast" in out["config"]["f64_form"]
    assert out["f64_bitwise"]["GLUPS"] > 0 and out["f64_single_step"]["GLUPS"] > 0
