"""The step kernels keep their state in registers: none of them may use
scratch (private) memory, which on gfx950 is a spill to HBM in the inner
z march.  A spill is silent (the results stay bitwise right) but costs a
third of the FP64 two-step kernel's throughput, so it is checked at build
time from the compiler's resource-usage remarks, on the CPU.

One variant is known to spill and is exempt: the FP64 64x16 two-step tile,
bitwise and fast forms (1024 threads need four waves per SIMD, 128 VGPRs),
which the autotuner can time but no default uses."""
import os
import re
import shutil
import subprocess

import pytest

from phasefield_coarsening_amd import build as B

HIPCC = B.HIPCC
EXEMPT = re.compile(r"^_Z8k_tiled2I(?:7PhysF64ILb[01]EE|5PhysCIdE)Li64ELi16ELi1E")


def _scratch_by_kernel(src, tmp_path):
    out = str(tmp_path / "probe.so")
    cmd = [HIPCC, *B.HIP_FLAGS, f"-I{B.INC}", f"-I{B.CSRC}", os.path.join(B.CSRC, src), "-o", out,
           "-Rpass-analysis=kernel-resource-usage"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    names = re.findall(r"remark: Function Name: (\S+)", r.stderr)
    scratch = [int(v) for v in re.findall(r"remark:\s+ScratchSize \[bytes/lane\]: (\d+)", r.stderr)]
    assert len(names) == len(scratch) and names
    return dict(zip(names, scratch))


@pytest.mark.skipif(not shutil.which(HIPCC) and not os.path.exists(HIPCC), reason="no hipcc")
@pytest.mark.parametrize("src, prefixes", [
    ("ppf.hip", ("_Z7k_tiled", "_Z8k_tiled2", "_Z9k_tiled2s")),
    ("ppf_ens.hip", ("_Z10k_ensemble", "_Z11k_ensemble2")),
])
def test_step_kernels_do_not_spill(src, prefixes, tmp_path):
    usage = _scratch_by_kernel(src, tmp_path)
    kernels = {k: v for k, v in usage.items() if k.startswith(prefixes)}
    assert kernels, f"no step kernels found in {src}"
    spilling = {k: v for k, v in kernels.items() if v and not EXEMPT.match(k)}
    assert not spilling, f"step kernels using scratch (bytes/lane): {spilling}"
