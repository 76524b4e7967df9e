"""ppf_sweep_report_noalloc (one sweep of caller-owned device arrays, its
partial sums in the library's fixed scratch array, nothing allocated)
against ppf_step_report on a context of the same grid: the new fields bit
for bit and the three sums bit for bit (same kernel, same tiling, same
reduction order), on two calls in a row, the second on a smaller grid, for
the full and the c-only sweep and both boundary conditions.  Device buffers
come from torch (plumbing only); every computation goes through the C ABI."""
import numpy as np
import pytest

from phasefield_coarsening_amd import solver as S

pytestmark = pytest.mark.gpu


def _state(shape, seed):
    rng = np.random.default_rng(seed)
    s = rng.uniform(0.05, 0.95, shape)
    l = np.clip(1 - s + rng.uniform(-0.1, 0.1, shape), 0.0, 1.0)
    c = 0.1 + 0.8 * s + rng.uniform(-0.02, 0.02, shape)
    return s, l, c


def _context_report(shape, state, bc, c_only):
    with S.Solver(shape, bc=bc) as sv:
        sv.set_fusion(1)
        sv.upload(*state)
        sums = sv.step_report(c_only=c_only)
        return sv.download(), sums


def _noalloc(shape, state, bc, c_only):
    import torch
    dev = torch.device("cuda:0")
    ins = [torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in state]
    outs = [torch.full(shape, -7.0, dtype=torch.float64, device=dev) for _ in range(3)]
    torch.cuda.synchronize()
    sums = S.sweep_report_noalloc(*(t.data_ptr() for t in ins), *(t.data_ptr() for t in outs),
                                  shape, bc=bc, c_only=c_only)
    got = [t.cpu().numpy() for t in outs]
    if c_only:  # only c is written; etaS and etaL stay the input's
        assert (got[0] == -7.0).all() and (got[1] == -7.0).all()
        got[0], got[1] = state[0], state[1]
    return got, sums


@pytest.mark.parametrize("c_only", [False, True])
@pytest.mark.parametrize("bc", [S.BC_PERIODIC, S.BC_NOFLUX])
def test_noalloc_equals_step_report_twice(bc, c_only):
    for shape, seed in (((40, 36, 130), 1), ((9, 20, 33), 2)):
        print(f"noalloc shape={shape} bc={bc} c_only={c_only}", flush=True)
        st = _state(shape, seed)
        want, wsums = _context_report(shape, st, bc, c_only)
        got, gsums = _noalloc(shape, st, bc, c_only)
        for g, w in zip(got, want):
            np.testing.assert_array_equal(g, w)
        assert gsums == wsums, (gsums, wsums)


def test_noalloc_rejects_a_plane_beyond_the_scratch():
    import torch
    buf = torch.zeros(8, dtype=torch.float64, device="cuda:0")
    p = buf.data_ptr()
    with pytest.raises(S.PpfError) as e:
        S.sweep_report_noalloc(p, p, p, p, p, p, (1, 8192, 8192))
    assert e.value.code == -1 and "scratch" in str(e.value)
