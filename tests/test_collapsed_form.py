"""The collapsed AS == AL form that the FP32 and fast FP64 kernels evaluate
(tests/fp32_ref.py restates it) against the oracle's increments, which are
bitwise the reference's (phasefield.c:324-384).

With AS == AL the reference's chemistry reduces exactly (in real
arithmetic) to cS - cL = cS0 - cL0, fS - fL = CS - CL and mu = A u; the
reference still forms fS and fL (~1e8 each) and subtracts them, so the two
evaluations differ by rounding only: ~1e-16 of the free energies, a few
1e-15 of the largest increment.  A dropped term, a wrong sign, a missing
factor 2 or a dropped gradient term is 2e-9 or more of it (the gradient
terms are scaled by -L dt ~ 5e-11).  CPU only.
"""
import numpy as np
import pytest

from oracle import pyoracle as O
from tests import fp32_ref


def _state(shape, seed):
    rng = np.random.default_rng(seed)
    s = rng.uniform(0.0, 1.0, shape)
    l = np.clip(1.0 - s + rng.uniform(-0.1, 0.1, shape), 0.0, 1.0)
    c = 0.1 + 0.8 * s + rng.uniform(-0.05, 0.05, shape)
    return s, l, c


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("shape", [(8, 9, 10), (16, 16, 16)])
def test_collapsed_increments_equal_the_reference_to_rounding(shape):
    s, l, c = _state(shape, 11)
    k = fp32_ref.constants(O.default_params(), dtype=np.float64)
    want = O.increments(s, l, c)
    got = fp32_ref.increments64(s, l, c, k)
    for g, w in zip(got, want):
        assert _rel(g, w) < 1e-14  # observed ~6e-16


def test_planted_mistakes_break_the_collapsed_form():
    """Each mistake planted in the restatement (not in any kernel) moves the
    increments far beyond the rounding bar above."""
    s, l, c = _state((10, 10, 10), 5)
    p = O.default_params()
    want = O.increments(s, l, c)
    base = fp32_ref.constants(p, dtype=np.float64)
    for name, factor in (("b0", 0.5), ("b1", -1.0), ("d0", 1.0 + 1e-6), ("tmp6", 1.0 + 1e-6),
                         ("a1", 0.0), ("a2", 0.0)):
        k = dict(base)
        k[name] = k[name] * factor
        got = fp32_ref.increments64(s, l, c, k)
        assert max(_rel(g, w) for g, w in zip(got, want)) > 1e-12, name  # observed >= 2.5e-9 (a1)


def test_float32_evaluation_stays_within_the_forms_rounding():
    """step32 (the FP32 kernel's operation order in float32) against step64
    on a sphere interface: etaS and etaL within one float32 ulp, c within
    ten (the rounding of u and of the flux sum; ~8 observed)."""
    shape = (16, 16, 16)
    z, y, x = np.meshgrid(*(np.arange(n) for n in shape), indexing="ij")
    d = np.sqrt((z - 8.3) ** 2 + (y - 7.1) ** 2 + (x - 8.8) ** 2) - 4.0
    s = 0.5 * (1 - np.tanh(d / 2))
    s, l, c = (a.astype(np.float32) for a in (s, 1 - s, 0.1 + 0.8 * s))
    k = fp32_ref.constants(O.default_params())
    a, b = fp32_ref.step32(s, l, c, k), fp32_ref.step64(s, l, c, k)
    for p, q, bound in zip(a, b, (1.0, 1.0, 10.0)):
        assert (np.abs(p.astype(np.float64) - q) / fp32_ref.ulp32(q)).max() <= bound
