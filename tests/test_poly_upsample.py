"""The identity behind the polyphase decoder convolution (U-Net kind 6),
in float64 on the CPU: a periodic 3x3x3 convolution of a nearest x2
upsampled tensor equals, at every fine site of parity p = (pz, py, px), a
2x2x2 convolution of the coarse tensor at offsets p - 1 + o (o in {0, 1}^3)
whose weights sum the fine taps d with floor((p + d) / 2) = p - 1 + o.
k_prep_wpoly (csrc/unet.hip) builds the summed weights by this rule; the
GPU tests compare kind 6 with the reference network."""
import itertools

import numpy as np


def _conv_direct(x, w):
    """x [D, H, W, Ci], w [27, Co, Ci], tap = (dz+1)*9 + (dy+1)*3 + (dx+1),
    periodic."""
    out = np.zeros(x.shape[:3] + (w.shape[1],))
    for dz, dy, dx in itertools.product((-1, 0, 1), repeat=3):
        tap = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)
        nb = np.roll(x, (-dz, -dy, -dx), axis=(0, 1, 2))  # nb[z] = x[z + dz]
        out += nb @ w[tap].T
    return out


def _merged(w, p):
    """Summed weights [8 (o), Co, Ci] for output parity p."""
    wp = np.zeros((8,) + w.shape[1:])
    for dz, dy, dx in itertools.product((-1, 0, 1), repeat=3):
        tap = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)
        o = [(pa + d + 2) // 2 - 1 - pa + 1 for pa, d in zip(p, (dz, dy, dx))]
        assert all(v in (0, 1) for v in o)
        wp[o[0] * 4 + o[1] * 2 + o[2]] += w[tap]
    return wp


def test_polyphase_equals_upsample_then_convolve():
    rng = np.random.default_rng(0)
    Dc, Hc, Wc, Ci, Co = 3, 4, 5, 3, 2
    xc = rng.normal(size=(Dc, Hc, Wc, Ci))
    w = rng.normal(size=(27, Co, Ci))
    fine = xc.repeat(2, 0).repeat(2, 1).repeat(2, 2)
    want = _conv_direct(fine, w)
    got = np.zeros_like(want)
    for p in itertools.product((0, 1), repeat=3):
        wp = _merged(w, p)
        for oz, oy, ox in itertools.product((0, 1), repeat=3):
            off = (p[0] - 1 + oz, p[1] - 1 + oy, p[2] - 1 + ox)
            nb = np.roll(xc, tuple(-v for v in off), axis=(0, 1, 2))  # xc[X + off]
            got[p[0]::2, p[1]::2, p[2]::2] += nb @ wp[oz * 4 + oy * 2 + ox].T
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)


def test_each_fine_tap_lands_in_exactly_one_group():
    """Per dimension: parity 0 maps d = -1 to o = 0 and d = 0, +1 to o = 1;
    parity 1 maps d = -1, 0 to o = 0 and d = +1 to o = 1.  So of the eight
    groups of a parity, one holds 1 tap, three hold 2, three hold 4 and one
    holds 8: 27 in all."""
    w = np.ones((27, 1, 1))
    for p in itertools.product((0, 1), repeat=3):
        wp = _merged(w, p)[:, 0, 0]
        assert wp.sum() == 27
        assert sorted(wp.tolist()) == [1, 2, 2, 2, 4, 4, 4, 8]
