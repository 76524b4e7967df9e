"""The GPU U-Net (BF16 MFMA forward, backward, Adam) against the torch CPU
float64 reference of the same network (tests/unet_ref.py)."""
import numpy as np
import pytest

from phasefield_coarsening_amd import solver as S
from phasefield_coarsening_amd.unet import UNet, layer_shapes
from tests import unet_ref

pytestmark = pytest.mark.gpu


def _params(seed, head=0.05, base=32):
    """He-scaled random weights (the regime the network runs in), small
    random biases, a head of scale `head`."""
    rng = np.random.default_rng(seed)
    out = []
    for name, s in layer_shapes(base):
        if name.startswith("w"):
            out.append(rng.normal(0, np.sqrt(2.0 / (27 * s[2])), int(np.prod(s))))
        elif name == "head_w":
            out.append(rng.normal(0, head, int(np.prod(s))))
        else:
            out.append(rng.normal(0, 0.01, int(np.prod(s))))
    return np.concatenate(out).astype(np.float32)


def _data(seed, shape=(2, 16, 16, 16)):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 1, shape + (3,)).astype(np.float32)
    y = (x + rng.normal(0, 0.05, x.shape)).astype(np.float32)
    return x, y


# convolution kernels: 2 row-blocked, 3 and 4 its fixed-tile variants, 5
# kind 2 with materialized upsampling, 6 (default) kind 2 with polyphase
# decoder convolutions at inference, 7 kind 6 without the z-marching
# polyphase kernel, 1 LDS-tiled, 0 direct
KINDS = [2, 3, 4, 5, 6, 7, 1, 0]


@pytest.mark.parametrize("shape", [(2, 16, 16, 16), (1, 24, 12, 40), (1, 8, 36, 20), (1, 32, 32, 32)])
def test_row_blocked_kinds_agree_bitwise(shape):
    """Kinds 2-5 use different tiles (8 or 4 rows, 32 or 64 channels per
    wave), the z-marching level-1 kernel (2, 5) or not (3, 4), and fused or
    materialized upsampling, but every output sums the same (chunk, dz, dx,
    dy) products in the same order: their forwards are bitwise equal.  At
    32^3 kinds 2 and 5 run the input-plane-stationary kernels (level-1
    layers, and the 32 -> 64 layer), with the pooling and the head in their
    epilogues; kinds 3 and 4 run the row-blocked kernels, k_pool_fwd and
    k_head_fwd."""
    p = _params(5)
    x, _ = _data(6, shape)
    out = {}
    with UNet(seed=3) as net:
        net.set_params(p)
        for kind in (2, 3, 4, 5):
            net.set_conv_kernel(kind)
            out[kind] = net.forward(x)
    for kind in (3, 4, 5):
        np.testing.assert_array_equal(out[kind], out[2])


@pytest.mark.parametrize("shape", [(2, 16, 16, 16), (1, 24, 12, 40)])
def test_fused_upsampling_is_bitwise_the_materialized_one(shape):
    """At inference the decoder convolutions read the upsampled channels from
    the coarse tensor (kind 2); kind 5 runs the same kernels on a
    materialized upsampled copy.  Same values, same sums: bitwise equal."""
    p = _params(2)
    x, _ = _data(3, shape)
    with UNet(seed=3) as net:
        net.set_params(p)
        net.set_conv_kernel(2)
        a = net.forward(x)
        net.set_conv_kernel(5)
        b = net.forward(x)
    np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("shape", [(2, 16, 16, 16), (1, 24, 12, 40), (1, 8, 36, 20), (1, 36, 20, 68)])
def test_zmarching_polyphase_matches_the_row_blocked_one(shape):
    """Kind 6 runs the last decoder convolution (L8) with the eight-wave
    z-marching polyphase kernel (both sources staged in LDS, the K dimension
    split over two waves and their partial sums added), kind 7 with the
    row-blocked polyphase kernel: the same products, summed in another
    association, so L8's outputs differ at most by FP32 rounding before
    their BF16 rounding: now and then a BF16 output lands one unit apart,
    which L9 and the head spread over their neighbourhoods.  The network
    outputs agree to 5e-3 of the head's scale at worst and 5e-5 on average,
    and most of them exactly (a wrong or missing product group would move
    every output by a sizeable fraction of the scale; agreement with the
    float64 reference is checked for every kind below).  The shapes cover tiles larger
    than the grid, ragged x and y (not multiples of the 32 x 8 tile) and z
    chunks shorter than 16 planes."""
    p = _params(11)
    x, _ = _data(12, shape)
    with UNet(seed=3) as net:
        net.set_params(p)
        net.set_conv_kernel(6)
        a = net.forward(x)
        net.set_conv_kernel(7)
        b = net.forward(x)
    scale = np.abs(b - x).max()
    d = np.abs(a - b)
    same = float(np.mean(d == 0))
    print("kind 6 vs 7: max |d| %.3e, mean %.3e, scale %.3e, exactly equal %.5f"
          % (d.max(), d.mean(), scale, same))
    assert d.max() <= 5e-3 * scale
    assert d.mean() <= 5e-5 * scale
    assert same >= 0.9


@pytest.mark.parametrize("base", [32, 64])
@pytest.mark.parametrize("shape", [(2, 16, 16, 16), (1, 24, 12, 40)])
def test_polyphase_decoder_close_to_kind2(shape, base):
    """Kind 6 runs the two decoder convolutions over upsampled channels as
    2x2x2 convolutions of the coarse tensor with summed weights.  It differs
    from kind 2 only by rounding each summed weight to BF16 once instead of
    each tap's weight: the outputs agree to a small fraction of the head's
    scale, and most sites agree exactly after the output rounding."""
    p = _params(4, base=base)
    x, _ = _data(5, shape)
    with UNet(base=base, seed=3) as net:
        net.set_params(p)
        net.set_conv_kernel(2)
        a = net.forward(x)
        net.set_conv_kernel(6)
        b = net.forward(x)
    scale = np.abs(a - x).max()
    d = np.abs(a - b)
    print("kind 6 vs 2: max |d| %.3e, mean %.3e, scale %.3e" % (d.max(), d.mean(), scale))
    assert d.max() < 2e-2 * scale and d.mean() < 2e-3 * scale


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("shape", [(2, 16, 16, 16), (1, 24, 12, 40)])
def test_forward_matches_reference(kind, shape):
    """Every convolution kernel, on a cube and on a grid whose x and y are
    not multiples of the kernels' tiles (coarse levels 6 x 3 x 10, smaller
    than a tile)."""
    p = _params(1)
    x, _ = _data(2, shape)
    with UNet(seed=3) as net:
        net.set_conv_kernel(kind)
        net.set_params(p)
        got = net.forward(x)
    want = unet_ref.predict(p, x, bf16=True)
    d = np.abs(got - want)
    scale = np.abs(want - x).max()
    print("forward: max |d| %.3e, mean |d| %.3e, head scale %.3e" % (d.max(), d.mean(), scale))
    # same bf16 rounding points; FP32 vs FP64 accumulation can flip a bf16
    # rounding here and there
    assert d.max() < 2e-2 * scale and d.mean() < 3e-3 * scale
    # and close to the unrounded float64 network
    assert np.abs(got - unet_ref.predict(p, x, bf16=False)).mean() < 2e-2 * scale


@pytest.mark.parametrize("shape", [(1, 32, 32, 32), (1, 64, 64, 64), (1, 16, 128, 64)])
def test_forward_32cubed_matches_reference(shape):
    """At 32^3 and 64^3 every layer of kinds 2, 6 and 7 runs a z-marching
    kernel where one exists (input-plane-stationary ones at levels 1 and 2,
    the pooling and the head fused into the level-1 layers).  Kind 6 also
    runs the 64 -> 64 layers L3 and L7 with the K dimension split over two
    waves (two partial sums), L6 as a polyphase pass over the upsampled
    channels into FP32 partial sums followed by the skip channels, and (at
    64^3, level 3 of 16^3) L4 and L5 by the same kernel per half of the
    output channels, L5 in two passes through FP32 partial sums.  At 16 x
    128 x 64 (level 2: 8 x 64 x 32) kind 6's polyphase pass of L6 is the
    one whose workgroups share the weights through LDS.  Kind 7
    keeps one running sum per output: a different FP32 association flips a
    BF16 rounding now and then, and the deeper layers spread each flip over
    their neighbourhoods, so kinds 6 and 7 differ slightly almost everywhere
    (mean 3.7e-4 of the head's scale 0.42 observed at 32^3).  The test holds
    each kind to the float64 reference and kind 6 to no more than 1.25 times
    the larger mean error of kinds 2 and 7: a wrong product group would move
    the outputs by a sizeable fraction of the scale."""
    p = _params(1)
    x, _ = _data(2, shape)
    want = unet_ref.predict(p, x, bf16=True)
    scale = np.abs(want - x).max()
    err = {}
    with UNet(seed=3) as net:
        net.set_params(p)
        for kind in (2, 6, 7):
            net.set_conv_kernel(kind)
            d = np.abs(net.forward(x) - want)
            err[kind] = d.mean()
            print("kind %d %s: max |d| %.3e, mean |d| %.3e, head scale %.3e"
                  % (kind, shape, d.max(), d.mean(), scale))
            assert d.max() < 2e-2 * scale and d.mean() < 3e-3 * scale
    assert err[6] <= 1.25 * max(err[2], err[7])


@pytest.mark.parametrize("kind", [2, 6, 0])
def test_forward_base64_matches_reference(kind):
    """The 64-channel network (first-layer kernel with four N tiles, 64-wide
    row tiles at every level) against the reference."""
    p = _params(6, base=64)
    x, _ = _data(7, (1, 16, 12, 24))
    with UNet(base=64, seed=3) as net:
        net.set_conv_kernel(kind)
        net.set_params(p)
        got = net.forward(x)
    want = unet_ref.predict(p, x, base=64, bf16=True)
    d = np.abs(got - want)
    scale = np.abs(want - x).max()
    print("base 64: max |d| %.3e, mean |d| %.3e, head scale %.3e" % (d.max(), d.mean(), scale))
    assert d.max() < 2e-2 * scale and d.mean() < 3e-3 * scale


def test_gradient_base64_matches_reference():
    """Every parameter gradient of the 64-channel network (default kernels)."""
    p = _params(8, base=64)
    x, y = _data(9, (1, 8, 8, 16))
    with UNet(base=64, seed=3) as net:
        net.set_params(p)
        loss, g = net.loss_grad(x, y)
    lr, gr = unet_ref.loss_and_grad(p, x, y, base=64, bf16=True)
    assert abs(loss - lr) / lr < 1e-3
    o = 0
    for name, shp in layer_shapes(64):
        n = int(np.prod(shp))
        a, b = g[o:o + n].astype(np.float64), gr[o:o + n]
        o += n
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        cos = a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30)
        print(f"{name}: rel {rel:.2e} cos {cos:.6f}")
        assert rel < 5e-2 and cos > 0.998, name


@pytest.mark.parametrize("kind", KINDS)
def test_gradient_matches_reference(kind):
    p = _params(4)
    x, y = _data(5)
    with UNet(seed=3) as net:
        net.set_conv_kernel(kind)
        net.set_params(p)
        loss, g = net.loss_grad(x, y)
    lr, gr = unet_ref.loss_and_grad(p, x, y, bf16=True)
    assert abs(loss - lr) / lr < 1e-3
    o = 0
    for name, shp in layer_shapes(32):
        n = int(np.prod(shp))
        a, b = g[o:o + n].astype(np.float64), gr[o:o + n]
        o += n
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        cos = a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30)
        print(f"{name}: rel {rel:.2e} cos {cos:.6f}")
        assert rel < 5e-2 and cos > 0.998, name


def _pairs(shape, seeds, steps):
    xs, ys = [], []
    for sd in seeds:
        cen, rad = S.spheres(sd, 6, shape, 3.0, 6.0)
        with S.Solver(shape, precision=S.F32) as sv:
            sv.init_spheres(cen, rad)
            xs.append(np.stack(sv.download(), axis=-1))
            sv.step(steps)
            ys.append(np.stack(sv.download(), axis=-1))
    return np.array(xs, np.float32), np.array(ys, np.float32)


def test_training_reduces_loss():
    """Real phase-field pairs: sphere states and the same states 20 FP32
    steps later.  Adam with the library's own kernels must cut the loss.
    At lr 5e-4 the 60-step loss is 0.035-0.037 of the first one, whatever
    the accumulation order of the forward; at 1e-3 this small problem's
    loss spikes (to 20x) on the way and lands anywhere in 0.5-1.2x."""
    x, y = _pairs((32, 32, 32), (11, 12), 20)
    with UNet(seed=7) as net:
        l0 = net.train_step(x, y, 5e-4)
        for _ in range(60):
            l1 = net.train_step(x, y, 5e-4)
    print("loss", l0, "->", l1)
    assert l1 < 0.2 * l0


def test_apply_jumps_state_and_conserves_solute():
    shape = (32, 32, 32)
    cen, rad = S.spheres(3, 8, shape, 3.0, 6.0)
    for prec in (S.F64, S.F32):
        with S.Solver(shape, precision=prec) as sv, UNet(seed=8) as net:
            sv.init_spheres(cen, rad)
            before = sv.download()
            c0 = sv.reduce()["sum_c"]
            net.set_params(_params(9, 0.01))
            net.apply(sv)
            after = sv.download()
            c1 = sv.reduce()["sum_c"]
            assert np.abs(after[0] - before[0]).max() > 1e-4
            assert abs(c1 - c0) / c0 < (1e-13 if prec == S.F64 else 1e-9)
            ms, flop = net.last_forward()
            assert ms > 0 and flop > 0


@pytest.mark.parametrize("shape", [(16, 24, 32), (32, 32, 32)])
@pytest.mark.parametrize("kind", [2, 6, 0])
def test_apply_equals_the_host_forward(kind, shape):
    """apply reads and writes the context's fields (row-blocked kernels) or
    goes through packed FP32 copies (the others); either way etaS and etaL
    are bitwise the network's output on the same state (forward from host
    arrays), and c is that output shifted uniformly onto the old total.  At
    32^3 the head runs in the last layer's epilogue (kinds 2 and 6), on the
    context's fields or on the packed output."""
    cen, rad = S.spheres(4, 6, shape, 3.0, 6.0)
    for prec in (S.F64, S.F32):
        with S.Solver(shape, precision=prec) as sv, UNet(seed=8) as net:
            net.set_conv_kernel(kind)
            net.set_params(_params(10, 0.02))
            sv.init_spheres(cen, rad)
            sv.step(3)
            x = np.stack(sv.download(), axis=-1).astype(np.float32)[None]
            y = net.forward(x)[0]
            net.apply(sv)
            s, l, c = sv.download()
            cast = np.float64 if prec == S.F64 else np.float32
            np.testing.assert_array_equal(s, y[..., 0].astype(cast).astype(np.float64))
            np.testing.assert_array_equal(l, y[..., 1].astype(cast).astype(np.float64))
            shift = c - y[..., 2].astype(cast).astype(np.float64)
            # one uniform shift (FP32: stochastically rounded, within an ulp)
            tol = 0.0 if prec == S.F64 else 2.0 * np.spacing(np.float32(np.abs(c).max()))
            assert np.ptp(shift) <= tol + 1e-15


def test_adam_steps_match_the_closed_form():
    """Two train_step calls against Adam written out in float32 numpy
    (beta1 0.9, beta2 0.999, eps 1e-8, bias corrections 1 - beta^t):
    m = b1 m + (1 - b1) g, v = b2 v + (1 - b2) g^2,
    p -= lr (m / c1) / (sqrt(v / c2) + eps), with g from loss_grad at the
    parameters each step starts from.  The first step moves every parameter
    with |g| >> eps by lr in the direction of -g (m/c1 = g, v/c2 = g^2)."""
    x, y = _pairs((16, 16, 16), (21,), 10)
    f = np.float32
    b1, b2, eps, lr = f(0.9), f(0.999), f(1e-8), f(1e-3)
    with UNet(seed=3) as net:
        p = net.get_params()
        m = np.zeros_like(p)
        v = np.zeros_like(p)
        for t in (1, 2):
            _, g = net.loss_grad(x, y)
            net.train_step(x, y, float(lr))
            got = net.get_params()
            m = b1 * m + (f(1) - b1) * g
            v = b2 * v + (f(1) - b2) * g * g
            c1 = f(1) - f(np.power(b1, f(t), dtype=np.float32))
            c2 = f(1) - f(np.power(b2, f(t), dtype=np.float32))
            want = p - lr * (m / c1) / (np.sqrt(v / c2) + eps)
            err = np.abs(got - want)
            ulps = err / np.spacing(np.abs(want))
            print(f"Adam step {t}: max {float(ulps.max()):.2f} ulps from the closed form",
                  flush=True)
            assert (ulps <= 2).all(), (t, float(ulps.max()))
            if t == 1:
                # m/c1 = g and v/c2 = g^2: no parameter moves by more than lr
                assert (np.abs(got - p) <= lr * 1.001).all()
                # |g| >> eps: the step is lr in the direction of -g
                steady = np.abs(g) > f(1e-6)
                assert steady.sum() > 20
                np.testing.assert_allclose((p - got)[steady], lr * np.sign(g[steady]), rtol=1e-2)
            p = got
