"""The controller's cheap judge.  For a tanh interface phi = (1 - tanh((r -
R)/2))/2 = 1/(1 + exp(r - R)), |grad phi| = phi (1 - phi) exactly, so
sum phi (1 - phi) measures interface area, and 3 sum phi / sum phi (1 - phi)
is a volume-to-area radius.  The Sommerfeld expansion of these Fermi-shaped
integrals gives, up to exponentially small terms,
    int phi dV           = (4 pi / 3) (R^3 + pi^2 R)
    int phi (1 - phi) dV = 4 pi (R^2 + pi^2 / 3)
so one sphere reads R (R^2 + pi^2) / (R^2 + pi^2 / 3) and a population
sum (R^3 + pi^2 R) / sum (R^2 + pi^2 / 3), a Sauter-type mean radius.
Checked with numpy on sampled spheres (no GPU); the library's reduction of
sum h (1 - h) is checked against numpy in test_gpu_parity.py."""
import numpy as np
import pytest

PI2 = np.pi ** 2


def _field(n, spheres):
    ax = np.arange(n) + 0.5
    z, y, x = np.meshgrid(ax, ax, ax, indexing="ij")
    phi = np.zeros((n, n, n))
    for (cz, cy, cx, r) in spheres:
        d = np.sqrt((z - cz) ** 2 + (y - cy) ** 2 + (x - cx) ** 2)
        phi = np.maximum(phi, 0.5 * (1.0 - np.tanh((d - r) / 2.0)))
    return phi


def _proxy(phi):
    return 3.0 * phi.sum() / (phi * (1.0 - phi)).sum()


@pytest.mark.parametrize("r", [4.0, 8.0, 12.0, 20.0])
def test_single_sphere_closed_form(r):
    phi = _field(64, [(32, 32, 32, r)])
    want = r * (r * r + PI2) / (r * r + PI2 / 3)
    assert abs(_proxy(phi) - want) / want < 5e-3
    # and the leading term is the radius itself
    assert abs(_proxy(phi) - r) / r < 0.35


def test_population_closed_form():
    sph = [(20, 20, 20, 6.0), (60, 60, 20, 10.0), (20, 60, 60, 14.0), (60, 20, 60, 8.0)]
    phi = _field(80, sph)
    r = np.array([s[3] for s in sph])
    want = (r ** 3 + PI2 * r).sum() / (r ** 2 + PI2 / 3).sum()
    assert abs(_proxy(phi) - want) / want < 5e-3
    # a coarsened population (the same volume in one particle) reads larger
    big = _field(80, [(40, 40, 40, float(np.cbrt((r ** 3).sum())))])
    assert _proxy(big) > 1.3 * _proxy(phi)
