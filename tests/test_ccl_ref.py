"""Known answers for the CPU labelling reference used by the GPU particle
tests."""
import numpy as np

from tests.ccl_ref import component_sizes, mean_radius


def test_two_blocks_and_a_diagonal_pair():
    m = np.zeros((8, 8, 8), bool)
    m[1:3, 1:3, 1:3] = True          # 8 sites
    m[5, 5, 5] = m[6, 6, 6] = True   # diagonal: not 6-connected -> two of 1
    assert component_sizes(m, periodic=False).tolist() == [8, 1, 1]


def test_periodic_wrap_joins_across_every_face():
    m = np.zeros((6, 7, 8), bool)
    m[0, 3, 3] = m[5, 3, 3] = True   # z faces
    m[2, 0, 1] = m[2, 6, 1] = True   # y faces
    m[4, 2, 0] = m[4, 2, 7] = True   # x faces
    assert component_sizes(m, periodic=True).tolist() == [2, 2, 2]
    assert component_sizes(m, periodic=False).tolist() == [1] * 6


def test_ring_closing_through_the_boundary():
    m = np.zeros((4, 4, 10), bool)
    m[1, 1, :] = True                # a line along x wraps into a ring
    m[1, 2, 0] = m[1, 2, 9] = True   # attached at both ends of the box
    assert component_sizes(m).tolist() == [12]
    m[1, 1, 5] = False               # cut: still one piece through the wrap
    assert component_sizes(m).tolist() == [11]


def test_mean_radius_of_unit_volume_spheres():
    v = 4.0 / 3.0 * np.pi * np.array([1.0, 8.0])
    assert abs(mean_radius(v) - 1.5) < 1e-12
