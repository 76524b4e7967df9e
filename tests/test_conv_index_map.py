"""The LDS-tiled convolution stages the halo of a 16 x 8 site brick; on the
coarse levels of small inputs the brick is larger than the grid, so the
staged coordinates can lie several periods outside it.  This restates the
kernel's coordinate map (wrapm in csrc/unet.hip) and checks, on the CPU,
that every staged site of every level of the grids the tests and the
surrogate use is in range, and that in-grid sites keep their periodic
neighbours."""
import pytest

BX, BY = 16, 8


def wrapm(i, n):
    r = i % n
    return r + n if r < 0 else r


@pytest.mark.parametrize("dims", [(16, 16, 16), (24, 12, 40), (32, 32, 32), (48, 48, 48),
                                  (64, 64, 64), (4, 4, 4), (12, 12, 12), (256, 256, 256)])
def test_halo_coordinates_in_range(dims):
    D, H, W = dims
    for lev in range(3):
        h, w = H >> lev, W >> lev
        if h < 1 or w < 1:
            continue
        for by in range((h + BY - 1) // BY):
            for bx in range((w + BX - 1) // BX):
                for hy in range(BY + 2):
                    for hx in range(BX + 2):
                        gy, gx = by * BY + hy - 1, bx * BX + hx - 1
                        y, x = wrapm(gy, h), wrapm(gx, w)
                        assert 0 <= y < h and 0 <= x < w
                        if -1 <= gy <= h and -1 <= gx <= w:
                            # a neighbour of an in-grid site: the periodic image
                            assert y == (gy + h) % h and x == (gx + w) % w
