"""CPU tests of the one-process ImplicitGlobalGrid functions the reference's multi-XPU examples
call (init_global_grid, nx_g, x_g, finalize_global_grid): closed forms, no GPU."""
import numpy as np
import pytest

import parallelstencil as ps
from parallelstencil import igg


@pytest.fixture(autouse=True)
def _fresh_grid():
    igg._grid = None
    yield
    igg._grid = None


def test_global_sizes_and_coordinates():
    me, dims, nprocs, coords, comm = igg.init_global_grid(64, 32, 16, select_device=False)
    assert (me, dims, nprocs, coords) == (0, (1, 1, 1), 1, (0, 0, 0))
    assert (igg.nx_g(), igg.ny_g(), igg.nz_g()) == (64, 32, 16)
    dx = 1.0 / (igg.nx_g() - 1)
    A = ps.ArgSpec((64, 32, 16))
    assert [igg.x_g(i, dx, A) for i in (1, 2, 64)] == [0.0, dx, 63 * dx]
    Vx = ps.ArgSpec((65, 32, 16))  # staggered: half a cell to the left
    assert igg.x_g(1, dx, Vx) == -0.5 * dx and igg.x_g(65, dx, Vx) == 63.5 * dx
    igg.finalize_global_grid()
    with pytest.raises(ps.ArgumentError, match="before init_global_grid"):
        igg.nx_g()


def test_periodic_coordinates_are_periodic_images():
    igg.init_global_grid(10, 8, 1, periodx=1, periody=1, select_device=False)
    assert (igg.nx_g(), igg.ny_g(), igg.nz_g()) == (8, 6, 1)
    dx = 0.5
    A = ps.ArgSpec((10, 8))
    xs = [igg.x_g(i, dx, A) for i in range(1, 11)]
    # cells 2..9 are the 8 global cells 0..7; cell 1 is the image of 7, cell 10 that of 0
    assert xs == [7 * dx] + [j * dx for j in range(8)] + [0.0]
    assert all(0 <= x <= (igg.nx_g() - 1) * dx for x in xs)


def test_grid_errors():
    with pytest.raises(ps.ArgumentError, match="one process"):
        igg.init_global_grid(16, 16, 16, dimx=2, select_device=False)
    igg.init_global_grid(16, 16, 16, select_device=False)
    with pytest.raises(ps.ArgumentError, match="already been initialized"):
        igg.init_global_grid(16, 16, 16, select_device=False)
    igg.finalize_global_grid()
    with pytest.raises(ps.ArgumentError, match="halowidth"):
        igg.init_global_grid(16, 16, 16, overlaps=(2, 2, 2), halowidths=(2, 1, 1), select_device=False)


def test_halo_boxes():
    # overlap 2, halo width 1, size 10: cell 1 <- cell 9, cell 10 <- cell 2
    assert igg._halo_boxes((10, 8, 6), 0, 2, 1) == [(((1, 1, 1), (1, 8, 6)), ((9, 1, 1), (9, 8, 6))),
                                                     (((10, 1, 1), (10, 8, 6)), ((2, 1, 1), (2, 8, 6)))]
    assert igg._halo_boxes((10, 8, 6), 2, 4, 2) == [(((1, 1, 1), (10, 8, 2)), ((1, 1, 3), (10, 8, 4))),
                                                     (((1, 1, 5), (10, 8, 6)), ((1, 1, 3), (10, 8, 4)))]
