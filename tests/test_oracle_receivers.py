"""compute_receivers of the oracle (receivers.py:15-58) against a vectorised
numpy closed form and hand cases."""

import numpy as np
import pytest

from oracle_util import receivers_closed_form
from pyfastflow_amd import synth


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("ny,nx,seed", [(2, 2, 1), (5, 7, 2), (33, 17, 3), (64, 64, 4), (100, 80, 5)])
def test_closed_form(ora, ny, nx, seed, mode):
    z = synth.dem(ny, nx, 2.0, 300.0, seed, n_pits=2, dx=10.0)
    rcv, grad = ora.compute_receivers(ora.Grid(nx, ny, 10.0, mode), z)
    r2, g2 = receivers_closed_form(z, ny, nx, mode, 10.0)
    np.testing.assert_array_equal(rcv, r2)
    np.testing.assert_array_equal(grad.view(np.uint32), g2.view(np.uint32))


def test_tilted_plane_drains_south(ora):
    ny, nx = 6, 5
    z = (np.arange(ny)[::-1, None] * 1.0 + np.zeros((1, nx))).astype(np.float32)
    rcv, grad = ora.compute_receivers(ora.Grid(nx, ny, 1.0, 0), z)
    i = np.arange(ny * nx)
    south = i + nx
    interior_rows = i // nx < ny - 1
    np.testing.assert_array_equal(rcv[interior_rows], south[interior_rows])
    np.testing.assert_array_equal(rcv[~interior_rows], i[~interior_rows])
    assert np.all(grad[interior_rows] == 1.0)


def test_flat_cells_are_their_own_receiver(ora):
    z = np.full((4, 4), 3.0, np.float32)
    rcv, grad = ora.compute_receivers(ora.Grid(4, 4, 1.0, 0), z)
    np.testing.assert_array_equal(rcv, np.arange(16))
    assert np.all(grad == 0)


def test_equal_drops_take_the_first_direction(ora):
    """top (0) before left (1) before right (2) before bottom (3)."""
    z = np.full((3, 3), 5.0, np.float32)
    z[1, 1] = 6.0  # centre; all four neighbours 1 lower
    rcv, _ = ora.compute_receivers(ora.Grid(3, 3, 1.0, 0), z)
    assert rcv[4] == 1  # top
    z[0, 1] = 5.5  # top now shallower: left wins
    rcv, _ = ora.compute_receivers(ora.Grid(3, 3, 1.0, 0), z)
    assert rcv[4] == 3


def test_periodic_wrap_column(ora):
    z = np.full((3, 4), 10.0, np.float32)
    z[1, 3] = 1.0  # east column, lowest
    z[1, 0] = 5.0  # west column: its left neighbour wraps to the east column
    rcv, _ = ora.compute_receivers(ora.Grid(4, 3, 1.0, 1), z)
    assert rcv[1 * 4 + 0] == 1 * 4 + 3
    rcv0, _ = ora.compute_receivers(ora.Grid(4, 3, 1.0, 0), z)
    assert rcv0[1 * 4 + 0] == 1 * 4 + 0  # no wrap in normal mode: self


def test_division_ambiguity_is_counted(ora):
    """The oracle reports cells where (zi-zj)/dx and (zi-zj)*fl(1/dx) pick
    different receivers; for the 512^2 plain workload (dx=1) there are none."""
    c = synth.plain_small(n=128)
    assert ora.receivers_div_ambiguity(ora.Grid(128, 128, 1.0, 0), c["z"]) == 0


def test_stochastic_draws_are_uniform(ora):
    """The seeded draws of RAND_RCV: mean 1/2, variance 1/12, and the
    fraction below 1/4 is 1/4, over 40 000 draws (5 sigma bounds)."""
    u = np.array([ora.uniform01(12345, i, k) for i in range(10_000) for k in range(4)])
    assert (u >= 0).all() and (u < 1).all()
    n = u.size
    assert abs(u.mean() - 0.5) < 5 * np.sqrt(1 / 12 / n)
    assert abs(u.var() - 1 / 12) < 5 * np.sqrt(1 / 180 / n)
    assert abs((u < 0.25).mean() - 0.25) < 5 * np.sqrt(0.25 * 0.75 / n)


def test_stochastic_draws_follow_splitmix64(ora):
    """Draw k of cell i with seed s is the top 24 bits of splitmix64 at the
    state s + (4i + k) * 0x9E3779B97F4A7C15.  Seed 0, cell 0, k = 1, 2, 3 are
    therefore splitmix64's first three outputs from state 0, published as
    0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F (Vigna's
    splitmix64.c, Java's SplittableRandom)."""
    for k, out in ((1, 0xE220A8397B1DCDAF), (2, 0x6E789E6AA1B965F4), (3, 0x06C45D188009454F)):
        assert ora.uniform01(0, 0, k) == (out >> 40) * 2.0**-24
    # k = 0 at seed 0 is the state 0 itself: splitmix64's finaliser maps 0 to 0
    assert ora.uniform01(0, 0, 0) == 0.0


def test_stochastic_receivers_two_slope_probability(ora):
    """On the plane z = -(2 row + col) every interior cell has two lower
    neighbours, bottom (slope 2/dx) and right (1/dx).  With the slopes
    weighted by independent uniform draws, bottom wins with probability
    P(2 u1 > u2) = 1 - 1/4 = 3/4.  The weighted slope is stored as gradient:
    max(2 u1, u2)/dx, with mean 13/12 / dx and variance 35/144 / dx^2."""
    ny, nx, dx = 300, 400, 2.0
    r, c = np.mgrid[0:ny, 0:nx]
    z = -(2.0 * r + 1.0 * c).astype(np.float32)
    g = ora.Grid(nx, ny, dx, 0)
    rcv, grad = ora.compute_receivers_stochastic(g, z, seed=7)
    inner = (r < ny - 1) & (c < nx - 1)
    i = np.arange(nx * ny).reshape(ny, nx)
    down = rcv.reshape(ny, nx)[inner] == (i + nx)[inner]
    right = rcv.reshape(ny, nx)[inner] == (i + 1)[inner]
    assert (down | right).all()
    p = down.mean()
    n = down.size
    assert abs(p - 0.75) < 5 * np.sqrt(0.75 * 0.25 / n)
    gi = grad.reshape(ny, nx)[inner]
    assert (gi > 0).all() and (gi < 2.0 / dx).all()
    assert abs(gi.mean() * dx - 13 / 12) < 5 * np.sqrt(35 / 144 / n)


def test_stochastic_receivers_reproducible_and_downhill(ora):
    """Same seed, same receivers; another seed, other receivers; every
    receiver is the cell itself or a strictly lower valid neighbour, and a
    cell with no lower neighbour keeps itself."""
    from pyfastflow_amd import synth

    z = synth.dem(64, 80, 2.0, 100.0, 3, n_pits=2, dx=10.0)
    g = ora.Grid(80, 64, 10.0, 1)
    a, ga = ora.compute_receivers_stochastic(g, z, seed=1)
    b, gb = ora.compute_receivers_stochastic(g, z, seed=1)
    c, _ = ora.compute_receivers_stochastic(g, z, seed=2)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(ga, gb)
    assert (a != c).any()
    det, _ = ora.compute_receivers(g, z)
    zf = z.ravel()
    moved = a != np.arange(a.size)
    assert (zf[a[moved]] < zf[moved]).all()
    np.testing.assert_array_equal(a[det == np.arange(a.size)], np.flatnonzero(det == np.arange(a.size)))
