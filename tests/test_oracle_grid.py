"""Boundary stencil of the oracle (neighbourer_flat.py) against geometric
closed forms built from (row, col) with numpy rolls, and hand cases for the
custom codes."""

import numpy as np
import pytest

SHAPES = [(2, 2), (2, 5), (5, 2), (3, 3), (4, 5), (5, 4), (6, 7)]
DIRS = [(-1, 0), (0, -1), (0, 1), (1, 0)]  # top, left, right, bottom


def geometric(ny, nx, mode):
    """Neighbour table: the 4-adjacent cell, wrapped east-west (mode 1) or
    north-south (mode 2), -1 where it leaves the grid."""
    r, c = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    out = np.full((4, ny * nx), -1, np.int64)
    for k, (dr, dc) in enumerate(DIRS):
        rr, cc = r + dr, c + dc
        if mode == 1:
            cc = cc % nx
        if mode == 2:
            rr = rr % ny
        ok = (rr >= 0) & (rr < ny) & (cc >= 0) & (cc < nx)
        out[k] = np.where(ok, rr * nx + cc, -1).ravel()
    return out


def leaves(ny, nx, mode):
    r, c = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    if mode == 0:
        m = (r == 0) | (r == ny - 1) | (c == 0) | (c == nx - 1)
    elif mode == 1:
        m = (r == 0) | (r == ny - 1)
    else:
        m = (c == 0) | (c == nx - 1)
    return m.ravel()


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("ny,nx", SHAPES)
def test_neighbours_and_outlets(ora, ny, nx, mode):
    g = ora.Grid(nx, ny, 1.0, mode)
    want = geometric(ny, nx, mode)
    got = np.array([[ora.neighbour(g, i, k) for i in range(g.N)] for k in range(4)])
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(np.array([ora.can_leave(g, i) for i in range(g.N)]), leaves(ny, nx, mode))


def test_which_edge_layout(ora):
    """The 9-class layout drawn in neighbourer_flat.py:79-83."""
    g = ora.Grid(5, 5, 1.0, 0)
    want = np.array([[1, 2, 2, 2, 3], [4, 0, 0, 0, 5], [4, 0, 0, 0, 5], [4, 0, 0, 0, 5], [6, 7, 7, 7, 8]])
    got = np.array([ora.which_edge(g, i) for i in range(25)]).reshape(5, 5)
    np.testing.assert_array_equal(got, want)


def test_custom_codes(ora):
    """Codes: 0 no data (no links, no links into it), 1 normal, 3 outlet,
    7 inlet (normal for routing), 9 periodic on its border."""
    ny, nx = 4, 5
    b = np.ones((ny, nx), np.uint8)
    b[1, 2] = 0        # hole
    b[2, 0] = 9        # wraps west -> east
    b[2, 4] = 9        # wraps east -> west
    b[0, 1] = 3        # outlet
    b[3, 3] = 7        # inlet
    g = ora.Grid(nx, ny, 1.0, 3, b)
    idx = lambda r, c: r * nx + c  # noqa: E731
    # the hole has no links, and nothing links into it
    assert [ora.neighbour(g, idx(1, 2), k) for k in range(4)] == [-1, -1, -1, -1]
    assert ora.neighbour(g, idx(1, 1), 2) == -1
    assert ora.neighbour(g, idx(0, 2), 3) == -1
    # periodic code on the west border wraps to the east column, and back
    assert ora.neighbour(g, idx(2, 0), 1) == idx(2, 4)
    assert ora.neighbour(g, idx(2, 4), 2) == idx(2, 0)
    # a normal west-border cell does not wrap
    assert ora.neighbour(g, idx(1, 0), 1) == -1
    # interior links as in the normal mode; inlets route normally
    assert ora.neighbour(g, idx(2, 2), 0) == -1  # top is the hole
    assert ora.neighbour(g, idx(3, 3), 0) == idx(2, 3)
    # only code 3 leaves the domain
    leave = [i for i in range(g.N) if ora.can_leave(g, i)]
    assert leave == [idx(0, 1)]
