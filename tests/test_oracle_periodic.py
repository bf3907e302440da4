"""Pins of the oracle's periodic ghost construction (CPU only).

- The reference's ghost-cell fixture (base/tests/test_cell.py:296-420):
  16 real cells and 12 ghost cells with PeriodicDomain(xmin=-0.2, xmax=1).
- Neighbour sets over the ghost-extended arrays, mapped back to the
  originals, equal brute-force minimum-image sets.
- A lattice filling a fully periodic box has no boundary: every particle's
  summation density is the same, and a uniform pressure gives no force.
"""
import numpy as np
import pytest

from oracle import oracle as O

X = np.array([-0.125, 0.125, 0.375, 0.625, 0.875] * 5)
Y = np.repeat([-0.125, 0.125, 0.375, 0.625, 0.875], 5)


def test_reference_ghost_cell_fixture():
    a1 = O.make_array(x=X, y=Y, h=np.ones(25) * 0.1)
    a2 = O.make_array(x=X, y=Y, h=np.ones(25) * 0.15)
    cs = 2.0 * 0.15  # max_radius_scale * max h (test_cell.py:351)
    cells, ghosts = O.ghost_cells([a1, a2], (1.2, 0.0, 0.0), (-0.2, -1000.0, -1000.0), cs, cs)
    assert len(cells) == 16
    assert len(cells | set(ghosts)) == 28  # test_cell.py:413
    # the x = -1 and x = 0 columns are copied up, the x = 2 column down
    assert {g[0] for g in ghosts} == {3, 4, -2}


def min_image_sets(dst, src, period, rs):
    out = []
    for i in range(len(dst["x"])):
        d2 = np.zeros(len(src["x"]))
        for k, L in zip(("x", "y", "z"), period):
            d = src[k] - dst[k][i]
            if L > 0:
                d = d - L * np.round(d / L)
            d2 = d2 + d * d
        out.append(sorted(np.nonzero(d2 < (dst["h"][i] * rs) ** 2)[0].tolist()))
    return out


def periodic_sets(dst, src, period, pmin, rs=2.0):
    cs = rs * max(dst["h"].max(), src["h"].max())
    ext, origin = O.with_ghosts(src, period, pmin, cs, cs)
    off, nb = O.nnps_csr(dst, ext, cell_size=cs, radius_scale=rs)
    return [sorted(origin[nb[off[i]:off[i + 1]]].tolist()) for i in range(len(dst["x"]))]


@pytest.mark.parametrize("dim,axes", [(2, (1, 0, 0)), (2, (1, 1, 0)), (3, (1, 1, 1)),
                                      (3, (0, 1, 0))])
def test_ghost_sets_equal_minimum_image(dim, axes):
    rng = np.random.Generator(np.random.PCG64(dim * 10 + sum(axes)))
    n = 600 if dim == 2 else 1500
    box = np.array([1.0, 0.8, 0.6])
    pos = rng.random((3, n)) * box[:, None]
    pos[dim:] = 0
    a = O.make_array(x=pos[0], y=pos[1], z=pos[2], h=0.03 + 0.01 * rng.random(n))
    period = tuple(float(box[d]) if (axes[d] and d < dim) else 0.0 for d in range(3))
    pmin = (0.0, 0.0, 0.0)
    assert periodic_sets(a, a, period, pmin) == min_image_sets(a, a, period, 2.0)
    b = O.make_array(x=pos[0][:200] * 0.99, y=pos[1][:200], z=pos[2][:200],
                     h=np.full(200, 0.035))
    assert periodic_sets(b, a, period, pmin) == min_image_sets(b, a, period, 2.0)


def lattice(nx, ny, dx, h):
    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    x = (i.ravel() + 0.5) * dx
    y = (j.ravel() + 0.5) * dx
    n = len(x)
    return O.make_array(x=x, y=y, h=np.full(n, h), m=np.full(n, 1000.0 * dx * dx),
                        rho=np.full(n, 1000.0), p=np.full(n, 500.0), cs=np.full(n, 10.0))


def test_fully_periodic_lattice_is_uniform():
    dx = 0.05
    a = lattice(20, 16, dx, 1.3 * dx)
    period, pmin = (20 * dx, 16 * dx, 0.0), (0.0, 0.0, 0.0)
    cs = 2.0 * 1.3 * dx
    ext, origin = O.with_ghosts(a, period, pmin, cs, cs)
    rho = O.sphrho(2, a, [ext], cell_size=cs)
    assert np.ptp(rho) <= 1e-12 * rho.mean()
    # summation density of an infinite lattice is close to the rest density
    assert abs(rho.mean() / 1000.0 - 1.0) < 0.02
    acc = O.momentum(2, a, [ext], alpha=0.0, beta=0.0, cell_size=cs)
    assert np.abs(acc).max() < 1e-9
    # without the ghosts the edges are depleted
    rho0 = O.sphrho(2, a, [a], cell_size=cs)
    assert np.ptp(rho0) > 0.1 * rho0.mean()
