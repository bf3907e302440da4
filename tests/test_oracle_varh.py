"""Pins of the oracle's variable-h locator and SPHDensityRate (CPU only).

- The symmetric sets reproduce the reference's own expectation table
  (base/tests/test_variable_h.py:71-81) and equal a brute-force all-pairs
  test d2 < (h_a rs)^2 or d2 < (h_b rs)^2; with one h they equal the
  FixedDest sets.
- SPHDensityRate (density_funcs.pyx:89-158) is the exact time derivative of
  the summation density for fixed h: a central difference of SPHRho along
  the velocity field pins its sign, its velocity difference and its
  gradient index.  Over the symmetric sets this holds with variable h as
  well; over FixedDest sets it does not (pairs with h_b > h_a are lost).
  A uniform velocity gives exactly zero, and a two-particle case is worked
  by hand.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

FAC = {1: 2.0 / 3.0, 2: 10.0 / (7.0 * math.pi), 3: 1.0 / math.pi}

# base/tests/test_variable_h.py:71-81
GOLDEN_1D = {0: [0, 1, 2], 1: [0, 1, 2, 3, 5], 2: [0, 1, 2, 3, 4, 5], 3: [1, 2, 3, 4, 5],
             4: [2, 3, 4, 5, 6], 5: [1, 2, 3, 4, 5, 6, 7, 8, 9], 6: [4, 5, 6, 7, 8],
             7: [5, 6, 7, 8, 9], 8: [5, 6, 7, 8, 9, 10], 9: [5, 7, 8, 9, 10], 10: [8, 9, 10]}


def golden_1d_array():
    """11 particles on [0, 1], h = dx + 1e-10, particle 5 with 2h
    (test_variable_h.py:14-39)."""
    x = np.linspace(0, 1, 11)
    dx = x[1] - x[0]
    h = np.ones_like(x) * (dx + 1e-10)
    h[5] = 2 * (dx + 1e-10)
    return O.make_array(x=x, h=h, m=np.ones_like(x) * dx, rho=np.ones_like(x))


def test_varh_reference_table():
    off, nb = O.nnps_csr(golden_1d_array(), radius_scale=2.0, varh=True)
    for i in range(11):
        assert sorted(nb[off[i]:off[i + 1]].tolist()) == GOLDEN_1D[i]


def test_fixed_dest_is_not_symmetric_on_the_reference_case():
    # particle 1 sees 5 only through 5's larger radius
    off, nb = O.nnps_csr(golden_1d_array(), radius_scale=2.0)
    assert 5 not in nb[off[1]:off[2]].tolist()


def random_varh(n, dim, seed, spread=2.0):
    rng = np.random.Generator(np.random.PCG64(seed))
    pos = [rng.random(n) for _ in range(dim)] + [np.zeros(n)] * (3 - dim)
    h = 0.04 * (1 + (spread - 1) * rng.random(n))
    return O.make_array(x=pos[0], y=pos[1], z=pos[2], h=h, m=rng.random(n) + 0.5,
                        u=rng.normal(size=n), v=rng.normal(size=n),
                        w=rng.normal(size=n) if dim == 3 else np.zeros(n),
                        rho=1000 + rng.random(n))


def brute_sets(dst, src, rs, sym):
    out = []
    for i in range(len(dst["x"])):
        d2 = ((src["x"] - dst["x"][i]) ** 2 + (src["y"] - dst["y"][i]) ** 2) + \
            (src["z"] - dst["z"][i]) ** 2
        r2 = (dst["h"][i] * rs) ** 2
        ok = d2 < r2
        if sym:
            ok |= d2 < (src["h"] * rs) ** 2
        out.append(sorted(np.nonzero(ok)[0].tolist()))
    return out


@pytest.mark.parametrize("dim,seed", [(1, 1), (2, 2), (3, 3)])
def test_varh_equals_brute_force(dim, seed):
    a = random_varh(300 if dim > 1 else 60, dim, seed)
    b = random_varh(200 if dim > 1 else 40, dim, seed + 10)
    for dst, src in ((a, a), (a, b), (b, a)):
        off, nb = O.nnps_csr(dst, src, radius_scale=2.0, varh=True)
        ref = brute_sets(dst, src, 2.0, True)
        got = [sorted(nb[off[i]:off[i + 1]].tolist()) for i in range(len(dst["x"]))]
        assert got == ref
        # no duplicates inside a row
        assert all(len(set(nb[off[i]:off[i + 1]])) == off[i + 1] - off[i]
                   for i in range(len(dst["x"])))


def test_varh_with_one_h_equals_fixed_dest():
    a = random_varh(400, 2, 5, spread=1.0)
    a["h"][:] = 0.05
    o1, n1 = O.nnps_csr(a, radius_scale=2.0, varh=True)
    o2, n2 = O.nnps_csr(a, radius_scale=2.0)
    assert np.array_equal(o1, o2)
    for i in range(len(a["x"])):
        assert sorted(n1[o1[i]:o1[i + 1]]) == sorted(n2[o2[i]:o2[i + 1]])


def test_density_rate_uniform_velocity_is_zero():
    a = random_varh(300, 2, 7)
    a["u"][:] = 1.25
    a["v"][:] = -0.5
    assert np.all(O.density_rate(2, a, [a], varh=True) == 0.0)


def test_density_rate_two_particles_by_hand():
    # 2D, h = 1, x_a - x_b = (0.5, 0) (q = 0.5), v_a - v_b = (2, 1), m_b = 3:
    # grad_a W = r_ab * fac * 3 (0.75 q - 1) / h^2 = (0.5, 0) * fac * -1.875
    a = O.make_array(x=[0.5], y=[0.0], h=[1.0], m=[1.0], u=[2.5], v=[1.0])
    b = O.make_array(x=[0.0], y=[0.0], h=[1.0], m=[3.0], u=[0.5], v=[0.0])
    got = O.density_rate(2, a, [b])
    assert got[0] == pytest.approx(3.0 * (2.0 * 0.5 * FAC[2] * -1.875), rel=1e-15)


def sph_density(a, dim, varh):
    return O.sphrho(dim, a, [a], varh=varh)


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_density_rate_is_the_time_derivative_of_summation_density(dim):
    a = random_varh(300 if dim > 1 else 80, dim, 20 + dim)
    vel = [a["u"], a["v"], a["w"]][:dim]
    eps = 1e-7

    def moved(sign):
        b = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in a.items()}
        for c, key in enumerate("xyz"[:dim]):
            b[key] = a[key] + sign * eps * vel[c]
        return b
    fd = (sph_density(moved(+1), dim, True) - sph_density(moved(-1), dim, True)) / (2 * eps)
    rate = O.density_rate(dim, a, [a], varh=True)
    assert np.abs(rate - fd).max() <= 1e-5 * np.abs(rate).max()
    # with fixed-dest sets, pairs seen only through the source's radius are
    # missing from both sums differently: the identity breaks
    rate_fd = O.density_rate(dim, a, [a], varh=False)
    assert np.abs(rate_fd - fd).max() > 1e-3 * np.abs(rate).max()
