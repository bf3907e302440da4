"""Pins of the FP64 oracle (CPU only).

Each check is independent of the oracle's own code: closed forms worked by
hand, quadrature of the kernel (base/tests/test_kernels.py:107-296), the
reference's test expectations (base/tests/test_cell.py:28-63,
test_nnps.py:94-270, sph/tests/test_density_funcs.py:86-137), brute-force
all-pairs neighbour sets, and invariants of the scheme (momentum
conservation, exact constant-acceleration motion under RK2).
"""
import math

import numpy as np
import pytest
from scipy import integrate

from oracle import oracle as O

FAC = {1: 2.0 / 3.0, 2: 10.0 / (7.0 * math.pi), 3: 1.0 / math.pi}


# ------------------------------------------------------------------ kernel
@pytest.mark.parametrize("dim", [1, 2, 3])
def test_cubic_spline_values_by_hand(dim):
    h = 0.5
    fac = FAC[dim] / h ** dim
    # q = 0: W = fac; q = 1: 1 - 1.5*(1)*(0.5) = 0.25; q = 1.5: 0.25*0.5^3;
    # q = 2 and beyond: 0
    for r, shape in ((0.0, 1.0), (0.5, 0.25), (0.75, 0.25 * 0.125), (1.0, 0.0), (1.3, 0.0)):
        w = O.cubic_w((r, 0, 0), (0, 0, 0), h, dim)
        assert w == pytest.approx(shape * fac, rel=1e-15, abs=1e-300)


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_cubic_spline_normalisation(dim):
    """int W dV = 1 (base/tests/test_kernels.py:132-150, 1D/2D/3D)."""
    h = 0.3
    def w(r):
        return O.cubic_w((r, 0, 0), (0, 0, 0), h, dim)
    if dim == 1:
        val = 2 * integrate.quad(w, 0, 2 * h, points=[h])[0]
    elif dim == 2:
        val = integrate.quad(lambda r: 2 * math.pi * r * w(r), 0, 2 * h, points=[h])[0]
    else:
        val = integrate.quad(lambda r: 4 * math.pi * r * r * w(r), 0, 2 * h, points=[h])[0]
    assert val == pytest.approx(1.0, abs=1e-10)


def test_cubic_gradient_by_hand_and_moments():
    # 2D, h = 1, r_ab = (0.5, 0): q = 0.5, val = 3 (0.75 q - 1)/h^2 = -1.875
    g = O.cubic_grad((0.5, 0, 0), (0, 0, 0), 1.0, 2)
    assert g[0] == pytest.approx(0.5 * -1.875 * FAC[2], rel=1e-15)
    assert g[1] == 0.0 and g[2] == 0.0
    # q = 1.5 branch: val = -0.75 (2 - q)^2 / (h r)
    g = O.cubic_grad((0.0, 1.5, 0), (0, 0, 0), 1.0, 3)
    assert g[1] == pytest.approx(1.5 * (-0.75 * 0.25 / 1.5) * FAC[3], rel=1e-15)
    # q <= 1e-14: zero gradient (self pair)
    assert np.all(O.cubic_grad((0, 0, 0), (0, 0, 0), 1.0, 3) == 0)
    # 1D moment int x dW/dx dx = -1 (test_kernels.py:77-93)
    h = 0.2
    val = integrate.quad(lambda x: x * O.cubic_grad((x, 0, 0), (0, 0, 0), h, 1)[0],
                         -2 * h, 2 * h, points=[-h, 0, h])[0]
    assert val == pytest.approx(-1.0, abs=1e-10)


# ------------------------------------------------------------------ binning
def test_real_to_int_reference_cases():
    """base/tests/test_cell.py:28-63."""
    cases = [(0.5, 0), (0.1, 0), (1.0, 1), (1.5, 1), (1.9, 1), (2.1, 2), (2.6, 2),
             (-2.01, -3), (-1.01, -2), (-0.5, -1)]
    for v, c in cases:
        assert O.real_to_int(v, 1.0) == c


def _dataset_1():
    """base/tests/common_data.py:8-31."""
    p1 = O.make_array(x=[0.25, 0.8, 0.5, 0.8, 0.2, 0.5, 1.5, 1.5],
                      y=[0.25, 0.1, 0.5, 0.8, 0.9, 1.5, 0.5, 1.5], z=[0.0] * 8, h=[1.0] * 8)
    p2 = O.make_array(x=[0.2, 1.2, 1.5, 0.4], y=[0.0] * 4, z=[1.6, 1.5, -0.5, 0.4],
                      h=[1.0] * 4)
    return p1, p2


def test_point_queries_reference_expectations():
    """base/tests/test_nnps.py:94-177 (cell size 1)."""
    p1, p2 = _dataset_1()
    assert list(O.point_query(p1, 1.0, (0.4, 0.0, 0.4), 0.5)) == [0]
    assert sorted(O.point_query(p1, 1.0, (0.4, 0.0, 0.4), 1.0)) == [0, 1, 2, 3]
    assert sorted(O.point_query(p1, 1.0, (1.5, 0.0, -0.5), 4.0)) == list(range(8))
    c = (1.0, 1.0, 1.0)
    for r, ids in [(0.1, []), (1.0, []), (1.136, [1]), (1.358, [1, 3]),
                   (1.4142135623730951, [0, 1, 3]), (1.88, [0, 1, 2, 3])]:
        assert sorted(O.point_query(p2, 1.0, c, r)) == ids, r
    assert sorted(O.point_query(p2, 1.0, c, 1.88, exclude_index=3)) == [0, 1, 2]


def test_fixed_dest_reference_expectations():
    """base/tests/test_nnps.py:221-270: dest parr2, source parr1."""
    p1, p2 = _dataset_1()
    for rs, dst_i, ids in [(0.5, 3, [0]), (1.0, 3, [0, 1, 2, 3]), (4.0, 2, list(range(8)))]:
        off, nb = O.nnps_csr(p2, p1, cell_size=1.0, radius_scale=rs)
        assert sorted(nb[off[dst_i]:off[dst_i + 1]]) == ids


def _brute_sets(dst, src, rs):
    """All-pairs sets with the reference's FP64 expression."""
    out = []
    for i in range(len(dst["x"])):
        dx = src["x"] - dst["x"][i]
        dy = src["y"] - dst["y"][i]
        dz = src["z"] - dst["z"][i]
        d2 = (dx * dx + dy * dy) + dz * dz
        r = dst["h"][i] * rs
        out.append(set(np.nonzero(d2 < r * r)[0].tolist()))
    return out


@pytest.mark.parametrize("seed", [0, 1])
def test_cell_search_equals_brute_force(seed):
    """bench/nnps_c++/test.cpp:22-60: the cell-binned sets equal all-pairs."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = 400
    a = O.make_array(x=rng.random(n), y=rng.random(n), z=rng.random(n),
                     h=0.05 + 0.05 * rng.random(n))
    b = O.make_array(x=rng.random(300), y=rng.random(300), z=rng.random(300),
                     h=0.08 * np.ones(300))
    off, nb = O.nnps_csr(a)
    assert [set(nb[off[i]:off[i + 1]].tolist()) for i in range(n)] == _brute_sets(a, a, 2.0)
    # self is kept last (nnps.pyx:612-618)
    assert all(nb[off[i + 1] - 1] == i for i in range(n))
    off, nb = O.nnps_csr(a, b, cell_size=0.2)
    assert [set(nb[off[i]:off[i + 1]].tolist()) for i in range(n)] == _brute_sets(a, b, 2.0)


def test_strict_cutoff_on_lattice():
    g = np.arange(6) * 0.5
    x, y = np.meshgrid(g, g, indexing="ij")
    a = O.make_array(x=x.ravel(), y=y.ravel(), h=np.full(36, 0.5))
    off, nb = O.nnps_csr(a)
    k = 2 * 6 + 2
    assert off[k + 1] - off[k] == 9  # distance exactly 2h = 1 excluded


# ------------------------------------------------------------------ functions
def test_sphrho_four_particles_closed_form():
    """sph/tests/test_density_funcs.py:86-137: unit square, h = 1, m = 1."""
    a = O.make_array(x=[0, 0, 1, 1.0], y=[0, 1, 1, 0.0])
    rho = O.sphrho(2, a, [a])
    fac = FAC[2]
    closed = fac * (1.0 + 2 * 0.25 + 0.25 * (2 - math.sqrt(2)) ** 3)
    assert np.allclose(rho, closed, rtol=0, atol=1e-15)


def test_tait_closed_form():
    a = O.make_array(x=np.zeros(3), rho=[990.0, 1000.0, 1010.0])
    p, cs = O.tait(a, 10.0, 1000.0, 7.0)
    B = 100.0 * 1000.0 / 7.0
    assert p[1] == 0.0 and cs[1] == 10.0
    assert p[2] == pytest.approx(B * (1.01 ** 7 - 1), rel=1e-14)
    assert cs[0] == pytest.approx(10.0 * 0.99 ** 3, rel=1e-14)


def test_momentum_two_particles_by_hand():
    """Two particles in 2D, h = 1, distance 0.5 along x, m = 1, rho = 1,
    p = (1, 3), approaching (u = (-1, 1)): Pi = (-alpha c mu + beta mu^2)/rho
    with mu = h v.r / (r^2 + eta^2 h^2)."""
    a = O.make_array(x=[0.5, 0.0], y=[0.0, 0.0], u=[-1.0, 1.0], p=[1.0, 3.0],
                     cs=[2.0, 2.0])
    alpha, beta, eta = 0.5, 0.25, 0.1
    acc = O.momentum(2, a, [a], alpha, beta, eta)
    mu = 1.0 * (-2.0 * 0.5) / (0.25 + 0.01)
    pi = (-alpha * 2.0 * mu + beta * mu * mu) / 1.0
    gx = 0.5 * -1.875 * FAC[2]  # grad_a W for r_ab = (0.5, 0) (see above)
    expect0 = -1.0 * (1.0 + 3.0 + pi) * gx
    assert acc[0][0] == pytest.approx(expect0, rel=1e-14)
    assert acc[0][1] == pytest.approx(-expect0, rel=1e-14)  # antisymmetric pair
    assert acc[1][0] == 0.0 and acc[1][1] == 0.0


def test_momentum_conservation_random_set():
    """sum_a m_a a_a = 0 for any closed set (pair forces antisymmetric)."""
    rng = np.random.Generator(np.random.PCG64(5))
    n = 300
    a = O.make_array(x=rng.random(n), y=rng.random(n), z=rng.random(n),
                     u=rng.normal(size=n), v=rng.normal(size=n), w=rng.normal(size=n),
                     h=np.full(n, 0.06), m=0.5 + rng.random(n),
                     rho=1.0 + 0.1 * rng.random(n), p=rng.random(n), cs=1.0 + rng.random(n))
    acc = O.momentum(3, a, [a], 0.7, 0.3, 0.1)
    tot = (acc * a["m"]).sum(axis=1)
    scale = (np.abs(acc) * a["m"]).sum()
    assert np.all(np.abs(tot) < 1e-13 * scale)


def test_xsph_invariants():
    """Uniform velocity gives no XSPH correction; two particles exchange
    equal and opposite mass-weighted corrections."""
    rng = np.random.Generator(np.random.PCG64(6))
    n = 200
    a = O.make_array(x=rng.random(n), y=rng.random(n), u=np.full(n, 2.0), v=np.full(n, -1.0),
                     h=np.full(n, 0.08))
    assert np.all(np.abs(O.xsph(2, a, [a], 0.5)) < 1e-15)
    b = O.make_array(x=[0.0, 0.5], y=[0.0, 0.0], u=[1.0, -1.0], m=[1.0, 2.0])
    dv = O.xsph(2, b, [b], 0.5)
    # m_a dv_a = eps m_a m_b W / rho_ab (v_b - v_a): antisymmetric
    assert b["m"][0] * dv[0][0] == pytest.approx(-b["m"][1] * dv[0][1], rel=1e-14)
    w = (1 - 1.5 * 0.25 * 0.75) * FAC[2]  # W at q = 0.5, h = 1
    assert dv[0][0] == pytest.approx(0.5 * 2.0 * w / 1.0 * (-2.0), rel=1e-14)


# ------------------------------------------------------------------ integrators
@pytest.mark.parametrize("integrator", ["euler", "rk2"])
def test_constant_acceleration_closed_form(integrator):
    """One isolated particle feels only gravity (grad W(0) = 0): RK2 is exact
    for constant acceleration, Euler lags by g dt^2 n/2."""
    a = O.make_array(x=[0.1], y=[0.2], u=[1.0], v=[-0.5], h=[0.01], m=[1.0], rho=[1.0])
    g = (0.0, -9.81, 0.0)
    dt, n = 1e-3, 50
    sch = O.scheme(2, integrator=integrator, dt=dt, co=1.0, ro=1.0, gamma=7.0, alpha=0.0,
                   beta=0.0, gravity=g, cell_size=1.0)
    O.scheme_step(sch, [a], n)
    t = n * dt
    y_exact = 0.2 - 0.5 * t + 0.5 * g[1] * t * t
    if integrator == "rk2":
        assert a["y"][0] == pytest.approx(y_exact, rel=1e-13)
    else:
        assert a["y"][0] == pytest.approx(y_exact - 0.5 * g[1] * dt * t, rel=1e-13)
    assert a["x"][0] == pytest.approx(0.1 + t, rel=1e-13)
    assert a["v"][0] == pytest.approx(-0.5 + g[1] * t, rel=1e-13)


def test_scheme_conserves_momentum_and_mass():
    """Elliptical-drop lattice (coarse), RK2, no gravity: total momentum
    stays at its initial value (zero by symmetry) to rounding."""
    dx = 0.1
    R = 10
    i, j = np.meshgrid(np.arange(-R, R + 1), np.arange(-R, R + 1), indexing="ij")
    k = (i * i + j * j) <= R * R
    x, y = i[k] * dx, j[k] * dx
    n = len(x)
    a = O.make_array(x=x, y=y, u=-100 * x, v=100 * y, m=np.full(n, dx * dx),
                     h=np.full(n, 1.3 * dx))
    c0 = O.conservation([a], co=100.0, ro=1.0)
    sch = O.scheme(2, integrator="rk2", dt=1e-5, co=100.0, ro=1.0, alpha=0.01, beta=0.0)
    O.scheme_step(sch, [a], 10)
    c1 = O.conservation([a], co=100.0, ro=1.0)
    assert c1[0] == c0[0]
    scale = np.sum(a["m"] * np.abs(a["u"]))
    assert abs(c1[1]) < 1e-12 * scale and abs(c1[2]) < 1e-12 * scale


def test_tait_elastic_energy_derivative():
    """d e / d rho = p / rho^2 for the elastic energy of the conservation
    sums (checked by central differences)."""
    co, ro, gamma = 10.0, 1000.0, 7.0
    B = co * co * ro / gamma
    def e_of(rho):
        a = O.make_array(x=[0.0], rho=[rho], m=[1.0])
        return O.conservation([a], co=co, ro=ro, gamma=gamma)[6]
    for rho in (995.0, 1000.0, 1012.0):
        h = 1e-3
        de = (e_of(rho + h) - e_of(rho - h)) / (2 * h)
        p = B * ((rho / ro) ** gamma - 1)
        assert de == pytest.approx(p / rho ** 2, rel=1e-6, abs=1e-12)
    assert e_of(ro) == pytest.approx(0.0, abs=1e-12)
