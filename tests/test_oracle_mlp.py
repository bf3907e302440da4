"""The FP64 network oracle (oracle/mlp_oracle.py) pinned on the CPU:
analytic gradients against central finite differences, and the physics
edge features against the SPH oracle's pressure and viscosity sums."""
import numpy as np

from oracle import mlp_oracle as M
from oracle import oracle as O

CONST = dict(rho0=1000.0, co=30.0, gamma=7.0, u_scale=3.0, g=9.81, rho_scale=100.0,
             alpha=0.1, eta=0.1, dim=2)


def _graph(n_side=5, dx=0.1, hdx=1.3, seed=0, jitter=0.2, vel=1.0):
    rng = np.random.Generator(np.random.PCG64(seed))
    i, j = np.meshgrid(np.arange(n_side), np.arange(n_side), indexing="ij")
    x = (i.ravel() + jitter * rng.uniform(-1, 1, n_side * n_side)) * dx
    y = (j.ravel() + jitter * rng.uniform(-1, 1, n_side * n_side)) * dx
    n = len(x)
    h = hdx * dx
    m = 1000.0 * dx * dx
    pos4 = np.stack([x, y, np.zeros(n), np.full(n, h)], 1)
    vel4 = np.stack([rng.normal(0, vel, n), rng.normal(0, vel, n), np.zeros(n), np.full(n, m)], 1)
    rho = 1000.0 * (1 + 0.01 * rng.standard_normal(n))
    btype = (rng.random(n) < 0.2).astype(np.int32)
    d2 = (x[:, None] - x[None, :]) ** 2 + (y[:, None] - y[None, :]) ** 2
    adj = d2 < (2 * h) ** 2
    off = np.concatenate([[0], np.cumsum(adj.sum(1))]).astype(np.int64)
    nbr = np.concatenate([np.nonzero(adj[k])[0] for k in range(n)]).astype(np.int64)
    return pos4, vel4, rho, btype, off, nbr


def _loss_of(theta, X, E, dst, src, n, T, btype):
    p = M.unflatten(theta)
    Y, _ = M.forward(p, X, E, dst, src, n)
    return M.loss(Y, T, btype)[0]


def test_backward_matches_finite_differences():
    pos4, vel4, rho, btype, off, nbr = _graph()
    n = len(rho)
    X, E, dst, src = M.features(pos4, vel4, rho, btype, off, nbr, CONST)
    rng = np.random.Generator(np.random.PCG64(3))
    p = M.init_params(1)
    for k in p:
        if p[k].ndim == 1:
            p[k] = rng.normal(0, 0.1, p[k].shape)
    theta = M.flatten(p)
    T = rng.normal(0, 1, (n, 2))
    Y, cache = M.forward(p, X, E, dst, src, n)
    L, dY = M.loss(Y, T, btype)
    g = M.flatten(M.backward(p, cache, dY))
    L_ = M.layout()
    # a few parameters of every tensor
    idx = []
    for k, v in L_.items():
        if k == "total" or k == "b_out":
            continue
        idx += list(v[0] + rng.choice(v[1], 4, replace=False))
    idx += [L_["b_out"][0], L_["b_out"][0] + 1]
    eps = 1e-6
    bad = []
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += eps
        tm[i] -= eps
        fd = (_loss_of(tp, X, E, dst, src, n, T, btype) -
              _loss_of(tm, X, E, dst, src, n, T, btype)) / (2 * eps)
        if abs(fd - g[i]) > 1e-6 * max(1.0, abs(fd)) + 1e-8:
            bad.append((i, fd, g[i]))
    assert not bad, bad[:5]
    # the gradient is not trivially zero
    assert np.count_nonzero(np.abs(g[idx]) > 1e-8) > len(idx) // 2


def test_physics_edge_features_sum_to_sph_accelerations():
    """sum_j of the pressure / viscosity edge features times g equals the
    pressure and artificial-viscosity parts of MomentumEquation
    (pressure_funcs.pyx:154-245) on the same particles."""
    pos4, vel4, rho, btype, off, nbr = _graph(n_side=7, seed=5)
    n = len(rho)
    btype[:] = 0
    X, E, dst, src = M.features(pos4, vel4, rho, btype, off, nbr, CONST)
    # node features 6..9 are the sums (x 0.01); edge features 8..11 the scaled terms
    fx = X[:, 6:10] * 100.0 * CONST["g"]
    h = pos4[:, 3][dst]
    ps = 2.0 * CONST["g"] * h / CONST["u_scale"] ** 2
    fe = np.zeros((n, 4))
    np.add.at(fe, dst, E[:, 8:12] / ps[:, None])
    assert np.allclose(fe * CONST["g"], fx, rtol=1e-12, atol=1e-9)
    B = CONST["co"] ** 2 * CONST["rho0"] / CONST["gamma"]
    p = B * ((rho / CONST["rho0"]) ** 7 - 1)
    cs = CONST["co"] * (rho / CONST["rho0"]) ** 3
    arr = O.make_array(x=pos4[:, 0], y=pos4[:, 1], u=vel4[:, 0], v=vel4[:, 1], h=pos4[:, 3],
                       m=vel4[:, 3], rho=rho, p=p, cs=cs)
    a_p = O.momentum(2, arr, [arr], alpha=0.0, beta=0.0, eta=0.1)
    a_v = O.momentum(2, arr, [arr], alpha=0.1, beta=0.0, eta=0.1) - a_p
    scale = np.abs(a_p[:2]).max()
    assert np.max(np.abs(fx[:, 0:2].T - a_p[:2])) < 1e-9 * scale
    assert np.max(np.abs(fx[:, 2:4].T - a_v[:2])) < 1e-9 * max(np.abs(a_v[:2]).max(), 1e-12)
    assert np.abs(a_v[:2]).max() > 0


def test_reverse_edge_features():
    """Features of (j <- i) are those of (i <- j) with the antisymmetric
    entries negated and the two densities swapped (the GPU backward pass
    relies on it)."""
    pos4, vel4, rho, btype, off, nbr = _graph(seed=9)
    n = len(rho)
    X, E, dst, src = M.features(pos4, vel4, rho, btype, off, nbr, CONST)
    index = {(int(a), int(b)): k for k, (a, b) in enumerate(zip(dst, src))}
    neg = [0, 1, 3, 4, 8, 9, 10, 11]
    for k, (a, b) in enumerate(zip(dst, src)):
        r = E[index[(int(b), int(a))]]
        t = E[k].copy()
        t[neg] = -t[neg]
        t[5], t[6] = E[k][6], E[k][5]
        assert np.allclose(r, t, rtol=0, atol=1e-15), (a, b)
