"""FP64 numpy restatement of the edge-MLP surrogate (test infrastructure).

Only tests/, the graft entry's smoke check and bench.py's CPU leg use this
module, as the checker of the GPU network kernels (sph_mlp.hip).  It follows
the same definition: node and edge features, two message-passing rounds of
hidden width 128, decoder to 2 outputs, mean squared error over fluid nodes,
and the analytic backward pass (checked against central finite differences
in tests/test_oracle_mlp.py).
"""
import numpy as np

H = 128
NIN = 16
NE = 16
ROUNDS = 2
NOUT = 2


def layout():
    """Offsets of the flat FP32 parameter vector (same order as the GPU)."""
    o = 0
    L = {}

    def take(name, size):
        nonlocal o
        L[name] = (o, size)
        o += size
    take("w_in", NIN * H)
    take("b_in", H)
    for r in range(ROUNDS):
        take("w_ab%d" % r, H * 2 * H)
        take("w_c%d" % r, NE * H)
        take("b_e%d" % r, H)
        take("w_u%d" % r, 2 * H * H)
        take("c_u%d" % r, H)
    take("w_out", H * NOUT)
    take("b_out", 4)
    L["total"] = o
    return L


SHAPES = {"w_in": (NIN, H), "b_in": (H,), "w_out": (H, NOUT), "b_out": (4,)}
for _r in range(ROUNDS):
    SHAPES.update({"w_ab%d" % _r: (H, 2 * H), "w_c%d" % _r: (NE, H), "b_e%d" % _r: (H,),
                   "w_u%d" % _r: (2 * H, H), "c_u%d" % _r: (H,)})


def unflatten(theta):
    L = layout()
    return {k: np.asarray(theta[v[0]:v[0] + v[1]], dtype=np.float64).reshape(SHAPES[k])
            for k, v in L.items() if k != "total"}


def flatten(params):
    L = layout()
    out = np.zeros(L["total"], dtype=np.float64)
    for k, v in L.items():
        if k != "total":
            out[v[0]:v[0] + v[1]] = np.asarray(params[k]).ravel()
    return out


def init_params(seed=0):
    """He-style initial weights (the GPU model loads the same vector)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    p = {}
    for k, shp in SHAPES.items():
        if len(shp) == 1:
            p[k] = np.zeros(shp)
        else:
            p[k] = rng.normal(0.0, np.sqrt(2.0 / shp[0]), shp)
    p["w_out"] *= 0.1
    return p


# ---------------------------------------------------------------------------
def tait_p(rho, c):
    B = c["co"] ** 2 * c["rho0"] / c["gamma"]
    return B * ((rho / c["rho0"]) ** c["gamma"] - 1.0)


def spline(r, h, dim):
    """Cubic spline W and gradient factor (grad W = r_ab * gf), kernels.pyx:367-412."""
    q = r / h
    fac = 10.0 / (7.0 * np.pi * h * h) if dim == 2 else 1.0 / (np.pi * h ** 3)
    t = 2.0 - q
    w = np.where(q > 2, 0.0, np.where(q > 1, fac * 0.25 * t ** 3,
                                      fac * (1 - 1.5 * q * q * (1 - 0.5 * q))))
    with np.errstate(divide="ignore", invalid="ignore"):
        outer = fac * (-0.75 * t * t / (h * r))
        inner = np.where(q > 1e-14, fac * 3.0 * (0.75 * q - 1.0) / (h * h), 0.0)
    gf = np.where(q > 2, 0.0, np.where(q > 1, outer, inner))
    return w, gf


def features(pos4, vel4, rho, btype, off, nbr, c):
    """Node features X [n x 16] and edge features E [edges x 16] (FP64).
    c: rho0, co, gamma, u_scale, g, rho_scale, alpha, eta, dim."""
    n = len(rho)
    X = np.zeros((n, NIN))
    X[:, 0] = vel4[:, 0] / c["u_scale"]
    X[:, 1] = vel4[:, 1] / c["u_scale"]
    X[:, 2] = (rho / c["rho0"] - 1.0) * c["rho_scale"]
    B = c["co"] ** 2 * c["rho0"] / c["gamma"]
    X[:, 3] = tait_p(rho, c) / B
    X[:, 4] = (btype != 0).astype(np.float64)
    X[:, 5] = 1.0
    dst = np.repeat(np.arange(n), np.diff(off))
    src = np.asarray(nbr, dtype=np.int64)
    pi, pj = pos4[dst], pos4[src]
    vi, vj = vel4[dst], vel4[src]
    ri, rj = rho[dst], rho[src]
    h = 0.5 * (pi[:, 3] + pj[:, 3])
    d = pi[:, :3] - pj[:, :3]
    r = np.sqrt((d ** 2).sum(1))
    w, gf = spline(r, h, c["dim"])
    dv = vi[:, :3] - vj[:, :3]
    mbar = 0.5 * (vi[:, 3] + vj[:, 3])
    pri, prj = tait_p(ri, c), tait_p(rj, c)
    pt = -mbar * (pri / ri ** 2 + prj / rj ** 2) * gf / c["g"]
    dot = (dv * d).sum(1)
    cs = lambda rr: c["co"] * (rr / c["rho0"]) ** (0.5 * (c["gamma"] - 1.0))  # noqa: E731
    cbar = 0.5 * (cs(ri) + cs(rj))
    mu = h * dot / (r * r + c["eta"] ** 2 * h * h)
    visc = np.where(dot < 0, -mbar * (-c["alpha"] * cbar * mu) / (0.5 * (ri + rj)) * gf / c["g"],
                    0.0)
    # node sums of the pair pressure / viscosity accelerations (units of g);
    # the network reads them scaled by 0.01, the decoder's skip path unscaled
    acc = np.zeros((n, 4))
    np.add.at(acc[:, 0], dst, pt * d[:, 0])
    np.add.at(acc[:, 1], dst, pt * d[:, 1])
    np.add.at(acc[:, 2], dst, visc * d[:, 0])
    np.add.at(acc[:, 3], dst, visc * d[:, 1])
    X[:, 6:10] = 0.01 * acc
    ps = 2.0 * c["g"] * h / c["u_scale"] ** 2
    E = np.zeros((len(src), NE))
    E[:, 0] = d[:, 0] / h
    E[:, 1] = d[:, 1] / h
    E[:, 2] = r / h
    E[:, 3] = dv[:, 0] / c["u_scale"]
    E[:, 4] = dv[:, 1] / c["u_scale"]
    E[:, 5] = (ri / c["rho0"] - 1.0) * c["rho_scale"]
    E[:, 6] = (rj / c["rho0"] - 1.0) * c["rho_scale"]
    E[:, 7] = w * h * h
    E[:, 8] = pt * d[:, 0] * ps
    E[:, 9] = pt * d[:, 1] * ps
    E[:, 10] = visc * d[:, 0] * ps
    E[:, 11] = visc * d[:, 1] * ps
    E[:, 12] = 1.0
    return X, E, dst, src


def relu(x):
    return np.maximum(x, 0.0)


def bf16(x):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    b = a.view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


def forward(p, X, E, dst, src, n, round_bf16=False):
    """Returns Y [n x 2] and the cache for backward.  With round_bf16 the
    operands of every matrix product are rounded to bfloat16 where the GPU
    kernels store them in that format (X, H, A and the GEMM weights)."""
    rb = bf16 if round_bf16 else (lambda v: v)
    cache = {"X": rb(X), "E": E, "dst": dst, "src": src, "rb": rb}
    Z0 = cache["X"] @ rb(p["w_in"]) + p["b_in"]
    Hc = relu(Z0)
    cache["H0"] = Hc
    for r in range(ROUNDS):
        Hb = rb(Hc)
        PQ = Hb @ rb(p["w_ab%d" % r])
        P, Q = PQ[:, :H], PQ[:, H:]
        S = P[dst] + Q[src] + E @ p["w_c%d" % r] + p["b_e%d" % r]
        A = np.zeros((n, H))
        np.add.at(A, dst, relu(S))
        HA = np.concatenate([Hb, rb(A)], axis=1)
        U = relu(HA @ rb(p["w_u%d" % r]) + p["c_u%d" % r])
        cache["r%d" % r] = dict(H=Hb, S=S, HA=HA, U=U)
        Hc = Hc + U
    cache["H_last"] = Hc
    # skip path: the node's SPH pressure + viscosity acceleration / g
    Y = Hc @ p["w_out"] + p["b_out"][:NOUT] + 100.0 * (X[:, 6:8] + X[:, 8:10])
    return Y, cache


def loss(Y, T, btype):
    fl = (btype == 0)
    nf = max(int(fl.sum()), 1)
    e = np.where(fl[:, None], Y - T, 0.0)
    return float((e ** 2).sum() / nf), 2.0 * e / nf


def backward(p, cache, dY):
    g = {k: np.zeros_like(v) for k, v in p.items()}
    dst, src, E = cache["dst"], cache["src"], cache["E"]
    rb = cache["rb"]
    Hl = cache["H_last"]
    g["w_out"] = Hl.T @ dY
    g["b_out"][:NOUT] = dY.sum(0)
    dH = dY @ p["w_out"].T
    n = dH.shape[0]
    for r in reversed(range(ROUNDS)):
        c = cache["r%d" % r]
        dZ = dH * (c["U"] > 0)
        g["c_u%d" % r] = dZ.sum(0)
        dZb = rb(dZ)
        g["w_u%d" % r] = c["HA"].T @ dZb
        dHA = dZb @ rb(p["w_u%d" % r]).T
        dH = dH + dHA[:, :H]
        dA = dHA[:, H:]
        dS = dA[dst] * (c["S"] > 0)
        dP = np.zeros((n, H))
        dQ = np.zeros((n, H))
        np.add.at(dP, dst, dS)
        np.add.at(dQ, src, dS)
        g["w_c%d" % r] = E.T @ dS
        g["b_e%d" % r] = dS.sum(0)
        dPQ = rb(np.concatenate([dP, dQ], axis=1))
        g["w_ab%d" % r] = c["H"].T @ dPQ
        dH = dH + dPQ @ rb(p["w_ab%d" % r]).T
    dZ0 = dH * (cache["H0"] > 0)
    g["b_in"] = dZ0.sum(0)
    g["w_in"] = cache["X"].T @ rb(dZ0)
    return g
