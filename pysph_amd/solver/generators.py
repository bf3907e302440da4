"""Seeded particle-set generators of the benchmark configurations.

Positions are generated in FP64 and stored at the array's precision; the
oracle is fed the stored values widened to FP64, so both sides see the same
bits.  All randomness comes from numpy's PCG64 generator seeded explicitly.
"""
import math

import numpy as np

from ..base.particle_array import get_particle_array
from ..base.particle_types import ParticleType


def _store(dt, a):
    return np.asarray(a, dtype=np.float64).astype(dt).astype(np.float64)


def elliptical_drop(dx=0.025, rho0=1.0, hdx=1.3, u0=100.0, precision="double"):
    """2D elliptical drop: lattice (i dx, j dx) with i^2 + j^2 <= (1/dx)^2
    (inside the unit circle), m = rho0 dx^2, h = hdx dx, u = -u0 x,
    v = u0 y, rho = rho0, p = cs = 0 (fluid_solver.py:16-51 is the
    reference's generator; it uses h = 2 dx and a non-zero initial p, cs)."""
    R = int(round(1.0 / dx))
    i, j = np.meshgrid(np.arange(-R, R + 1), np.arange(-R, R + 1), indexing="ij")
    i = i.ravel()
    j = j.ravel()
    keep = i * i + j * j <= R * R
    dt = np.float64 if precision == "double" else np.float32
    x = _store(dt, i[keep] * dx)
    y = _store(dt, j[keep] * dx)
    n = len(x)
    return get_particle_array(
        cl_precision=precision, name="fluid", type=ParticleType.Fluid,
        x=x, y=y, z=np.zeros(n), u=-u0 * x, v=u0 * y, w=np.zeros(n),
        m=np.full(n, rho0 * dx * dx), h=np.full(n, hdx * dx), rho=np.full(n, rho0),
        p=np.zeros(n), cs=np.zeros(n))


def dam_break_3d(dx=1.0 / 135, column=(81, 81, 162), tank=(216, 81, 162), layers=3,
                 hdx=1.2, rho0=1000.0, jitter=0.01, seed=0, precision="single"):
    """3D dam break: a fluid column of column[0] x column[1] x column[2]
    lattice cells ((i+1/2)dx, ...) with uniform jitter of +-jitter*dx, inside
    a tank of tank[...] cells lined with `layers` layers of fixed boundary
    particles on the floor and the four walls (open top).  Returns (fluid,
    boundary, info)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    dt = np.float64 if precision == "double" else np.float32
    nx, ny, nz = column
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    pos = [(a.ravel() + 0.5) * dx for a in (i, j, k)]
    nf = len(pos[0])
    pos = [p + rng.uniform(-jitter * dx, jitter * dx, nf) for p in pos]
    fx, fy, fz = (_store(dt, p) for p in pos)
    m = rho0 * dx ** 3
    h = hdx * dx
    fluid = get_particle_array(
        cl_precision=precision, name="fluid", type=ParticleType.Fluid,
        x=fx, y=fy, z=fz, u=np.zeros(nf), v=np.zeros(nf), w=np.zeros(nf),
        m=np.full(nf, m), h=np.full(nf, h), rho=np.full(nf, rho0), p=np.zeros(nf),
        cs=np.zeros(nf))
    tx, ty, tz = tank
    L = layers
    i, j, k = np.meshgrid(np.arange(-L, tx + L), np.arange(-L, ty + L), np.arange(-L, tz),
                          indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    outside = (i < 0) | (i >= tx) | (j < 0) | (j >= ty) | (k < 0)
    bx, by, bz = ((a[outside] + 0.5) * dx for a in (i, j, k))
    nb = len(bx)
    boundary = get_particle_array(
        cl_precision=precision, name="boundary", type=ParticleType.Solid,
        x=_store(dt, bx), y=_store(dt, by), z=_store(dt, bz), u=np.zeros(nb), v=np.zeros(nb),
        w=np.zeros(nb), m=np.full(nb, m), h=np.full(nb, h), rho=np.full(nb, rho0),
        p=np.zeros(nb), cs=np.zeros(nb))
    H = nz * dx
    g = 9.81
    co = 10.0 * math.sqrt(2.0 * g * H)
    info = dict(dx=dx, h=h, m=m, rho0=rho0, co=co, gamma=7.0, alpha=0.1, beta=0.0,
                eta=0.1, eps=0.5, g=(0.0, 0.0, -g), dt=0.25 * h / co, n_fluid=nf,
                n_boundary=nb, domain=((-L * dx, -L * dx, -L * dx),
                                       ((tx + L) * dx, (ty + L) * dx, (tz + 2 * L) * dx)))
    return fluid, boundary, info


def dam_break_2d(width=0.5, height=1.0, dx=0.01, tank_width=None, layers=3, hdx=1.3,
                 rho0=1000.0, jitter=0.0, seed=0, precision="single"):
    """2D dam break: a width x height water column at the left wall of a
    tank (default 4x the column width, as tall as the column) lined with
    `layers` boundary layers on the floor and both walls.  Returns (fluid,
    boundary, info)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    dt = np.float64 if precision == "double" else np.float32
    tank_width = tank_width or 4.0 * width
    nx, ny = int(round(width / dx)), int(round(height / dx))
    tx, ty = int(round(tank_width / dx)), ny
    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    x = (i.ravel() + 0.5) * dx
    y = (j.ravel() + 0.5) * dx
    nf = len(x)
    if jitter:
        x = x + rng.uniform(-jitter * dx, jitter * dx, nf)
        y = y + rng.uniform(-jitter * dx, jitter * dx, nf)
    m = rho0 * dx * dx
    h = hdx * dx
    fluid = get_particle_array(
        cl_precision=precision, name="fluid", type=ParticleType.Fluid,
        x=_store(dt, x), y=_store(dt, y), z=np.zeros(nf), u=np.zeros(nf), v=np.zeros(nf),
        w=np.zeros(nf), m=np.full(nf, m), h=np.full(nf, h), rho=np.full(nf, rho0),
        p=np.zeros(nf), cs=np.zeros(nf))
    L = layers
    i, j = np.meshgrid(np.arange(-L, tx + L), np.arange(-L, ty), indexing="ij")
    i, j = i.ravel(), j.ravel()
    outside = (i < 0) | (i >= tx) | (j < 0)
    bx, by = ((a[outside] + 0.5) * dx for a in (i, j))
    nb = len(bx)
    boundary = get_particle_array(
        cl_precision=precision, name="boundary", type=ParticleType.Solid,
        x=_store(dt, bx), y=_store(dt, by), z=np.zeros(nb), u=np.zeros(nb), v=np.zeros(nb),
        w=np.zeros(nb), m=np.full(nb, m), h=np.full(nb, h), rho=np.full(nb, rho0),
        p=np.zeros(nb), cs=np.zeros(nb))
    g = 9.81
    co = 10.0 * math.sqrt(2.0 * g * height)
    info = dict(dx=dx, h=h, m=m, rho0=rho0, co=co, gamma=7.0, alpha=0.1, beta=0.0,
                eta=0.1, eps=0.5, g=(0.0, -g, 0.0), dt=0.25 * h / co, n_fluid=nf,
                n_boundary=nb, domain=((-L * dx, -L * dx, 0.0),
                                       ((tx + L) * dx, (ty + 2 * L) * dx, 0.0)))
    return fluid, boundary, info


def cube_h(n, neighbours=50):
    """h such that a sphere of radius 2h holds `neighbours` particles of a
    uniform unit-cube set of n: (4/3) pi (2h)^3 n = neighbours."""
    return 0.5 * (3.0 * neighbours / (4.0 * math.pi * n)) ** (1.0 / 3.0)


def cube_points(n, kind="uniform", seed=0, n_clusters=64, sigma=0.02, cluster_frac=0.073):
    """Points in [0,1)^3: 'uniform', or 'clustered' = a mixture of a uniform
    remainder and n_clusters Gaussian clusters (sigma) with centres in
    U[0.1,0.9]^3 holding cluster_frac of the points, clipped into the cube."""
    rng = np.random.Generator(np.random.PCG64(seed))
    if kind == "uniform":
        return rng.random((3, n))
    nc = int(round(cluster_frac * n))
    centres = rng.uniform(0.1, 0.9, (n_clusters, 3))
    which = rng.integers(0, n_clusters, nc)
    cl = centres[which] + rng.normal(0.0, sigma, (nc, 3))
    un = rng.random((n - nc, 3))
    pts = np.concatenate([cl, un]).T
    # 1 - 2^-24 is the largest value below 1 in FP32 and FP64 alike
    return np.clip(pts, 0.0, 1.0 - 2.0 ** -24)


def cube_array(n, kind="uniform", seed=0, neighbours=50, precision="single"):
    """Unit-cube particle set of the neighbour/density microbenchmark:
    uniform masses 1/n, h from cube_h."""
    dt = np.float64 if precision == "double" else np.float32
    x, y, z = (_store(dt, a) for a in cube_points(n, kind, seed))
    return get_particle_array(cl_precision=precision, name="cube", type=ParticleType.Fluid,
                              x=x, y=y, z=z, m=np.full(n, 1.0 / n),
                              h=np.full(n, cube_h(n, neighbours)))


def elliptical_drop_2d(n_target=5000, aspect=1.0, u0=100.0, hdx=1.3, rho0=1.0, alpha=0.1,
                       c_ratio=14.0, precision="single"):
    """2D elliptical drop of about n_target particles on a square lattice
    inside the ellipse x^2 + (y/aspect)^2 <= 1, u = -u0 x, v = u0 y,
    c0 = c_ratio * u0 (the reference's drop: c0 = 1400 at u0 = 100),
    dt = 0.25 h / c0.  Returns (fluid, None, info)."""
    a, b = 1.0, float(aspect)
    dx = math.sqrt(math.pi * a * b / n_target)
    R = int(math.ceil(max(a, b) / dx)) + 1
    i, j = np.meshgrid(np.arange(-R, R + 1), np.arange(-R, R + 1), indexing="ij")
    x, y = i.ravel() * dx, j.ravel() * dx
    keep = (x / a) ** 2 + (y / b) ** 2 <= 1.0
    x, y = x[keep], y[keep]
    n = len(x)
    dt_ = np.float64 if precision == "double" else np.float32
    h = hdx * dx
    co = c_ratio * u0
    fluid = get_particle_array(
        cl_precision=precision, name="fluid", type=ParticleType.Fluid,
        x=_store(dt_, x), y=_store(dt_, y), z=np.zeros(n), u=-u0 * x, v=u0 * y,
        w=np.zeros(n), m=np.full(n, rho0 * dx * dx), h=np.full(n, h), rho=np.full(n, rho0),
        p=np.zeros(n), cs=np.zeros(n))
    info = dict(dx=dx, h=h, rho0=rho0, co=co, gamma=7.0, alpha=alpha, beta=0.0, eta=0.1,
                eps=0.0, g=None, dt=0.25 * h / co, n_fluid=n, n_boundary=0, u0=u0,
                domain=((-1.2 * a, -1.6 * b, 0.0), (1.2 * a, 1.6 * b, 0.0)))
    return fluid, None, info


def batch_scenes(n_scenes=512, seed=0, n_min=1000, n_max=64000, precision="single"):
    """Scenes of the batched benchmark: half elliptical drops, half 2D dam
    breaks, fluid counts log-uniform in [n_min, n_max], per-scene h/dx in
    [1.2, 1.5], drop aspect ratio in [0.5, 1], column height/width in
    [0.5, 2] (width in [0.3, 1] m) and drop velocity scale u0 in [50, 150],
    all drawn from `seed`.  Returns a list of (fluid, boundary, info)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for k in range(n_scenes):
        n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        hdx = float(rng.uniform(1.2, 1.5))
        if k % 2 == 0:
            f, b, info = elliptical_drop_2d(n, aspect=float(rng.uniform(0.5, 1.0)),
                                            u0=float(rng.uniform(50.0, 150.0)), hdx=hdx,
                                            precision=precision)
        else:
            width = float(rng.uniform(0.3, 1.0))
            height = width * float(rng.uniform(0.5, 2.0))
            dx = math.sqrt(width * height / n)
            f, b, info = dam_break_2d(width=width, height=height, dx=dx, hdx=hdx,
                                      precision=precision)
        info["kind"] = "drop" if k % 2 == 0 else "dam"
        out.append((f, b, info))
    return out
