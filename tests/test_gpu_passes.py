"""The pair-sum passes on the GPU against the FP64 oracle.

Tolerances: density and acceleration within 1e-5
relative in FP32 storage (acceleration normwise, ||a_gpu - a_ref|| /
||a_ref||, because a near-hydrostatic pressure-gradient sum cancels);
FP64 storage within 1e-12.  Inputs are the same stored bits on both sides.
"""
import numpy as np
import pytest
import torch

import pysph_amd.base.api as base
import pysph_amd.sph.api as sph
import pysph_amd.solver.api as solver
from oracle import oracle as O
from pysph_amd.sph import _ops
from pysph_amd import _abi
from pysph_amd.solver import generators as gen
from tests.util import max_rel, rel_norm, to_oracle

pytestmark = pytest.mark.gpu

TOL = {"single": 1e-5, "double": 1e-12}


def _dam2d(precision, seed=0):
    f, b, info = gen.dam_break_2d(width=0.3, height=0.4, dx=0.01, jitter=0.05, seed=seed,
                                  precision=precision)
    return f, b, info


def _randomise_state(pa, seed, info, dim):
    rng = np.random.Generator(np.random.PCG64(seed))
    n = pa.get_number_of_particles()
    dt = np.float32 if pa.dtype.itemsize == 4 else np.float64
    vel = {c: (rng.normal(0, 0.5, n) if k < dim else np.zeros(n)).astype(dt)
           for k, c in enumerate("uvw")}
    rho = (info["rho0"] * (1 + 0.01 * rng.standard_normal(n))).astype(dt)
    pa.set(u=vel["u"], v=vel["v"], w=vel["w"], rho=rho)


def _eos(pa, info, ptype):
    ref = to_oracle(pa, ptype)
    p, cs = O.tait(ref, info["co"], info["rho0"], 7.0)
    pa.set(p=p.astype(np.float64), cs=cs.astype(np.float64))


@pytest.mark.parametrize("precision", ["single", "double"])
def test_sphrho_two_sources(precision):
    f, b, info = _dam2d(precision)
    particles = base.Particles([f, b])
    rho_f = sph.SPHRho.withargs()
    calc = sph.SPHCalc(particles, [f, b], f, base.CubicSplineKernel(dim=2),
                       [rho_f.get_func(f, f), rho_f.get_func(b, f)], updates=["rho"])
    calc.sph("rho")
    got = f.get("rho")
    ref = O.sphrho(2, to_oracle(f), [to_oracle(f), to_oracle(b)],
                   cell_size=particles.cell_manager.cell_size)
    assert max_rel(got, ref) < TOL[precision]


@pytest.mark.parametrize("case", ["varied_fluid_h", "uniform_but_different_h"])
def test_sphrho_sources_of_uniform_and_varied_h(case):
    """The FP32 density pass reads one (x,y,z,m) record per pair from a
    source whose h is uniform (h taken from its h range) and the pos4 / vel4
    pair otherwise; both against the oracle, with fluid and boundary h
    different from each other."""
    f, b, info = _dam2d("single", seed=3)
    rng = np.random.Generator(np.random.PCG64(9))
    hf = f.get("h")
    if case == "varied_fluid_h":
        f.set(h=(hf * (0.9 + 0.2 * rng.random(len(hf)))).astype(np.float32))
    else:
        f.set(h=(hf * 1.1).astype(np.float32))
    b.set(h=(b.get("h") * 0.95).astype(np.float32))
    particles = base.Particles([f, b])
    rho_f = sph.SPHRho.withargs()
    for dst, srcs in ((f, [f, b]), (b, [f, b])):
        calc = sph.SPHCalc(particles, srcs, dst, base.CubicSplineKernel(dim=2),
                           [rho_f.get_func(s, dst) for s in srcs], updates=["rho"])
        calc.sph("rho")
        ref = O.sphrho(2, to_oracle(dst), [to_oracle(s) for s in srcs],
                       cell_size=particles.cell_manager.cell_size)
        assert max_rel(dst.get("rho"), ref) < TOL["single"]


@pytest.mark.parametrize("precision", ["single", "double"])
def test_sphrho_exclude_self(precision):
    """exclude_self (sph_func.pyx:386-389) drops the self pair of the source
    that is the destination and leaves other sources alone: the fluid sum
    is the full oracle sum less m_a W(0, h_a) = m_a 10/(7 pi h_a^2) (2D),
    and a source that is not the destination cannot be flagged."""
    f, b, info = _dam2d(precision, seed=2)
    particles = base.Particles([f, b])
    rho_f = sph.SPHRho.withargs(exclude_self=True)
    calc = sph.SPHCalc(particles, [f, b], f, base.CubicSplineKernel(dim=2),
                       [rho_f.get_func(f, f), rho_f.get_func(b, f)], updates=["rho"])
    calc.sph("rho")
    cm = particles.cell_manager
    full = O.sphrho(2, to_oracle(f), [to_oracle(f), to_oracle(b)], cell_size=cm.cell_size)
    h = f.get("h").astype(np.float64)
    w0 = f.get("m").astype(np.float64) * 10.0 / (7.0 * np.pi * h * h)
    assert max_rel(f.get("rho"), full - w0) < TOL[precision]
    # the boundary as destination: its own self pair goes, the fluid source
    # is not the destination and keeps every pair
    calc_b = sph.SPHCalc(particles, [f, b], b, base.CubicSplineKernel(dim=2),
                         [rho_f.get_func(f, b), rho_f.get_func(b, b)], updates=["rho"])
    calc_b.sph("rho")
    full_b = O.sphrho(2, to_oracle(b), [to_oracle(f), to_oracle(b)], cell_size=cm.cell_size)
    hb = b.get("h").astype(np.float64)
    w0b = b.get("m").astype(np.float64) * 10.0 / (7.0 * np.pi * hb * hb)
    assert max_rel(b.get("rho"), full_b - w0b) < TOL[precision]
    prm = _ops.make_params(2, exclude_self=0b10)  # source 1 (boundary) is not f
    with pytest.raises(Exception, match="exclude_self"):
        _ops.density(cm, prm, f, [f, b], f.get_carray("rho"))


def test_reference_four_particles_exclude_self():
    """The unit square of density_funcs' test (h = 1, m = 1) without the
    self term: rho = fac (2 W(1) + W(sqrt 2)) = fac (0.5 + 0.25 (2 - sqrt 2)^3)."""
    x = np.array([0, 0, 1, 1.0])
    y = np.array([0, 1, 1, 0.0])
    pa = base.get_particle_array(name="test", x=x, y=y, z=np.zeros(4), m=np.ones(4))
    particles = base.Particles([pa])
    calc = sph.SPHCalc(particles, [pa], pa, base.CubicSplineKernel(dim=2),
                       [sph.SPHRho.withargs(exclude_self=True).get_func(pa, pa)],
                       updates=["rho"])
    pa.add_property({"name": "_tmpx"})
    calc.sph("_tmpx")
    fac = 10 / (7 * np.pi)
    assert np.allclose(pa.get("_tmpx"), fac * (0.5 + 0.25 * (2 - np.sqrt(2)) ** 3),
                       rtol=0, atol=1e-14)


def _tag(pa, frac, value, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    n = pa.get_number_of_particles()
    t = np.where(rng.random(n) < frac, value, base.LocalReal).astype(np.int64)
    pa.set(tag=t)
    # setting tags re-aligns the particles (LocalReal first)
    t = pa.get("tag", only_real_particles=False)
    assert np.all(t[:pa.num_real_particles] == base.LocalReal)
    assert np.all(t[pa.num_real_particles:] != base.LocalReal)
    return t


def _aux4(pa):
    """(rho, cs, p/rho^2, p) per slot as the density pass writes it (p/rho^2
    formed in FP64 from the stored FP32 values)."""
    n = pa.get_number_of_particles()
    rho = pa.get_carray("rho")[:n].double()
    p = pa.get_carray("p")[:n].double()
    cs = pa.get_carray("cs")[:n].double()
    return torch.stack([rho, cs, p / (rho * rho), p], 1).float().contiguous()


@pytest.mark.parametrize("precision", ["single", "double"])
def test_non_local_destinations_get_zero(precision):
    """Destinations that are not LocalReal get 0 in every output and still
    act as sources (sph_func.pyx:213-219): 10 % of the fluid RemoteReal and
    10 % of the boundary LocalDummy.  Momentum + XSPH through the SoA kernels
    and (FP32) the packed kernels, then SPHRho + Tait, each against the
    oracle with the same tags."""
    f, b, info = _dam2d(precision, seed=5)
    _randomise_state(f, 6, info, 2)
    _eos(f, info, 0)
    _eos(b, info, 1)
    tf = _tag(f, 0.1, base.RemoteReal, 7)
    tb = _tag(b, 0.1, base.LocalDummy, 8)
    assert 0 < np.count_nonzero(tf) < len(tf) and 0 < np.count_nonzero(tb) < len(tb)
    particles = base.Particles([f, b])
    cm = particles.cell_manager
    cm.update()
    rf, rb = to_oracle(f), to_oracle(b)
    tol = TOL[precision]

    # momentum + XSPH on the preset state
    prm = _ops.make_params(2, alpha=0.1, beta=0.0, eta=0.1, eps=0.5,
                           pair_prec=_abi.SPH_PAIR_F64)
    flags = [_abi.SPH_SRC_MOMENTUM | _abi.SPH_SRC_XSPH, _abi.SPH_SRC_MOMENTUM]
    a_o = O.momentum(2, rf, [rf, rb], 0.1, 0.0, 0.1, cell_size=cm.cell_size)[:2]
    d_o = O.xsph(2, rf, [rf], 0.5, cell_size=cm.cell_size)[:2]
    assert np.all(a_o[:, tf != 0] == 0) and np.all(d_o[:, tf != 0] == 0)
    order = f._order_np()
    n = f.get_number_of_particles()
    overrides = [None]
    if precision == "single":
        ov = _ops.packed_override([f, b])
        for pa in (f, b):
            ov[id(pa)]["aux4"] = _aux4(pa)
        overrides.append(ov)
    for ov in overrides:
        acc = [f.get_carray("u").new_full((n,), 7.0) for _ in range(3)]
        dv = [f.get_carray("u").new_full((n,), 7.0) for _ in range(3)]
        _ops.momentum_xsph(cm, prm, f, [f, b], flags, acc=acc, dv=dv, override=ov)
        def user(t):
            a = t.cpu().numpy().astype(np.float64)
            out = np.empty_like(a)
            out[order] = a
            return out
        a_g = np.stack([user(t) for t in acc[:2]])
        d_g = np.stack([user(t) for t in dv[:2]])
        assert np.all(a_g[:, tf != 0] == 0) and np.all(d_g[:, tf != 0] == 0)
        assert rel_norm(a_g, a_o) < tol
        assert rel_norm(d_g, d_o) < tol

    # SPHRho + Tait: rho, p and cs of non-local destinations are 0
    prm = _ops.make_params(2, co=info["co"], ro=info["rho0"], gamma=7.0)
    for dst, tags, ref in ((f, tf, rf), (b, tb, rb)):
        rho_o = O.sphrho(2, ref, [rf, rb], cell_size=cm.cell_size)
        assert np.all(rho_o[tags != 0] == 0)
        _ops.density(cm, prm, dst, [f, b], dst.get_carray("rho"), dst.get_carray("p"),
                     dst.get_carray("cs"))
        rho = dst.get("rho", only_real_particles=False)
        assert np.all(rho[tags != 0] == 0)
        assert max_rel(rho, rho_o, floor=1.0) < tol
        p_o, cs_o = O.tait(to_oracle(dst), info["co"], info["rho0"], 7.0)
        for got, want in ((dst.get("p", only_real_particles=False), p_o),
                          (dst.get("cs", only_real_particles=False), cs_o)):
            assert np.all(got[tags != 0] == 0)
            assert max_rel(got, want, floor=float(np.max(np.abs(want)))) < 1e-6


@pytest.mark.parametrize("precision", ["single", "double"])
def test_tait_closed_form(precision):
    """p = B((rho/ro)^7 - 1), cs = co (rho/ro)^3 against an independent
    evaluation of the closed form (eos_funcs.pyx:105-116)."""
    rho = np.linspace(990.0, 1010.0, 1001)
    pa = base.get_particle_array(cl_precision=precision, x=np.zeros(1001), rho=rho)
    prm = _ops.make_params(3, co=48.52, ro=1000.0, gamma=7.0)
    p, cs = pa.get_carray("p"), pa.get_carray("cs")
    _ops.tait(prm, pa, p, cs)
    r = pa.get("rho")
    B = 48.52 ** 2 * 1000.0 / 7.0
    ratio = r / 1000.0
    p_ref = B * (ratio ** 7 - 1.0)
    cs_ref = 48.52 * ratio ** 3
    tol = 1e-6 if precision == "single" else 1e-13
    assert np.max(np.abs(pa.get("p") - p_ref)) <= tol * B
    assert max_rel(pa.get("cs"), cs_ref) < tol


@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("pair_prec", [_abi.SPH_PAIR_F64, _abi.SPH_PAIR_NATIVE])
def test_momentum_xsph_vs_oracle(precision, pair_prec):
    f, b, info = _dam2d(precision, seed=3)
    _randomise_state(f, 4, info, 2)
    _eos(f, info, 0)
    _eos(b, info, 1)
    particles = base.Particles([f, b])
    cm = particles.cell_manager
    cm.update()
    prm = _ops.make_params(2, alpha=0.1, beta=0.0, eta=0.1, eps=0.5, pair_prec=pair_prec)
    n = f.get_number_of_particles()
    acc = [f.get_carray("u").new_empty(n) for _ in range(3)]
    dv = [f.get_carray("u").new_empty(n) for _ in range(3)]
    flags = [_abi.SPH_SRC_MOMENTUM | _abi.SPH_SRC_XSPH, _abi.SPH_SRC_MOMENTUM]
    _ops.momentum_xsph(cm, prm, f, [f, b], flags, acc=acc, dv=dv)
    order = f._order_np()
    def user(t):
        a = t.cpu().numpy().astype(np.float64)
        out = np.empty_like(a)
        out[order] = a
        return out
    a_g = np.stack([user(t) for t in acc[:2]])
    d_g = np.stack([user(t) for t in dv[:2]])
    rf, rb = to_oracle(f), to_oracle(b)
    a_o = O.momentum(2, rf, [rf, rb], 0.1, 0.0, 0.1, cell_size=cm.cell_size)[:2]
    d_o = O.xsph(2, rf, [rf], 0.5, cell_size=cm.cell_size)[:2]
    tol = TOL[precision]
    print("momentum/xsph %s pair_prec=%d: normwise rel err %.3e / %.3e"
          % (precision, pair_prec, rel_norm(a_g, a_o), rel_norm(d_g, d_o)))
    assert rel_norm(a_g, a_o) < tol
    assert rel_norm(d_g, d_o) < tol


@pytest.mark.parametrize("precision,pair_prec", [("single", _abi.SPH_PAIR_F64),
                                                 ("single", _abi.SPH_PAIR_NATIVE),
                                                 ("double", _abi.SPH_PAIR_F64)])
def test_momentum_3d_dam_break_hydrostatic(precision, pair_prec, monkeypatch):
    """3D dam break at rest: the pressure sum cancels strongly.  Normwise
    within the contract's 1e-5 (FP32 storage) with FP64 and with all-FP32
    pair arithmetic; the error is printed."""
    monkeypatch.setattr(_ops, "DEFAULT_PAIR_PREC", pair_prec)
    f, b, info = gen.dam_break_3d(dx=1.0 / 40, column=(12, 12, 24), tank=(30, 12, 24),
                                  seed=2, precision=precision)
    particles = base.Particles([f, b])
    s = solver.wcsph_solver(3, co=info["co"], ro=info["rho0"], alpha=0.1, gravity=info["g"],
                            xsph_eps=0.5)
    s.setup_integrator(particles)
    # run the density + eos calcs through the generic path
    for calc in s.integrator.calcs[:4]:
        calc.sph(*calc.updates)
    rf, rb = to_oracle(f), to_oracle(b)
    rho_o = O.sphrho(3, rf, [rf, rb], cell_size=particles.cell_manager.cell_size)
    assert max_rel(f.get("rho"), rho_o) < TOL[precision]
    mom = s.integrator.calcs[4]
    mom.sph("_k1_u40", "_k1_v41", "_k1_w42")
    a_g = np.stack([f.get("_k1_u40"), f.get("_k1_v41"), f.get("_k1_w42")])
    a_o = O.momentum(3, rf, [rf, rb], 0.1, 0.0, 0.1, cell_size=particles.cell_manager.cell_size)
    err = rel_norm(a_g, a_o)
    print("hydrostatic momentum %s pair_prec=%d: normwise rel err %.3e, max rel rho err %.3e"
          % (precision, pair_prec, err, max_rel(f.get("rho"), rho_o)))
    assert err < TOL[precision]


def test_reference_four_particle_density_and_momentum():
    """sph/tests/test_density_funcs.py:86-137 and test_pressure_funcs.py:
    209-301: unit square, h = 1, m = 1, u = (1,0,0,-1), p = (0,0,1,1)."""
    x = np.array([0, 0, 1, 1.0])
    y = np.array([0, 1, 1, 0.0])
    pa = base.get_particle_array(name="test", x=x, y=y, z=np.zeros(4), m=np.ones(4),
                                 u=np.array([1, 0, 0, -1.0]), p=np.array([0, 0, 1, 1.0]))
    particles = base.Particles([pa])
    k = base.CubicSplineKernel(dim=2)
    calc = sph.SPHCalc(particles, [pa], pa, k, [sph.SPHRho.withargs().get_func(pa, pa)],
                       updates=["rho"])
    pa.add_property({"name": "_tmpx"})
    calc.sph("_tmpx")
    fac = 10 / (7 * np.pi)
    closed = fac * (1 + 2 * 0.25 + 0.25 * (2 - np.sqrt(2)) ** 3)
    assert np.allclose(pa.get("_tmpx"), closed, rtol=0, atol=1e-14)
    mom = sph.MomentumEquation.withargs(alpha=1.0, beta=1.0, gamma=1.4, eta=0.1)
    func = mom.get_func(pa, pa)
    func.kernel = k
    func.nbr_locator = base.Particles.get_neighbor_particle_locator(pa, pa)
    for nm in ("tmpx", "tmpy", "tmpz"):
        pa.add_property({"name": nm})
    func.nbr_locator.cell_manager.update()
    func.eval(k, pa.get_carray("tmpx"), pa.get_carray("tmpy"), pa.get_carray("tmpz"))
    ref = O.momentum(2, to_oracle(pa), [to_oracle(pa)], 1.0, 1.0, 0.1)
    got = np.stack([pa.get("tmpx"), pa.get("tmpy"), pa.get("tmpz")])
    # reference test: assertAlmostEqual to 7 places
    assert np.allclose(got[:2], ref[:2], atol=1e-7, rtol=0)


@pytest.mark.parametrize("cells_per_radius", [1, 2])
def test_fast_path_particles_outside_fixed_domain(cells_per_radius):
    """Particles outside a fixed binning domain are clamped into the border
    cells; the culled stencil rows must still reach them (and they must
    still find themselves)."""
    rng = np.random.Generator(np.random.PCG64(11))
    n = 4000
    x, y = rng.random(n), rng.random(n)
    pa = base.get_particle_array(cl_precision="single", name="cloud", x=x, y=y,
                                 h=np.full(n, 0.02), m=np.full(n, 1.0 / n))
    particles = base.Particles([pa], update_particles=False)
    cm = particles.cell_manager
    cm.set_fixed_domain((0.3, 0.3, 0.0), (0.7, 0.7, 0.0))
    cm.set_cells_per_radius(cells_per_radius)
    calc = sph.SPHCalc(particles, [pa], pa, base.CubicSplineKernel(dim=2),
                       [sph.SPHRho.withargs().get_func(pa, pa)], updates=["rho"])
    calc.sph("rho")
    ref = to_oracle(pa)
    rho_o = O.sphrho(2, ref, [ref])
    assert max_rel(pa.get("rho"), rho_o) < 1e-5


def test_fused_step_with_particles_outside_fixed_domain():
    f, b, info = gen.dam_break_2d(width=0.2, height=0.3, dx=0.01, jitter=0.05, seed=5,
                                  precision="single")
    particles = base.Particles([f, b], update_particles=False)
    cm = particles.cell_manager
    # a domain that leaves the top of the column and the right wall outside
    cm.set_fixed_domain((-0.03, -0.03, 0.0), (0.5, 0.2, 0.0))
    cm.check_jumps = False
    s = solver.wcsph_solver(2, co=info["co"], ro=info["rho0"], alpha=0.1, gravity=info["g"],
                            xsph_eps=0.5)
    s.setup_integrator(particles)
    s.integrator.plan.set_cells_per_radius(2)
    rf, rb = to_oracle(f), to_oracle(b)
    s.integrator.integrate(info["dt"])
    sch = O.scheme(2, integrator="rk2", dt=info["dt"], co=info["co"], ro=info["rho0"],
                   gamma=7.0, alpha=0.1, beta=0.0, eta=0.1, gravity=info["g"], xsph_eps=0.5)
    O.scheme_step(sch, [rf, rb], 1)
    assert max_rel(f.get("rho"), rf["rho"]) < 1e-5
    assert rel_norm(np.stack([f.get("u"), f.get("v")]), np.stack([rf["u"], rf["v"]])) < 1e-4


@pytest.mark.parametrize("precision", ["single", "double"])
def test_eval_accumulates_into_local_and_zeroes_the_rest(precision):
    """SPHFunction.eval (sph_func.pyx:199-234) adds the function's sum to
    the outputs of LocalReal destinations and sets the outputs of the other
    destinations to 0, whatever they held before."""
    f, b, info = _dam2d(precision, seed=12)
    _randomise_state(f, 13, info, 2)
    _eos(f, info, 0)
    tf = _tag(f, 0.15, base.RemoteReal, 14)
    assert 0 < np.count_nonzero(tf) < len(tf)
    k = base.CubicSplineKernel(dim=2)
    func = sph.MomentumEquation.withargs(alpha=0.1, beta=0.0, eta=0.1, dim=2).get_func(f, f)
    func.kernel = k
    func.nbr_locator = base.Particles.get_neighbor_particle_locator(f, f)
    func.nbr_locator.cell_manager.update()
    n = f.get_number_of_particles()
    rng = np.random.Generator(np.random.PCG64(15))
    pre = {nm: rng.normal(0, 3.0, n) for nm in ("tmpx", "tmpy", "tmpz")}
    for nm, v in pre.items():
        f.add_property({"name": nm})
        f.set(**{nm: v})
    func.eval(k, f.get_carray("tmpx"), f.get_carray("tmpy"), f.get_carray("tmpz"))
    ref = O.momentum(2, to_oracle(f), [to_oracle(f)], 0.1, 0.0, 0.1)
    local = tf == base.LocalReal
    every = {"only_real_particles": False}
    for c, nm in enumerate(("tmpx", "tmpy")):
        got = f.get(nm, **every).astype(np.float64)
        assert np.all(got[~local] == 0.0), nm
        want = pre[nm].astype(got.dtype if f.dtype.itemsize == 8 else np.float32).astype(
            np.float64)[local] + ref[c][local]
        scale = np.max(np.abs(ref[c][local]))
        assert np.max(np.abs(got[local] - want)) < TOL[precision] * scale * 10, nm


def test_density_xyz_records_with_uniform_mass():
    """A density source of uniform h and uniform mass is read from the
    12-byte (x, y, z) copy, its mass taken from the pack's mass-uniformity
    pair: the same bits as the 16-byte (x, y, z, m) records (pair withheld),
    with one and two sources and in the fused launch.  A source whose masses
    differ keeps the 16-byte records (its pair is unequal) and matches the
    oracle."""
    f, b, info = gen.dam_break_3d(dx=1.0 / 40, column=(10, 10, 20), tank=(24, 10, 20), seed=7,
                                  precision="single")
    particles = base.Particles([f, b])
    cm = particles.cell_manager
    cm.set_cells_per_radius(2)
    cm.update()
    prm = _ops.make_params(3, co=info["co"], ro=info["rho0"], gamma=7.0)
    ov = _ops.packed_override([f, b])
    for pa in (f, b):
        mr = ov[id(pa)]["mrange"].cpu().numpy()
        assert mr[0] == mr[1] == np.float32(pa.get("m")[0]).view(np.int32), pa.name
    ov16 = {k: dict(v, mrange=None) for k, v in ov.items()}

    def rho_of(dst, srcs, o):
        r = torch.full_like(dst.get_carray("rho"), -1.0)
        _ops.density(cm, prm, dst, srcs, r, override=o)
        torch.cuda.synchronize()
        return r.cpu().numpy()

    for dst, srcs in ((f, [f]), (f, [f, b]), (b, [f, b])):
        a, c = rho_of(dst, srcs, ov), rho_of(dst, srcs, ov16)
        assert np.array_equal(a, c), (dst.name, len(srcs))
    outs = []
    for o in (ov, ov16):
        res = [torch.full_like(pa.get_carray("rho"), -1.0) for pa in (f, b)]
        _ops.density_multi(cm, prm, [f, b], [f, b], res, [None, None], [None, None], override=o)
        torch.cuda.synchronize()
        outs.append([t.cpu().numpy() for t in res])
    for a, c in zip(*outs):
        assert np.array_equal(a, c)
    # masses that differ: the pair is unequal, the 16-byte records are read
    n = f.get_number_of_particles()
    rng = np.random.Generator(np.random.PCG64(5))
    f.set(m=f.get("m") * (1.0 + 0.2 * rng.random(n)))
    ov2 = _ops.packed_override([f, b])
    mr = ov2[id(f)]["mrange"].cpu().numpy()
    assert mr[0] != mr[1]
    got = np.empty(n)
    got[f._order_np()] = rho_of(f, [f, b], ov2)
    rf, rb = to_oracle(f), to_oracle(b)
    assert max_rel(got, O.sphrho(3, rf, [rf, rb])) < 1e-5


def test_density_lds_rows_same_bits():
    """The opt-in LDS-staged stencil rows (sph_params.lds_rows: the wave
    loads the union of its lanes' candidate ranges into LDS, every lane
    screens its own range from there) give the same bits as the per-lane
    loads: one and two sources, uniform and varied masses, the fused launch,
    and RK2 steps of the bench's plan."""
    f, b, info = gen.dam_break_3d(dx=1.0 / 40, column=(10, 10, 20), tank=(24, 10, 20), seed=7,
                                  precision="single")
    particles = base.Particles([f, b])
    cm = particles.cell_manager
    cm.set_cells_per_radius(2)
    cm.update()

    def run(lds, massvar=False):
        prm = _ops.make_params(3, co=info["co"], ro=info["rho0"], gamma=7.0, lds_rows=lds)
        ov = _ops.packed_override([f, b])
        out = []
        for dst, srcs in ((f, [f]), (f, [f, b]), (b, [f, b])):
            r = torch.full_like(dst.get_carray("rho"), -1.0)
            _ops.density(cm, prm, dst, srcs, r, override=ov)
            out.append(r)
        res = [torch.full_like(pa.get_carray("rho"), -1.0) for pa in (f, b)]
        _ops.density_multi(cm, prm, [f, b], [f, b], res, [None, None], [None, None], override=ov)
        torch.cuda.synchronize()
        return [t.cpu().numpy() for t in out + res]

    for a, c in zip(run(False), run(True)):
        assert np.array_equal(a, c)
    n = f.get_number_of_particles()
    rng = np.random.Generator(np.random.PCG64(9))
    f.set(m=f.get("m") * (1.0 + 0.2 * rng.random(n)))
    for a, c in zip(run(False), run(True)):
        assert np.array_equal(a, c)
    import bench
    outs = []
    for lds in (False, True):
        s, parts, fl, bd, inf, col, tank = bench.build_case(scale=0.2)
        s.integrator.plan.set_lds_rows(lds)
        for _ in range(3):
            s.integrator.integrate(inf["dt"])
        torch.cuda.synchronize()
        outs.append({c: fl.get(c) for c in ("x", "y", "z", "u", "v", "w", "rho", "p")})
    for c in outs[0]:
        assert np.array_equal(outs[0][c], outs[1][c]), c


def test_density_multi_one_launch_equals_two():
    """sph_density_multi (the fluid's and the boundary's SPHRho + Tait in one
    launch) writes the same bits as two sph_density calls, and matches the
    oracle; the fused WCSPH plan takes it."""
    f, b, info = gen.dam_break_3d(dx=1.0 / 40, column=(10, 10, 20), tank=(24, 10, 20), seed=7,
                                  precision="single")
    particles = base.Particles([f, b])
    cm = particles.cell_manager
    cm.set_cells_per_radius(2)
    cm.update()
    prm = _ops.make_params(3, co=info["co"], ro=info["rho0"], gamma=7.0)
    ov = _ops.packed_override([f, b])
    outs = []
    for fused in (False, True):
        res = {pa.name: [torch.full_like(pa.get_carray("rho"), -1.0) for _ in range(3)]
               for pa in (f, b)}
        if fused:
            _ops.density_multi(cm, prm, [f, b], [f, b], [res[pa.name][0] for pa in (f, b)],
                               [res[pa.name][1] for pa in (f, b)],
                               [res[pa.name][2] for pa in (f, b)], override=ov)
        else:
            for pa in (f, b):
                _ops.density(cm, prm, pa, [f, b], *res[pa.name], override=ov)
        torch.cuda.synchronize()
        outs.append({k: [t.cpu().numpy() for t in v] for k, v in res.items()})
    for k in outs[0]:
        for a, c in zip(outs[0][k], outs[1][k]):
            assert np.array_equal(a, c), k
    order = f._order_np()
    rho_f = np.empty(f.get_number_of_particles())
    rho_f[order] = outs[1]["fluid"][0]
    rf, rb = to_oracle(f), to_oracle(b)
    assert max_rel(rho_f, O.sphrho(3, rf, [rf, rb])) < 1e-5
    s = solver.wcsph_solver(3, co=info["co"], ro=info["rho0"], alpha=0.1, gravity=info["g"],
                            xsph_eps=0.5)
    s.setup_integrator(particles)
    plan = s.integrator.plan
    assert plan._fuse_density()
