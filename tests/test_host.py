"""Host-side logic on the CPU: the reference's operation grouping, solver
bookkeeping, calc checks and integrator property names, the fused-plan
matcher, ParticleArray user-order semantics, and the C-ABI library
(loads, exports every declared symbol, validates arguments before any GPU
call)."""
import ctypes
import os
import re

import numpy as np
import pytest

import pysph_amd.base.api as base
import pysph_amd.sph.api as sph
import pysph_amd.solver.api as solver
from pysph_amd import _abi

Fluid = base.ParticleType.Fluid
Solid = base.ParticleType.Solid
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _two_arrays():
    x = np.linspace(0, 1, 11)
    h = np.ones_like(x) * 2 * (x[1] - x[0])
    f = base.get_particle_array(name="fluid", type=Fluid, x=x, h=h)
    s = base.get_particle_array(name="solid", type=Solid, x=x, h=h)
    return f, s, base.Particles(arrays=[f, s])


def test_operation_grouping_reference_expectations():
    """solver/tests/test_sph_equation.py:54-178."""
    f, s, particles = _two_arrays()
    op = solver.SPHOperation(sph.SPHRho, on_types=[Fluid], updates=["rho"], id="sd",
                             from_types=[Fluid, Solid])
    data = op.get_calc_data(particles)
    assert len(data) == 1 and data[0]["dnum"] == 0
    assert data[0]["sources"] == [f, s]
    funcs = data[0]["funcs"]
    assert funcs[0].dest is f and funcs[0].source is f
    assert funcs[1].dest is f and funcs[1].source is s
    grav = solver.SPHIntegration(sph.GravityForce.withargs(gy=-9.81), on_types=[Solid, Fluid],
                                 updates=["u", "v"], id="gravity")
    calcs = grav.get_calcs(particles=particles, kernel=base.CubicSplineKernel())
    assert [c.dnum for c in calcs] == [0, 1]
    assert calcs[0].dest is f and calcs[1].dest is s
    for c in calcs:
        assert len(c.sources) == 0 and c.integrates and not c.nbr_info
        assert c.updates == ["u", "v"]


def test_calc_checks_num_outputs():
    """A 2D momentum with the default dim=3 raises, as the reference does
    (sph_calc.pyx:174-178)."""
    f, s, particles = _two_arrays()
    op = solver.SPHIntegration(sph.MomentumEquation.withargs(), on_types=[Fluid],
                               from_types=[Fluid], updates=["u", "v"], id="mom")
    with pytest.raises(ValueError):
        op.get_calcs(particles, base.CubicSplineKernel(dim=2))
    op = solver.SPHIntegration(sph.MomentumEquation.withargs(dim=2), on_types=[Fluid],
                               from_types=[Fluid], updates=["u", "v"], id="mom")
    assert len(op.get_calcs(particles, base.CubicSplineKernel(dim=2))) == 1


def test_solver_operation_order():
    """Solver.add_operation / replace / remove (solver.py:184-300)."""
    s = solver.Solver(dim=2, integrator_type=solver.RK2Integrator)
    mk = lambda i: solver.SPHOperation(sph.SPHRho, on_types=[Fluid], from_types=[Fluid],  # noqa
                                       updates=["rho"], id=i)
    s.add_operation(mk("a"))
    s.add_operation(mk("c"))
    s.add_operation(mk("b"), before=True, id="c")
    s.add_operation(mk("d"), id="a")
    assert s.order == ["a", "d", "b", "c"]
    s.replace_operation("d", mk("e"))
    assert s.order == ["a", "e", "b", "c"]
    s.remove_operation("b")
    assert s.order == ["a", "e", "c"]
    with pytest.raises(AssertionError):
        s.add_operation(mk("a"))


def test_wcsph_setup_calcs_and_integrator_props():
    f, b, info = solver.generators.dam_break_2d(width=0.05, height=0.05, dx=0.01)
    particles = base.Particles([f, b])
    s = solver.wcsph_solver(2, co=info["co"], ro=1000.0, alpha=0.1, gravity=info["g"],
                            xsph_eps=0.5)
    s.setup_integrator(particles)
    ids = [c.id for c in s.integrator.calcs]
    assert ids == ["density_fluid", "density_boundary", "eos_fluid", "eos_boundary",
                   "mom_fluid", "gravity_fluid", "step_fluid", "xsph_fluid"]
    # _p_0 shared per property, _k{n}_{p}{i}{j} per calc (integrator.py:229-337)
    names = set(f.get_property_names())
    for nm in ("_u_0", "_v_0", "_x_0", "_y_0", "_k1_u40", "_k2_v41", "_k1_x60", "_k2_y71"):
        assert nm in names, nm
    assert "_rho_0" not in names
    assert s.integrator.plan is not None
    assert s.integrator.plan.fluid is f
    assert [p.name for p in s.integrator.plan.srcs] == ["fluid", "boundary"]
    assert list(s.integrator.plan.flags) == [_abi.SPH_SRC_MOMENTUM | _abi.SPH_SRC_XSPH,
                                             _abi.SPH_SRC_MOMENTUM]


def test_plan_not_matched_for_other_orders():
    f, b, info = solver.generators.dam_break_2d(width=0.05, height=0.05, dx=0.01)
    particles = base.Particles([f, b])
    s = solver.wcsph_solver(2, co=info["co"], ro=1000.0, alpha=0.1)
    s.set_order(["eos", "density", "mom", "step"])
    s.setup_integrator(particles)
    assert s.integrator.plan is None  # eos before density: the generic path


def test_particle_array_user_order_api():
    pa = base.get_particle_array(x=np.array([3.0, 1.0, 2.0]), name="a")
    assert pa.get_number_of_particles() == 3
    assert list(pa.get("idx")) == [0, 1, 2]
    assert list(pa.x) == [3.0, 1.0, 2.0]
    pa.set(u=np.array([1.0, 2.0, 3.0]))
    assert pa.is_dirty
    pa.add_property({"name": "t", "default": 4.0})
    assert list(pa.get("t")) == [4.0] * 3
    pa.remove_particles([1])
    assert list(pa.x) == [3.0, 2.0] and list(pa.u) == [1.0, 3.0]
    sub = pa.extract_particles([1])
    assert list(sub.x) == [2.0]
    pa.append_parray(sub)
    assert list(pa.x) == [3.0, 2.0, 2.0]
    assert pa.dtype.itemsize == 8
    assert base.get_particle_array(cl_precision="single", x=[1.0]).dtype.itemsize == 4


def test_cell_size_rules():
    """compute_cell_size (cell.pyx:893-925)."""
    f, s, particles = _two_arrays()
    cm = particles.cell_manager
    assert cm.cell_size == pytest.approx(2 * 0.2)
    cm2 = base.CellManager(arrays_to_bin=[f], min_cell_size=1.0, max_cell_size=0.5)
    assert cm2.cell_size == 0.5
    assert base.py_real_to_int(-2.01, 1.0) == -3
    assert base.find_cell_id(base.Point(-2.01, -1.01, 1), 1.0) == base.IntPoint(-3, -2, 1)
    nb = base.construct_immediate_neighbor_list(base.IntPoint(0, 0, 0))
    assert len(nb) == 27 and nb[0] == base.IntPoint(-1, -1, -1) and nb[1] == base.IntPoint(-1, -1, 0)


def _header_functions():
    src = open(os.path.join(ROOT, "include", "sph_mi355x.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(sph_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = _abi.load()
    declared = _header_functions()
    assert len(declared) >= 15
    for name in declared:
        assert hasattr(L, name), name
        assert name in _abi.EXPORTS, name
    assert L.sph_abi_version() == 1


def test_abi_validates_before_touching_the_gpu():
    L = _abi.load()
    prm = _abi.SphParams()
    prm.dim = 4
    v = _abi.SphView()
    g = _abi.SphGrid()
    g.cell_size = 1.0
    for d in range(3):
        g.dims[d] = 1
    rc = L.sph_density(None, ctypes.byref(g), ctypes.byref(prm), ctypes.byref(v), 0, None,
                       1.0, None, None, None, None, None)
    assert rc == _abi.SPH_EUNSUPPORTED
    assert b"dimension" in L.sph_last_error()
    prm.dim = 3
    prm.radius_scale = 2.0
    g.cell_size = -1.0
    rc = L.sph_density(None, ctypes.byref(g), ctypes.byref(prm), ctypes.byref(v), 0, None,
                       1.0, None, None, None, None, None)
    assert rc == _abi.SPH_EINVAL
    assert L.sph_bin_workspace_bytes(1000, 5000) > 4 * (5000 + 2000)
    rc = L.sph_update(None, 7, 10, 0, 1.0, 0, None, None)
    assert rc == _abi.SPH_EUNSUPPORTED


def test_compute_without_gpu_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    f, s, particles = _two_arrays()
    from pysph_amd._device import GPURequired
    with pytest.raises(GPURequired):
        particles.cell_manager.update()
