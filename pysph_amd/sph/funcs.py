"""The WCSPH operators of the hot path.

  SPHRho            sph/funcs/density_funcs.pyx:15-84
  SPHDensityRate    sph/funcs/density_funcs.pyx:86-158
  TaitEquation      sph/funcs/eos_funcs.pyx:58-116
  MomentumEquation  sph/funcs/pressure_funcs.pyx:106-247
  XSPHCorrection    sph/funcs/xsph_funcs.pyx:8-88
  PositionStepping  sph/funcs/position_funcs.pyx:10-60
  GravityForce      sph/funcs/external_force.pyx:14-63
  MoveCircleX/Y     sph/funcs/external_force.pyx:99-162

Constructor signatures, ids, tags and num_outputs follow the reference so
that operation set-up code written for it works unchanged.
"""
from .. import _abi
from . import _ops
from .sph_func import CSPHFunctionParticle, SPHFunction, SPHFunctionParticle, grid_of


class SPHRho(CSPHFunctionParticle):
    """Summation density rho_a = sum_b m_b W(r_a - r_b, h_ab)."""

    def __init__(self, source, dest, setup_arrays=True, **kwargs):
        CSPHFunctionParticle.__init__(self, source, dest, setup_arrays=True, **kwargs)
        self.id = "sphrho"
        self.tag = "density"
        self.num_outputs = 1

    def set_src_dst_reads(self):
        self.src_reads = ["x", "y", "z", "h", "m"]
        self.dst_reads = ["x", "y", "z", "h", "tag"]

    @classmethod
    def calc_eval(cls, calc, outputs):
        cm = grid_of(calc)
        # sources whose self pair is dropped: the intended semantics of
        # exclude_self (sph_func.pyx:381-394; see SPHFunctionParticle)
        excl = sum(1 << s for s, (f, src) in enumerate(zip(calc.funcs, calc.sources))
                   if getattr(f, "exclude_self", False) and src is calc.dest)
        prm = _ops.calc_params(calc, exclude_self=excl)
        _ops.density(cm, prm, calc.dest, calc.sources, outputs[0])


class SPHDensityRate(SPHFunctionParticle):
    """Continuity equation: drho_a/dt = sum_b m_b (v_a - v_b) . grad_a W_ab."""

    def __init__(self, source, dest, setup_arrays=True, **kwargs):
        SPHFunctionParticle.__init__(self, source, dest, setup_arrays=True, **kwargs)
        self.id = "densityrate"
        self.tag = "density"
        self.num_outputs = 1

    def set_src_dst_reads(self):
        self.src_reads = ["x", "y", "z", "h", "m", "u", "v", "w"]
        self.dst_reads = ["x", "y", "z", "h", "tag", "u", "v", "w"]

    @classmethod
    def calc_eval(cls, calc, outputs):
        cm = grid_of(calc)
        _ops.density_rate(cm, _ops.calc_params(calc), calc.dest, calc.sources, outputs[0])


class TaitEquation(SPHFunction):
    """p = B((rho/ro)^gamma - 1), cs = co (rho/ro)^((gamma-1)/2),
    B = co^2 ro / gamma."""

    def __init__(self, source, dest=None, setup_arrays=True, co=1.0, ro=1000.0, gamma=7.0,
                 **kwargs):
        kwargs.pop("hks", None)
        SPHFunction.__init__(self, source, dest, setup_arrays, **kwargs)
        self.co = co
        self.ro = ro
        self.gamma = gamma
        self.B = co * co * ro / gamma
        self.id = "tait"
        self.tag = "state"
        self.num_outputs = 2

    def set_src_dst_reads(self):
        self.src_reads = []
        self.dst_reads = ["rho"]

    def params(self, dim):
        return _ops.make_params(dim, co=self.co, ro=self.ro, gamma=self.gamma)

    @classmethod
    def calc_eval(cls, calc, outputs):
        f = calc.funcs[0]
        _ops.tait(f.params(calc.kernel.dim), calc.dest, outputs[0], outputs[1])


class MomentumEquation(SPHFunctionParticle):
    """Pressure gradient with Monaghan artificial viscosity:
    a_a = -sum_b m_b (P_a/rho_a^2 + P_b/rho_b^2 + Pi_ab) grad_a W_ab."""

    def __init__(self, source, dest, setup_arrays=True, alpha=1, beta=1, gamma=1.4, eta=0.1,
                 dim=3, **kwargs):
        SPHFunctionParticle.__init__(self, source, dest, setup_arrays, **kwargs)
        self.alpha = alpha
        self.beta = beta
        self.gamma = gamma
        self.eta = eta
        self.id = "momentumequation"
        self.tag = "velocity"
        self.num_outputs = dim

    def set_src_dst_reads(self):
        self.src_reads = ["x", "y", "z", "h", "m", "rho", "p", "u", "v", "w", "cs"]
        self.dst_reads = ["x", "y", "z", "h", "rho", "p", "u", "v", "w", "cs", "tag"]

    @classmethod
    def calc_eval(cls, calc, outputs):
        f = calc.funcs[0]
        cm = grid_of(calc)
        prm = _ops.calc_params(calc, alpha=f.alpha, beta=f.beta, eta=f.eta)
        acc = list(outputs) + [None] * (3 - len(outputs))
        _ops.momentum_xsph(cm, prm, calc.dest, calc.sources,
                           [_abi.SPH_SRC_MOMENTUM] * len(calc.sources), acc=acc)


class XSPHCorrection(CSPHFunctionParticle):
    """dv_a = eps sum_b m_b W_ab / rho_ab (v_b - v_a)."""

    def __init__(self, source, dest, setup_arrays=True, eps=0.5, dim=3, **kwargs):
        CSPHFunctionParticle.__init__(self, source, dest, setup_arrays, **kwargs)
        self.eps = eps
        self.id = "xsph"
        self.tag = "position"
        self.num_outputs = dim

    def set_src_dst_reads(self):
        self.src_reads = ["x", "y", "z", "h", "m", "rho", "u", "v", "w"]
        self.dst_reads = ["x", "y", "z", "h", "rho", "u", "v", "w"]

    @classmethod
    def calc_eval(cls, calc, outputs):
        f = calc.funcs[0]
        cm = grid_of(calc)
        prm = _ops.calc_params(calc, eps=f.eps)
        dv = list(outputs) + [None] * (3 - len(outputs))
        _ops.momentum_xsph(cm, prm, calc.dest, calc.sources,
                           [_abi.SPH_SRC_XSPH] * len(calc.sources), dv=dv)


class PositionStepping(SPHFunction):
    """k_x = u (the velocity at evaluation time)."""

    def __init__(self, source, dest=None, setup_arrays=True, dim=3):
        SPHFunction.__init__(self, source, dest, setup_arrays)
        self.id = "positionstepper"
        self.tag = "position"
        self.num_outputs = dim

    def set_src_dst_reads(self):
        self.dst_reads = ["u", "v", "w"][:self.num_outputs]

    @classmethod
    def calc_eval(cls, calc, outputs):
        pa = calc.dest
        vel = [pa.get_carray(c) for c in ("u", "v", "w")]
        _ops.update(pa.dtype, pa.get_number_of_particles(), 3, 0.0,
                    [(o, v, None, None) for o, v in zip(outputs, vel)],
                    tag=pa.get_carray("tag"))


class GravityForce(SPHFunction):
    """k_v = (gx, gy, gz)."""

    def __init__(self, source, dest=None, setup_arrays=True, gx=0.0, gy=0.0, gz=0.0,
                 hks=False):
        SPHFunction.__init__(self, source, dest, setup_arrays)
        self.id = "gravityforce"
        self.tag = "velocity"
        self.gx, self.gy, self.gz = gx, gy, gz

    @classmethod
    def calc_eval(cls, calc, outputs):
        f = calc.funcs[0]
        pa = calc.dest
        g = (f.gx, f.gy, f.gz)
        for o, gc in zip(outputs, g):
            _ops.update(pa.dtype, pa.get_number_of_particles(), 4, gc, [(o, None, None, None)],
                        tag=pa.get_carray("tag"))


class _MoveCircle(SPHFunction):
    axis = 0

    def __init__(self, source, dest=None, setup_arrays=True, *args, **kwargs):
        SPHFunction.__init__(self, source, dest, setup_arrays)
        self.tag = "position"
        self.num_outputs = 2

    def set_src_dst_reads(self):
        self.src_reads = []
        self.dst_reads = ["x", "y", "z", "tag"]

    @classmethod
    def calc_eval(cls, calc, outputs):
        pa = calc.dest
        o = list(outputs) + [None] * (2 - len(outputs))
        stream = _ops.require_gpu(pa.get_carray("x"))
        _abi.call("sph_move_circle", stream, pa.view(), cls.axis, _ops.ptr(o[0]), _ops.ptr(o[1]))


class MoveCircleX(_MoveCircle):
    """(-sin(angle), 0) with angle = arccos(x/|p|), sign flipped for y < 0:
    the x velocity of unit-speed motion on a circle."""
    axis = 0

    def __init__(self, source, dest=None, setup_arrays=True, *args, **kwargs):
        _MoveCircle.__init__(self, source, dest, setup_arrays)
        self.id = "circlex"


class MoveCircleY(_MoveCircle):
    """(0, cos(angle)): the y velocity of unit-speed motion on a circle."""
    axis = 1

    def __init__(self, source, dest=None, setup_arrays=True, *args, **kwargs):
        _MoveCircle.__init__(self, source, dest, setup_arrays)
        self.id = "circley"
