"""The SPH operator interface (reference: sph/sph_func.pyx).

A function class describes one SPH operator between a source and a
destination array.  In the reference each instance loops over destinations
and neighbours in Cython (SPHFunction.eval, sph_func.pyx:199-234;
SPHFunctionParticle.eval_single, :381-394).  Here each class provides
`calc_eval(calc, outputs)`, which evaluates the operator for ALL sources of
an SPHCalc in one GPU launch and writes the outputs; that equals the
reference's zero-then-accumulate over sources (sph_calc.pyx:223-276).
`eval(kernel, out1, out2, out3)` keeps the reference's per-source
accumulate-into-outputs contract on top of it.
"""
from .. import _abi
from ..base.particle_types import LocalReal  # noqa: F401
from . import _ops


class Function:
    """A function class with bound parameters (sph_func.pyx:25-77)."""

    def __init__(self, sph_func, *args, **kwargs):
        self.sph_func = sph_func
        self.args = args
        self.kwargs = kwargs

    def get_func_class(self):
        return self.sph_func

    def get_func(self, source, dest):
        return self.sph_func(source, dest, *self.args, **self.kwargs)


class SPHFunction:
    """Base of every operator (sph_func.pyx:80-243).  `dest` is ignored and
    equals `source`, as in the reference (:104-108)."""

    def __init__(self, source, dest=None, setup_arrays=True, **kwargs):
        self.name = ""
        self.id = ""
        self.tag = ""
        self.source = source
        self.dest = source
        self.x, self.y, self.z = "x", "y", "z"
        self.u, self.v, self.w = "u", "v", "w"
        self.m, self.rho, self.h, self.p, self.e, self.cs = "m", "rho", "h", "p", "e", "cs"
        self.num_outputs = 3
        self.src_reads = []
        self.dst_reads = []
        self.kernel = None
        self.nbr_locator = None
        self.hks = False
        self.set_src_dst_reads()

    @classmethod
    def withargs(cls, *args, **kwargs):
        return Function(cls, *args, **kwargs)

    @classmethod
    def get_func(cls, source, dest):
        return Function(cls).get_func(source, dest)

    @classmethod
    def get_func_class(cls):
        return cls

    def set_src_dst_reads(self):
        pass

    def setup_arrays(self):
        pass

    def setup_iter_data(self):
        pass

    # ---------------------------------------------------------------- GPU
    @classmethod
    def calc_eval(cls, calc, outputs):
        raise NotImplementedError("%s has no GPU implementation" % cls.__name__)

    def _cell_manager(self):
        if self.nbr_locator is not None and self.nbr_locator.cell_manager is not None:
            return self.nbr_locator.cell_manager
        raise RuntimeError("%s: no neighbour locator set" % type(self).__name__)

    def eval(self, kernel, output1=None, output2=None, output3=None):
        """Accumulate this function's contribution from its one source into
        the outputs (device tensors of the destination, slot order) of
        LocalReal destinations, and set the outputs of the other destinations
        to 0 (sph_func.pyx:199-234)."""
        outs = [o for o in (output1, output2, output3) if o is not None][:self.num_outputs]
        if self.kernel is None:
            self.kernel = kernel
        tmp = [o.new_empty(o.shape) for o in outs]
        calc = _SingleCalc(self, kernel)
        type(self).calc_eval(calc, tmp)
        n = self.dest.get_number_of_particles()
        _ops.update(self.dest.dtype, n, 0, 1.0, [(o, o, t, None) for o, t in zip(outs, tmp)],
                    tag=self.dest.get_carray("tag"))


class _SingleCalc:
    """The minimal calc a single function's eval() runs through."""

    def __init__(self, func, kernel):
        self.funcs = [func]
        self.sources = [func.source]
        self.dest = func.dest
        self.kernel = kernel
        self.dim = kernel.dim
        loc = func.nbr_locator
        self.cell_manager = loc.cell_manager if loc is not None else None
        # the pair sets follow the function's locator
        self.symmetric = bool(getattr(loc, "symmetric", False))


class SPHFunctionParticle(SPHFunction):
    """Operators that sum over neighbours (sph_func.pyx:252-403).  hks
    kernel symmetrisation is not provided.

    exclude_self drops the self pair when the source is the destination, the
    intent of sph_func.pyx:386-389.  That reference code tests `self.src`,
    an attribute SPHFunctionParticle does not declare: the AttributeError is
    swallowed by the `cdef void eval_single` before `result` is zeroed, so
    the reference adds an unset result and no reference output exists for
    this option.  This class implements the intended semantics on purpose
    (pinned by the closed-form self term m_a W(0, h_a), not by the
    reference).  Of the functions here only the summation density changes,
    since the self pair adds 0 to the gradient and velocity-difference
    sums."""

    def __init__(self, source, dest, setup_arrays=True, hks=False, exclude_self=False,
                 nbr_locator=None, **kwargs):
        SPHFunction.__init__(self, source, setup_arrays=False, **kwargs)
        if hks:
            raise NotImplementedError("hks kernel symmetrisation is not supported")
        self.dest = dest
        self.exclude_self = exclude_self
        self.bonnet_and_lok_correction = False
        self.rkpm_first_order_correction = False
        self.hks = hks
        self.nbr_locator = nbr_locator


class CSPHFunctionParticle(SPHFunctionParticle):
    """Corrected-SPH variant (sph_func.pyx:405-493): the denominator stays 0
    without kernel correction, so the result is the plain sum."""


def grid_of(calc):
    cm = calc.cell_manager
    cm.update()
    return cm


SRC_MOMENTUM = _abi.SPH_SRC_MOMENTUM
SRC_XSPH = _abi.SPH_SRC_XSPH
