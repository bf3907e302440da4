"""SPH smoothing kernels (reference: base/kernels.pyx).

The kernel object is a descriptor: the neighbour passes evaluate the cubic
spline on the GPU (csrc/sph_common.h, w_shape / grad_shape).  The scalar
`py_function` / `py_gradient` methods exist for scripts and tests that
evaluate a single pair, as in the reference; no compute path calls them.
"""
import math

from .point import Point

CUBIC_SPLINE = 1  # kernels.pyx:37


def h_dim(h, dim):
    """1/h^dim (kernels.pyx:47-53)."""
    if dim == 1:
        return 1 / h
    if dim == 2:
        return 1 / (h * h)
    return 1 / (h * h * h)


class KernelBase:
    def __init__(self, dim=1, constant_h=-1):
        self.dim = dim
        self.constant_h = constant_h
        self.fac = self._fac(1.0)

    def radius(self):
        raise NotImplementedError("KernelBase::radius")

    def dimension(self):
        return self.dim

    def _fac(self, h):
        raise NotImplementedError("KernelBase::_fac")

    def py_fac(self, h):
        return self._fac(h)


class DummyKernel(KernelBase):
    """Kernel placeholder of the position-stepping operation
    (kernels.pyx:328-340)."""

    def _fac(self, h):
        return 1.0

    def radius(self):
        return 2.0

    def py_function(self, pa, pb, h):
        return 1.0

    def py_gradient(self, pa, pb, h, grad):
        grad.set(0.0, 0.0, 0.0)


class CubicSplineKernel(KernelBase):
    """Monaghan (1992) cubic spline, support radius 2h
    (kernels.pyx:345-449)."""

    def _fac(self, h):
        if self.dim == 1:
            return (2.0 / 3.0) / h
        if self.dim == 2:
            return 10 / (7 * math.pi) / (h * h)
        if self.dim == 3:
            return 1.0 / math.pi / (h * h * h)
        raise ValueError("dimension must be 1, 2 or 3")

    def radius(self):
        return 2

    def get_type(self):
        return CUBIC_SPLINE

    def py_function(self, pa, pb, h):
        rab = math.sqrt((pa.x - pb.x) * (pa.x - pb.x) + (pa.y - pb.y) * (pa.y - pb.y) +
                        (pa.z - pb.z) * (pa.z - pb.z))
        q = rab / h
        fac = h_dim(h, self.dim) * self.fac
        if q > 2.0:
            val = 0.0
        elif q > 1.0:
            val = 0.25 * (2 - q) * (2 - q) * (2 - q)
        else:
            val = 1 - 1.5 * (q * q) * (1 - 0.5 * q)
        return val * fac

    def py_gradient(self, pa, pb, h, grad=None):
        rab = math.sqrt((pa.x - pb.x) * (pa.x - pb.x) + (pa.y - pb.y) * (pa.y - pb.y) +
                        (pa.z - pb.z) * (pa.z - pb.z))
        q = rab / h
        fac = h_dim(h, self.dim) * self.fac
        val = 0.0
        if q > 2.0:
            pass
        elif q >= 1.0:
            val = -0.75 * (2 - q) * (2 - q) / (h * rab)
        elif q > 1e-14:
            val = 3.0 * (0.75 * q - 1) / (h * h)
        out = Point((pa.x - pb.x) * (val * fac), (pa.y - pb.y) * (val * fac),
                    (pa.z - pb.z) * (val * fac))
        if grad is not None:
            grad.set(out.x, out.y, out.z)
        return out
