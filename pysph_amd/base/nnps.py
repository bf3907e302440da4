"""Neighbour locators (reference: base/nnps.pyx).

FixedDestNbrParticleLocator answers `get_nearest_particles(i)` with the
source indices j for which
    ((xj-xi)^2 + (yj-yi)^2) + (zj-zi)^2 < (h_i * radius_scale)^2
evaluated in FP64 exactly as the reference does (nnps.pyx:311-354, :552,
point.pxd:35-37), self included and moved to the last slot when source is
destination (nnps.pyx:612-618).  The lists are computed on the GPU for all
destinations at once (csrc/sph_bin.hip, nbr_kernel) and cached until either
array moves, as the reference's particle cache is (nnps.pyx:561-620).

The hot SPH passes never materialise these lists: they walk the same cell
stencil with the same exact test.  The lists exist for the reference's query
API, for parity checks and for the neighbour-graph consumers.
"""
import numpy as np
import torch

from .. import _abi
from .._device import require_gpu
from .carray import LongArray
from .particle_array import ParticleArray


class NeighborLocatorType:
    PointNeighborLocator = 0
    SPHNeighborLocator = 1
    NSquareNeighborLocator = 2


def _csr_user_order(pa_dst, off_slot, nbr, n):
    """Reorder CSR rows from destination slot order to user order."""
    order = pa_dst._order_np()
    counts = np.diff(off_slot)
    if order is None:
        return off_slot, nbr
    ucounts = np.empty_like(counts)
    ucounts[order] = counts
    uoff = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(ucounts, out=uoff[1:])
    # entry e of slot row s moves to uoff[order[s]] + (e - off_slot[s])
    row = np.repeat(np.arange(n, dtype=np.int64), counts)
    dest = uoff[order[row]] + (np.arange(len(nbr), dtype=np.int64) - off_slot[row])
    out = np.empty_like(nbr)
    out[dest] = nbr
    return uoff, out


def _self_last(off, nb):
    """Move each row's own index to the row's last slot (nnps.pyx:612-618);
    rows hold an index at most once."""
    n = len(off) - 1
    row = np.repeat(np.arange(n, dtype=np.int64), np.diff(off))
    pos = np.nonzero(nb == row)[0]
    last = off[row[pos] + 1] - 1
    nb[pos] = nb[last]
    nb[last] = row[pos]
    return nb


def neighbor_csr(cell_manager, src, dst, radius_scale, symmetric=False):
    """All neighbour lists of dst from src on the GPU: (offsets, nbrs) in user
    index space, rows in user order, each row sorted ascending.  With
    `symmetric` the rows are the variable-h sets of VarHNbrParticleLocator
    (nnps.pyx:654-749): d2 < (h_dst rs)^2 or d2 < (h_src rs)^2."""
    cm = cell_manager
    cm.update()
    n = dst.get_number_of_particles()
    if n == 0:
        return np.zeros(1, dtype=np.int64), np.zeros(0, dtype=np.int64)
    stream = require_gpu(dst.get_carray("x"))
    dev = dst.device
    counts = torch.empty(n, dtype=torch.int64, device=dev)
    vd = dst.view()
    vs = src.view()
    if symmetric:
        rmax = float(radius_scale) * cm.max_h
        _abi.call("sph_varh_neighbor_counts", stream, cm.grid, vd, vs, float(radius_scale),
                  rmax, counts.data_ptr())
    else:
        _abi.call("sph_neighbor_counts", stream, cm.grid, vd, vs, float(radius_scale),
                  counts.data_ptr())
    offsets = torch.empty(n + 1, dtype=torch.int64, device=dev)
    total = _abi.ctypes.c_int64(0)
    _abi.call("sph_scan_counts", stream, n, counts.data_ptr(), offsets.data_ptr(),
              _abi.ctypes.byref(total))
    total = int(total.value)
    nbrs = torch.empty(max(total, 1), dtype=torch.int64, device=dev)
    ids = src._order if src._order is not None else torch.arange(
        src.get_number_of_particles(), dtype=torch.int64, device=dev)
    needed = _abi.ctypes.c_int64(0)
    if symmetric:
        _abi.call("sph_varh_neighbor_fill", stream, cm.grid, vd, vs, float(radius_scale),
                  rmax, ids.data_ptr(), offsets.data_ptr(), nbrs.data_ptr(), total, total,
                  _abi.ctypes.byref(needed))
    else:
        _abi.call("sph_neighbor_fill", stream, cm.grid, vd, vs, float(radius_scale),
                  ids.data_ptr(), offsets.data_ptr(), nbrs.data_ptr(), total, total,
                  _abi.ctypes.byref(needed))
    off = offsets.cpu().numpy()
    nb = nbrs[:total].cpu().numpy()
    return _csr_user_order(dst, off, nb, n)


class NbrParticleLocatorBase:
    """Point queries against one source array (nnps.pyx:100-356)."""

    def __init__(self, source, cell_manager=None, locator_type=NeighborLocatorType.SPHNeighborLocator):
        self.source = source
        self.cell_manager = cell_manager
        self.locator_type = locator_type
        self.source_index = cell_manager.arrays_to_bin.index(source) if cell_manager else -1

    def py_get_nearest_particles_to_point(self, pnt, radius, output_array,
                                          exclude_index=-1):
        """Append the source indices within `radius` of pnt (strict <)."""
        if self.locator_type == NeighborLocatorType.NSquareNeighborLocator:
            n = self.source.get_number_of_particles()
            ids = [i for i in range(n) if i != exclude_index]
            for i in ids:
                output_array.append(i)
            return 0
        src = self.source
        cm = self.cell_manager
        cm.update()
        q = ParticleArray(name="__query__", cl_precision=src.cl_precision,
                          x={"data": [pnt.x]}, y={"data": [pnt.y]}, z={"data": [pnt.z]},
                          h={"data": [radius]})
        # the query point is binned on the manager's grid
        cm.bin_array(q, cm.grid)
        off, nb = neighbor_csr(cm, src, q, 1.0)
        for j in nb:
            if j != exclude_index:
                output_array.append(int(j))
        return 0


class FixedDestNbrParticleLocator(NbrParticleLocatorBase):
    """Cached neighbour lists of a fixed destination array
    (nnps.pyx:385-649)."""

    def __init__(self, source, dest, radius_scale, cell_manager=None, h="h",
                 locator_type=NeighborLocatorType.SPHNeighborLocator):
        NbrParticleLocatorBase.__init__(self, source, cell_manager, locator_type)
        self.dest = dest
        self.radius_scale = radius_scale
        self.h = h
        self.is_dirty = True
        self._off = None
        self._nbr = None

    def update_status(self):
        if not self.is_dirty:
            self.is_dirty = self.source.is_dirty or self.dest.is_dirty

    py_update_status = update_status

    def update(self):
        if self.cell_manager is not None:
            self.cell_manager.update_status()
            if self.cell_manager.is_dirty:
                self.is_dirty = True
        if self.is_dirty or self._off is None:
            off, nb = neighbor_csr(self.cell_manager, self.source, self.dest,
                                   self.radius_scale, symmetric=self.symmetric)
            if self.source is self.dest:
                # keep the particle itself in the last slot (nnps.pyx:612-618)
                nb = _self_last(off, nb)
            self._off, self._nbr = off, nb
            self.is_dirty = False
        return 0

    def py_update(self):
        return self.update()

    def get_csr(self):
        """(offsets, neighbour indices) of every destination, user order."""
        self.update()
        return self._off, self._nbr

    def py_get_nearest_particles(self, dest_p_index, exclude_self=False):
        if self.locator_type == NeighborLocatorType.NSquareNeighborLocator:
            n = self.source.get_number_of_particles()
            ids = np.arange(n, dtype=np.int64)
            if exclude_self and self.source is self.dest:
                ids = ids[ids != dest_p_index]
            return LongArray(ids)
        self.update()
        row = self._nbr[self._off[dest_p_index]:self._off[dest_p_index + 1]].copy()
        if exclude_self and self.source is self.dest:
            row = row[:-1]
        return LongArray(row)

    get_nearest_particles = py_get_nearest_particles

    symmetric = False


class VarHNbrParticleLocator(FixedDestNbrParticleLocator):
    """Symmetric neighbour lists for variable smoothing lengths
    (nnps.pyx:654-749): j is a neighbour of i when i's search finds j or j's
    search (with h_j) finds i.  The reference builds the union by appending
    the reverse locator's hits; here one GPU pass tests both radii per pair.
    Rows hold the same sets, sorted, with the particle itself last when
    source is destination (the reference leaves it where its forward search
    put it)."""

    symmetric = True


class NNPSManager:
    """Hands out cached locators keyed by (src, dst, radius_scale)
    (nnps.pyx:754-895)."""

    def __init__(self, cell_manager=None, variable_h=False,
                 locator_type=NeighborLocatorType.SPHNeighborLocator):
        self.cell_manager = cell_manager
        self.variable_h = variable_h
        self.locator_type = locator_type
        self.particle_locator_cache = {}
        self.polygon_cache = {}

    def get_neighbor_particle_locator(self, source, dest=None, radius_scale=1.0):
        """A base locator without dest, else the cached locator of
        (source, dest, radius_scale) (nnps.pyx:789-807)."""
        if dest is None:
            return NbrParticleLocatorBase(source, self.cell_manager, self.locator_type)
        key = (id(source), id(dest), radius_scale)
        loc = self.particle_locator_cache.get(key)
        if loc is None:
            cls = VarHNbrParticleLocator if self.variable_h else FixedDestNbrParticleLocator
            loc = cls(source, dest, radius_scale, self.cell_manager,
                      locator_type=self.locator_type)
            self.particle_locator_cache[key] = loc
        return loc

    def add_interaction(self, source, dest, radius_scale):
        self.get_neighbor_particle_locator(source, dest, radius_scale)

    def update(self):
        for loc in self.particle_locator_cache.values():
            loc.update_status()
        if self.cell_manager is not None:
            self.cell_manager.update_status()
        return 0

    def py_update(self):
        return self.update()
