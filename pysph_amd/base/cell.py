"""CellManager: GPU cell binning of all particle arrays of a simulation.

Reference: base/cell.pyx.  The reference keeps a Python dict of Cell
objects keyed by IntPoint(floor(x/cs)) and moves escaped particles between
cells incrementally (cell.pyx:319-432, :927-1015).  Here every array is
re-binned on the GPU by a stable counting sort on the linear cell key of a
dense grid (csrc/sph_bin.hip) and its properties are permuted into cell
order, so each stencil row of cells is one contiguous particle range.

The cell of a particle is the reference's floor(x/cell_size)
(cell.pyx:42-82); the grid offset `lo` is an integer, so the key is exact.
The cell size follows compute_cell_size (cell.pyx:893-925).
"""
import math

import numpy as np
import torch

from .. import _abi
from .._device import dtype_code, require_gpu
from .point import IntPoint


def real_to_int(real_val, step):
    """floor(real_val/step) as an int (cell.pyx:42-59)."""
    return int(math.floor(real_val / step))


# "no cell yet" in the jump check's per-particle reference cells
_REF_NONE = -(1 << 63)


def find_cell_id(pnt, cell_size):
    """Cell index of a point (cell.pyx:61-82)."""
    return IntPoint(real_to_int(pnt.x, cell_size), real_to_int(pnt.y, cell_size),
                    real_to_int(pnt.z, cell_size))


py_real_to_int = real_to_int
py_find_cell_id = find_cell_id


def construct_immediate_neighbor_list(cell_id, include_self=True, distance=1):
    """Ids of the (2d+1)^3 cells around cell_id, x outer, z inner
    (cell.pyx:93-116)."""
    out = []
    for i in range(-distance, distance + 1):
        for j in range(-distance, distance + 1):
            for k in range(-distance, distance + 1):
                out.append(IntPoint(cell_id.x + i, cell_id.y + j, cell_id.z + k))
    if not include_self:
        n = len(out)
        out[n // 2] = out[n - 1]
        out.pop()
    return out


def grid_for_domain(lo, hi, cell_size, slack=2):
    """Cell grid covering [lo, hi] with `slack` spare cells on each side.
    An axis on which the domain has no extent (z of a 2D set) gets a single
    layer of cells, so that 2D stencils do not visit empty planes."""
    cs = float(cell_size)
    g = _abi.SphGrid()
    g.cell_size = cs
    for d in range(3):
        s = 0 if lo[d] == hi[d] else slack
        c0 = math.floor(lo[d] / cs) - s
        c1 = math.floor(hi[d] / cs) + s
        g.lo[d] = int(max(min(c0, 2**30), -2**30))
        g.dims[d] = int(max(1, min(c1 - c0 + 1, 2**20)))
    ncells = g.dims[0] * g.dims[1] * g.dims[2]
    if ncells > 2**31 - 2:
        raise MemoryError("cell grid of %d cells is too large" % ncells)
    return g


class PeriodicDomain:
    """Limits of a periodic box (cell.pyx:1568-1596).  An axis whose limits
    stay at the default +-1000 has translate 0 and is not periodic."""

    def __init__(self, xmin=-1000, xmax=1000, ymin=-1000, ymax=1000, zmin=-1000, zmax=1000):
        self.xmin, self.xmax = float(xmin), float(xmax)
        self.ymin, self.ymax = float(ymin), float(ymax)
        self.zmin, self.zmax = float(zmin), float(zmax)
        self.xtranslate = 0.0 if (self.xmax - self.xmin) == 2000 else self.xmax - self.xmin
        self.ytranslate = 0.0 if (self.ymax - self.ymin) == 2000 else self.ymax - self.ymin
        self.ztranslate = 0.0 if (self.zmax - self.zmin) == 2000 else self.zmax - self.zmin

    def periods(self):
        """((period per axis, 0 = not periodic), (lower face per axis))."""
        return ((self.xtranslate, self.ytranslate, self.ztranslate),
                (self.xmin, self.ymin, self.zmin))


class CellManager:
    """Bins every array of `arrays_to_bin` on one shared grid.

    With a periodic_domain the stored coordinates are wrapped into the box
    before each binning (x >= max moves by -L, x < min by +L, the
    reference's binning-point rule of cell.pyx:392-398 applied to the
    coordinate, and to the integrator's `_x_0` copies by the same amount),
    the jump check is off as in the reference, and the neighbour passes see
    sources across each periodic face at x_j +- L, which is where the
    reference's ghost cells (cell.pyx:1277-1376) put their copies; no ghost
    particles are created."""

    def __init__(self, arrays_to_bin=(), min_cell_size=-1.0, max_cell_size=0,
                 periodic_domain=None, initialize=True, max_radius_scale=2.0):
        self.periodic_domain = periodic_domain
        self.arrays_to_bin = list(arrays_to_bin)
        self.min_cell_size = min_cell_size
        self.max_cell_size = max_cell_size
        self.max_radius_scale = max_radius_scale
        self.jump_tolerance = 1
        # relative margin of the cell size over the search radius; a binning
        # stays reusable while particles move less than skin*radius/2
        self.skin = 0.0
        # True: count on the device at every binning, read back without a
        # wait and raised at a later binning or in verify_jumps();
        # "deferred": checked in verify_jumps() only; False: no check
        self.check_jumps = True
        self.cell_size = 0.0
        # binning grid bins per reference cell (set_cells_per_radius)
        self.bins_per_cell = 1
        self.min_h = self.max_h = 0.0
        self.is_dirty = True
        self.initialized = False
        self.grid = None          # _abi.SphGrid of the last binning
        self.fixed_domain = None  # optional ((xmin,ymin,zmin),(xmax,ymax,zmax))
        self.grid_slack = 2       # extra cells on every side of the bounds
        self.coord_x, self.coord_y, self.coord_z = "x", "y", "z"
        self.num_binnings = 0
        self._ws = None
        self._jump_acc = None
        # asynchronous read-backs of the default path (no host waits per
        # binning): the jump count of check_jumps=True and the bounds that
        # size the grid, copied to pinned host memory behind an event
        self.async_bounds = True
        self._jump_read = None
        self._jump_seen = 0
        self._bounds_read = None
        self._h_versions = None
        # set while a step is captured into a HIP graph: binning keeps the
        # current grid and starts no host read-back (the graph's owner reads
        # bounds and jump counts outside the graph)
        self.capturing = False
        if initialize and self.arrays_to_bin:
            self.initialize()

    # ------------------------------------------------------------------
    def _compute_minmax_h(self):
        mn, mx = math.inf, 0.0
        for pa in self.arrays_to_bin:
            if pa.get_number_of_particles() == 0:
                continue
            h = pa.get_carray("h")
            mn = min(mn, float(h.min()))
            mx = max(mx, float(h.max()))
        return (0.0 if mn == math.inf else mn), mx

    def compute_cell_size(self, min_size=0, max_size=0):
        """cell_size = max_radius_scale * max h unless min_size > 0, capped by
        max_size, 1.0 when zero (cell.pyx:893-925)."""
        self.min_h, self.max_h = self._compute_minmax_h()
        if min_size <= 0:
            self.cell_size = self.max_radius_scale * self.max_h * (1.0 + self.skin)
        else:
            self.cell_size = min_size
        if max_size > 0 and self.cell_size > max_size:
            self.cell_size = max_size
        if self.cell_size == 0.0:
            self.cell_size = 1.0
        return self.cell_size

    def initialize(self):
        self.compute_cell_size(self.min_cell_size, self.max_cell_size)
        self.initialized = True
        self.is_dirty = True
        for pa in self.arrays_to_bin:
            pa.set_dirty(True)

    def set_skin(self, skin):
        """Cell size = max_radius_scale * max h * (1 + skin); re-bins."""
        self.skin = float(skin)
        if self.min_cell_size <= 0:
            self.initialize()
            self.grid = None

    def set_cells_per_radius(self, k):
        """Bin the particles on a grid k times finer than the reference's
        cells (k = 1: the bins are the cells).  cell_size, the cell ids it
        defines and the jump tolerance keep the reference's cell; smaller
        bins test fewer candidates per neighbour at the price of more
        stencil rows, and neighbour sets do not change."""
        k = int(k)
        if k < 1:
            raise ValueError("cells per radius must be >= 1, got %d" % k)
        self.bins_per_cell = k
        self.grid = None
        self.is_dirty = True
        for pa in self.arrays_to_bin:
            pa.set_dirty(True)
        return self.cell_size / k if self.cell_size else 0.0

    @property
    def bin_size(self):
        """Edge of one bin of the binning grid: cell_size / bins_per_cell."""
        return (self.cell_size or 1.0) / self.bins_per_cell

    def set_fixed_domain(self, lo, hi):
        """Bin on a fixed box (no bounds read-back): needed for graph capture.
        Particles outside it are clamped into border cells, which keeps the
        search complete."""
        self.fixed_domain = (tuple(lo), tuple(hi))
        self.grid = None
        self.is_dirty = True

    def update_status(self):
        for pa in self.arrays_to_bin:
            if pa.is_dirty:
                self.is_dirty = True

    # ------------------------------------------------------------------
    def _bounds(self):
        lo = np.full(3, np.inf)
        hi = np.full(3, -np.inf)
        hmax = 0.0
        for pa in self.arrays_to_bin:
            if pa.get_number_of_particles() == 0:
                continue
            stream = require_gpu(pa.get_carray("x"))
            out = torch.empty(_abi.SPH_BOUNDS_WORDS, dtype=torch.float64, device=pa.device)
            v = pa.view(("x", "y", "z", "h"))
            _abi.call("sph_bounds", stream, v, out.data_ptr())
            b = out[:7].cpu().numpy()
            lo = np.minimum(lo, b[:3])
            hi = np.maximum(hi, b[3:6])
            hmax = max(hmax, float(b[6]))
        if hmax > 0:
            self.max_h = hmax
        if not np.all(np.isfinite(lo)) or not np.all(np.isfinite(hi)):
            bad = ~(np.isfinite(lo) & np.isfinite(hi))
            lo[bad] = 0.0
            hi[bad] = 0.0
        return lo, hi

    def _h_state(self):
        return tuple(pa.__dict__.get("_h_version", 0) for pa in self.arrays_to_bin)

    def _launch_bounds(self):
        """Bounds of every array into one device buffer, copied to pinned
        host memory on the stream behind an event (no wait)."""
        arrs = [pa for pa in self.arrays_to_bin if pa.get_number_of_particles() > 0]
        if not arrs:
            return None
        dev = arrs[0].device
        stream = require_gpu(arrs[0].get_carray("x"))
        buf = torch.empty((len(arrs), _abi.SPH_BOUNDS_WORDS), dtype=torch.float64, device=dev)
        for k, pa in enumerate(arrs):
            _abi.call("sph_bounds", stream, pa.view(("x", "y", "z", "h")), buf[k].data_ptr())
        return _AsyncRead(buf)

    @staticmethod
    def _reduce_bounds(b):
        """lo[3], hi[3] and max h over rows of sph_bounds output."""
        lo = np.min(b[:, :3], axis=0)
        hi = np.max(b[:, 3:6], axis=0)
        bad = ~(np.isfinite(lo) & np.isfinite(hi))
        lo[bad] = 0.0
        hi[bad] = 0.0
        return lo, hi, float(np.max(b[:, 6]))

    def make_grid(self, lo, hi):
        g = grid_for_domain(lo, hi, self.bin_size, self.grid_slack)
        g.bins_per_cell = self.bins_per_cell
        self._set_periodic(g)
        return g

    def _set_periodic(self, g):
        if self.periodic_domain is None:
            return g
        per, pmin = self.periodic_domain.periods()
        for d in range(3):
            g.period[d] = float(per[d])
            g.pmin[d] = float(pmin[d])
        return g

    def _check_period(self):
        per, _ = self.periodic_domain.periods()
        r = self.max_radius_scale * self.max_h
        for d in range(3):
            if per[d] > 0 and not per[d] > 2.0 * r:
                raise ValueError("periodic axis %d: the period %g must exceed twice the "
                                 "search radius %g" % (d, per[d], r))

    def wrap_periodic(self, pa):
        """Wrap pa's coordinates (and its `_x_0`-style integrator copies)
        into the periodic box on the device."""
        if self.periodic_domain is None or pa.get_number_of_particles() == 0:
            return
        g = self._set_periodic(_abi.SphGrid())
        g.cell_size = self.cell_size or 1.0
        for d in range(3):
            g.dims[d] = 1
        pos = [pa.get_carray(c) for c in ("x", "y", "z")]
        extra = []
        if all(pa.has_array("_%s_0" % c) for c in ("x", "y", "z")):
            extra.append([pa.get_carray("_%s_0" % c) for c in ("x", "y", "z")])
        stream = require_gpu(pos[0])
        _abi.call("sph_periodic_wrap", stream, g, dtype_code(pa.dtype),
                  pa.get_number_of_particles(), _abi.ptr_array([t.data_ptr() for t in pos]),
                  len(extra), _abi.ptr_array([t.data_ptr() for e in extra for t in e])
                  if extra else None)

    @staticmethod
    def grid_key(g):
        return (g.cell_size, tuple(g.lo), tuple(g.dims))

    def ncells(self):
        g = self.grid
        return g.dims[0] * g.dims[1] * g.dims[2]

    def _workspace(self, n, ncells, device):
        need = _abi.load().sph_bin_workspace_bytes(n, ncells)
        if self._ws is None or self._ws.numel() < need or self._ws.device != device:
            self._ws = torch.empty(need, dtype=torch.uint8, device=device)
        return self._ws

    def update(self):
        """Re-bin dirty arrays (all arrays when the grid changes).

        The default path waits on the host for nothing.  The jump check of
        check_jumps=True counts on the device and its count is read back
        asynchronously: a violation raises the reference's JumpToleranceError
        (cell.pyx:1422-1437) at the first binning after the count reached the
        host, and verify_jumps() (called at the end of Solver.solve) waits
        for it.  The bounds that size the grid are read asynchronously too:
        the grid follows the latest bounds that have reached the host, which
        may be a step old; particles outside the grid are clamped into its
        border cells, so the neighbour search stays exact.  Bounds are read
        synchronously at the first binning and whenever an array's h has
        changed, since h sets the search radius."""
        if not self.initialized:
            self.initialize()
        if not self.capturing:
            self._poll_jumps(wait=False)
        self.update_status()
        if not self.is_dirty and self.grid is not None:
            return 0
        if self.periodic_domain is not None:
            for pa in self.arrays_to_bin:
                if pa.is_dirty or pa._grid_key is None:
                    self.wrap_periodic(pa)
        if self.fixed_domain is not None:
            lo, hi = self.fixed_domain
        elif self.capturing and self.grid is not None:
            lo = hi = None
        elif (self.async_bounds and self.grid is not None and
              self._h_versions == self._h_state()):
            lo = hi = None
            res = self._bounds_read.value() if self._bounds_read is not None else None
            if res is not None:
                self._bounds_read = None
                lo, hi, hmax = self._reduce_bounds(res)
                if hmax > 0 and hmax != self.max_h:
                    # h was changed in place on the device: size the grid
                    # from bounds read now
                    lo, hi = self._bounds()
            if self._bounds_read is None:
                self._bounds_read = self._launch_bounds()
        else:
            lo, hi = self._bounds()
            self._h_versions = self._h_state()
            self._bounds_read = None
        if lo is None:
            g = self.grid
        else:
            g = self.make_grid(lo, hi)
        if self.grid is not None and self.fixed_domain is None and self._covers(self.grid, g):
            # keep the previous grid when it still covers the bounds, so that
            # unchanged arrays keep valid bins and the jump check applies
            g = self.grid
        if self.periodic_domain is not None:
            self._check_period()
        key = self.grid_key(g)
        for pa in self.arrays_to_bin:
            if pa.is_dirty or pa._grid_key != key:
                self.bin_array(pa, g)
        self.grid = g
        self.is_dirty = False
        return 0

    def _covers(self, old, g):
        """Whether grid `old` holds grid `g` (made from newer bounds) with at
        least one cell of slack left on every side."""
        return old.cell_size == g.cell_size and all(
            old.lo[d] <= g.lo[d] + self.grid_slack - 1 and
            old.lo[d] + old.dims[d] >= g.lo[d] + g.dims[d] - self.grid_slack + 1
            for d in range(3))

    def grid_holds_bounds(self):
        """For steps replayed from a captured graph, which keep the grid:
        False when h changed or the latest bounds that reached the host are
        not held by the current grid (the next binning outside a graph then
        makes a new grid from them); otherwise True, and a new asynchronous
        bounds read is started."""
        if self.fixed_domain is not None:
            return self.grid is not None
        if self.grid is None or self._h_versions != self._h_state():
            return False
        if self._bounds_read is not None:
            res = self._bounds_read.value()
            if res is None:
                return True
            lo, hi, hmax = self._reduce_bounds(res)
            if (hmax > 0 and hmax != self.max_h) or not self._covers(
                    self.grid, self.make_grid(lo, hi)):
                return False
            self._bounds_read = None
        self._bounds_read = self._launch_bounds()
        return True

    def start_jump_read(self):
        """Start the asynchronous read-back of the device jump count (steps
        replayed from a graph, whose binnings start none)."""
        if self.check_jumps is True and self._jump_acc is not None and self._jump_read is None:
            self._jump_read = _AsyncRead(self._jump_acc)

    def bin_array(self, pa, g, pos=None, extra=(), extra_out=None):
        """Counting-sort pa by cell on grid g and permute its properties.
        `pos` (three tensors) bins by other positions than x, y, z (a
        predictor stage); `extra` per-particle tensors are permuted along and
        returned (into `extra_out` when given)."""
        n = pa.get_number_of_particles()
        key = self.grid_key(g)
        ncells = g.dims[0] * g.dims[1] * g.dims[2]
        dev = pa.device
        if n == 0:
            cell_start = torch.empty(ncells + 1, dtype=torch.int32, device=dev)
            stream = require_gpu(cell_start)
            _abi.call("sph_fill_i32", stream, cell_start.data_ptr(), ncells + 1, 0)
            pa.__dict__["_cell"] = torch.empty(0, dtype=torch.int32, device=dev)
            pa.__dict__["_cell_start"] = cell_start
            pa.__dict__["_grid_key"] = key
            pa.set_dirty(False)
            return
        stream = require_gpu(pa.get_carray("x"))
        # binning buffers kept per array: the permutation, the cell starts
        # and two cell arrays (each binning writes the one the previous left)
        bb = pa.__dict__.get("_bin_bufs")
        if bb is None or bb["n"] != n or bb["ncells"] != ncells or bb["device"] != dev:
            rb = bb["ref"] if bb is not None and bb["n"] == n and bb["device"] == dev else [
                torch.full((n,), _REF_NONE, dtype=torch.int64, device=dev) for _ in range(2)]
            bb = {"n": n, "ncells": ncells, "device": dev,
                  "perm": torch.empty(n, dtype=torch.int32, device=dev),
                  "cell": [torch.empty(n, dtype=torch.int32, device=dev) for _ in range(2)],
                  "cell_start": torch.empty(ncells + 1, dtype=torch.int32, device=dev),
                  # the jump check's reference cells (two buffers, as the
                  # cells); they outlive grid changes
                  "ref": rb}
            pa.__dict__["_bin_bufs"] = bb
        perm = bb["perm"]
        cell = bb["cell"][1] if pa._cell is bb["cell"][0] else bb["cell"][0]
        cell_start = bb["cell_start"]
        ws = self._workspace(n, ncells, dev)
        v = pa.view(("x", "y", "z", "h"))
        if pos is not None:
            v.x, v.y, v.z = (t.data_ptr() for t in pos)
        extra = list(extra)
        if extra_out is None:
            extra_out = [torch.empty_like(t) for t in extra]
        # the jump check (cell.pyx:1422-1437) on the reference's unbounded
        # cells, kept per particle and permuted with the particles; none
        # across a periodic wrap (cell.pyx:404-407)
        check = bool(self.check_jumps) and self.periodic_domain is None
        ref = pa.__dict__.get("_refcell")
        if ref is not bb["ref"][0] and ref is not bb["ref"][1]:
            ref = bb["ref"][0]
        ref_next = bb["ref"][1] if ref is bb["ref"][0] else bb["ref"][0]
        if check:
            if not bb.get("ref_valid", True):
                # not kept while the check was off: start again
                ref.fill_(_REF_NONE)
            bb["ref_valid"] = True
            # accumulated on the device; "deferred" reads it in
            # verify_jumps(), True also asynchronously after every binning
            if self._jump_acc is None or self._jump_acc.device != dev:
                self._jump_acc = torch.zeros(1, dtype=torch.int32, device=dev)
            _abi.call("sph_jump_check", stream, v, float(self.cell_size or 1.0), ref.data_ptr(),
                      self._jump_acc.data_ptr())
            if self.check_jumps is True and self._jump_read is None and not self.capturing:
                self._jump_read = _AsyncRead(self._jump_acc)
        else:
            bb["ref_valid"] = False
        _abi.call("sph_bin", stream, v, g, perm.data_ptr(), cell.data_ptr(),
                  cell_start.data_ptr(), None, None, ws.data_ptr(), ws.numel())
        if check:
            moved = pa.apply_permutation(perm, [ref] + extra, [ref_next] + list(extra_out))
            pa.__dict__["_refcell"] = moved[0]
            moved = moved[1:]
        else:
            moved = pa.apply_permutation(perm, extra, extra_out)
        pa.__dict__["_cell"] = cell
        pa.__dict__["_cell_start"] = cell_start
        pa.__dict__["_grid_key"] = key
        pa.__dict__["_bin_version"] = pa.__dict__.get("_bin_version", 0) + 1
        pa.set_dirty(False)
        self.num_binnings += 1
        return moved

    def _poll_jumps(self, wait):
        """Raise when the device jump count (read back asynchronously, or
        waited for) has grown since it was last seen."""
        if self._jump_read is None:
            return
        v = self._jump_read.value(wait=wait)
        if v is None:
            return
        self._jump_read = None
        nj = int(v[0])
        if nj > self._jump_seen:
            self._jump_seen = nj
            raise _abi.JumpToleranceError(
                "CellManager.update", _abi.SPH_EJUMP,
                "%d particle move(s) of more than one cell width (jump tolerance %d)"
                % (nj, self.jump_tolerance))

    def verify_jumps(self):
        """Raise if the jump check counted a particle that moved more than
        one cell between two binnings (waits for the device count)."""
        self._jump_read = None
        if self._jump_acc is not None and self.check_jumps:
            nj = int(self._jump_acc.item())
            if nj > self._jump_seen:
                self._jump_seen = nj
                raise _abi.JumpToleranceError(
                    "CellManager.update", _abi.SPH_EJUMP,
                    "%d particle move(s) of more than one cell width" % nj)
        return 0

    def get_number_of_particles(self):
        return sum(pa.get_number_of_particles() for pa in self.arrays_to_bin)

    def get_particle_representation(self, fname, fmt="pickle"):
        """Write the particles of every cell (cell.pyx:1505-1514, :702-731).

        fmt "pickle": the reference's dict {(i, j, k): {'cid', 'centroid',
        'positions': {array name: {'x', 'y', 'z', 'idx'}}}} of plain Python
        types, centroid = (cid + 0.5) * cell_size.  fmt "npz": per array
        `<name>_x/_y/_z/_idx` (float64 / int64, user order) and `<name>_cid`
        (int32 [n,3], floor(coordinate / cell_size)), plus cell_size.
        Returns the file name written."""
        import pickle

        cs = self.cell_size
        per = {}
        for pa in self.arrays_to_bin:
            x, y, z = (pa.get(c, only_real_particles=False).astype(np.float64)
                       for c in ("x", "y", "z"))
            idx = pa.get("idx", only_real_particles=False).astype(np.int64) \
                if pa.has_array("idx") else \
                np.arange(len(x), dtype=np.int64)
            cid = np.stack([np.floor(a / cs) for a in (x, y, z)], 1).astype(np.int32)
            per[pa.name] = (x, y, z, idx, cid)
        if fmt == "npz":
            out = {"cell_size": cs}
            for name, (x, y, z, idx, cid) in per.items():
                out.update({name + "_x": x, name + "_y": y, name + "_z": z, name + "_idx": idx,
                            name + "_cid": cid})
            if not fname.endswith(".npz"):
                fname = fname + ".npz"
            np.savez(fname, **out)
            return fname
        if fmt != "pickle":
            raise ValueError("fmt must be 'pickle' or 'npz'")
        ret = {}
        for name, (x, y, z, idx, cid) in per.items():
            for k in range(len(x)):
                key = tuple(int(c) for c in cid[k])
                cell = ret.get(key)
                if cell is None:
                    cell = {"cid": key,
                            "centroid": tuple((c + 0.5) * cs for c in key),
                            "positions": {p.name: {"x": [], "y": [], "z": [], "idx": []}
                                          for p in self.arrays_to_bin}}
                    ret[key] = cell
                d = cell["positions"][name]
                d["x"].append(float(x[k]))
                d["y"].append(float(y[k]))
                d["z"].append(float(z[k]))
                d["idx"].append(int(idx[k]))
        with open(fname, "wb") as f:
            pickle.dump(ret, f, protocol=2)
        return fname

    def cells_of(self, pa):
        """Host view of the binning of pa: dict IntPoint -> user indices (the
        reference's cells_dict content for one array)."""
        self.update()
        g = self.grid
        cs = pa._cell_start.cpu().numpy()
        order = pa._order_np()
        out = {}
        for c in np.nonzero(np.diff(cs))[0]:
            cx = c % g.dims[0]
            t = c // g.dims[0]
            cy = t % g.dims[1]
            cz = t // g.dims[1]
            ids = np.arange(cs[c], cs[c + 1])
            if order is not None:
                ids = order[ids]
            out[IntPoint(cx + g.lo[0], cy + g.lo[1], cz + g.lo[2])] = np.sort(ids)
        return out


class _AsyncRead:
    """A small device tensor copied to pinned host memory on the current
    stream behind an event; value() is the host copy once the copy has
    completed (None before, unless wait)."""

    def __init__(self, t):
        self.host = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
        self.host.copy_(t, non_blocking=True)
        self.event = torch.cuda.Event()
        self.event.record()

    def value(self, wait=False):
        if wait:
            self.event.synchronize()
        elif not self.event.query():
            return None
        return self.host.numpy()
