"""The radius graph of a particle set on the GPU, as the network reads it.

Fluid and boundary particles live in one combined ParticleArray (user order =
fluid then boundary) with an int property `btype` (0 fluid, 1 boundary).
Building the graph bins the array (CellManager), exports the neighbour CSR
(within 2h, self included; symmetric because h is uniform) and computes the
summation density of every node with the plain density pass.
"""
import ctypes

import numpy as np
import torch

from .. import _abi
from .._device import require_gpu
from ..base.cell import CellManager
from ..base.particle_array import get_particle_array
from ..sph import _ops

FIELDS = ("x", "y", "z", "u", "v", "w", "h", "m", "rho", "p", "cs")


def combined_array(fluid, boundary):
    """One FP32 ParticleArray holding fluid then boundary particles."""
    parts = [fluid] + ([boundary] if boundary is not None else [])
    data = {f: np.concatenate([pa.get(f).astype(np.float64) for pa in parts]) for f in FIELDS}
    nf = fluid.get_number_of_particles()
    n = len(data["x"])
    pa = get_particle_array(cl_precision="single", name="graph", **data)
    bt = np.zeros(n, dtype=np.int32)
    bt[nf:] = 1
    pa.add_property({"name": "btype", "type": "int", "data": bt})
    return pa


class ParticleGraph:
    """Graph builder over a combined array; `struct` is the sph_mlp_graph of
    the last build."""

    def __init__(self, pa, info, domain=None, cells_per_radius=2, rho_scale=10.0, max_per=48):
        self.pa = pa
        self.info = dict(info)
        self.device = pa.device
        self.cm = CellManager(arrays_to_bin=[pa])
        if domain is not None:
            self.cm.set_fixed_domain(*domain)
        self.cm.check_jumps = False
        self.cm.set_cells_per_radius(cells_per_radius)
        self.rho_scale = rho_scale
        self.prm = _ops.make_params(2, radius_scale=2.0)
        self.n = pa.get_number_of_particles()
        bt = pa.get("btype")
        self.n_fluid = int((bt == 0).sum())
        self.off = torch.empty(self.n + 1, dtype=torch.int64, device=self.device)
        self.counts = torch.empty(max(self.n, 1), dtype=torch.int64, device=self.device)
        self.ids = torch.arange(max(self.n, 1), dtype=torch.int64, device=self.device)
        # edge capacity: a bound on the neighbours per node (2D, 2h support,
        # h = 1.3 dx: ~21 at rest density) with 2x headroom; the CSR is
        # filled without reading its total back, and an overflow is flagged
        # on the device and raised at the next check
        self.max_per = max_per
        self.cap = max(self.n, 1) * max_per
        self.nbr = torch.empty(self.cap, dtype=torch.int64, device=self.device)
        self.overflow = torch.zeros(1, dtype=torch.int32, device=self.device)
        self.struct = _abi.SphMlpGraph()

    def build(self, edges=True):
        """Bin the array, export the radius-2h CSR and sum the density.
        With edges=False only the binning and the summation density run
        (what a density check needs); the CSR and the network's graph
        struct keep the last full build."""
        pa, cm = self.pa, self.cm
        cm.update()
        stream = require_gpu(self.off)
        if not edges:
            _ops.density(cm, self.prm, pa, [pa], pa.get_carray("rho"),
                         override=_ops.packed_override([pa]))
            return self
        rmax = 2.0 * cm.max_h
        v = pa.view()
        _abi.call("sph_neighbor_counts", stream, cm.grid, v, v, 2.0, self.counts.data_ptr())
        # the scan without a device allocation, so that a build can be
        # captured into a HIP graph
        scan = ("sph_scan_counts_noalloc" if self.n <= _abi.SPH_SCAN_NOALLOC_MAX_COUNTS
                else "sph_scan_counts")
        _abi.call(scan, stream, self.n, self.counts.data_ptr(), self.off.data_ptr(), None)
        _abi.call("sph_neighbor_fill_capped", stream, cm.grid, v, v, 2.0, self.ids.data_ptr(),
                  self.off.data_ptr(), self.nbr.data_ptr(), self.cap,
                  self.overflow.data_ptr())
        # the network sizes its edge buffers by the capacity; the exact
        # count stays on the device (nedges_exact() reads it)
        ne = self.cap
        # summation density of every node, packed copies for the network
        ov = _ops.packed_override([pa])
        _ops.density(cm, self.prm, pa, [pa], pa.get_carray("rho"), override=ov)
        sc = pa.__dict__["_pack_scratch"]
        # the density pass packed (x,y,z,h), (u,v,w,m) into sc
        s = self.struct
        inf = self.info
        s.n, s.nedges, s.n_fluid, s.dim = self.n, ne, self.n_fluid, 2
        s.pos4 = sc["pos4"].data_ptr()
        s.vel4 = sc["vel4"].data_ptr()
        s.rho = pa.get_carray("rho").data_ptr()
        s.btype = pa.get_carray("btype").data_ptr()
        s.off = self.off.data_ptr()
        s.nbr = self.nbr.data_ptr()
        s.rho0, s.co, s.gamma = inf["rho0"], inf["co"], inf.get("gamma", 7.0)
        s.u_scale = inf["co"] / 10.0
        s.g = 9.81
        s.rho_scale = self.rho_scale
        s.alpha, s.eta = inf.get("alpha", 0.1), inf.get("eta", 0.1)
        self.nedges = ne
        _ = rmax
        return self

    def nedges_exact(self):
        """Edges of the last build (synchronises)."""
        return int(self.off[self.n].item())

    def check_overflow(self):
        """Raise if a build exceeded the edge capacity (synchronises)."""
        if int(self.overflow.item()) != 0:
            raise RuntimeError("surrogate graph: a node has more than %d neighbours; build the "
                               "ParticleGraph with a larger max_per" % self.max_per)

    # ---- user-order access on the device
    def user_to_slot(self, t):
        o = self.pa._order
        return t if o is None else t[o]

    def slot_to_user(self, t):
        o = self.pa._order
        if o is None:
            return t.clone()
        out = torch.empty_like(t)
        out[o] = t
        return out

    def user_rows_to_slot(self, t):
        """Rows of a per-node [n x k] tensor from user to slot order."""
        o = self.pa._order
        return t if o is None else t[o].contiguous()

    def slot_rows_to_user(self, t):
        o = self.pa._order
        if o is None:
            return t
        out = torch.empty_like(t)
        out[o] = t
        return out

    def get_state(self):
        """x, y, u, v in user order (device tensors)."""
        return {f: self.slot_to_user(self.pa.get_carray(f)) for f in ("x", "y", "u", "v")}

    def set_state(self, st):
        for f, t in st.items():
            self.pa.get_carray(f).copy_(self.user_to_slot(t))
        self.pa.set_dirty(True)
