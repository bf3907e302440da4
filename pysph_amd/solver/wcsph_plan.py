"""Fused execution of the standard WCSPH calc set.

Matched calc order (the order the reference's dam-break and elliptical-drop
set-ups produce, examples/dam-break/dam_break.py:270-326 with summation
density):

    SPHRho       on any arrays (non-integrating, updates rho)
    TaitEquation on arrays that have an SPHRho calc (updates p, cs)
    MomentumEquation on the moving array F     (integrates u, v[, w])
    GravityForce     on F, optional            (integrates u, v[, w])
    PositionStepping on F                      (integrates x, y[, z])
    XSPHCorrection   on F, optional            (integrates x, y[, z])

Per RK2 step the plan issues, on one stream (FP32 storage, the defaults):

    re-bin F (counting sort + permutation)          sph_bin / sph_permute
    pack F's float4 copies                          sph_pack
    density + Tait of every array at X, one launch  sph_density_multi
    momentum + XSPH + predictor update -> X1, k1    sph_wcsph_stage(1)
    pack F's copies from X1                         sph_pack
    density + Tait of every array at X1             sph_density_multi
    momentum + XSPH + final update -> X (in place)  sph_wcsph_stage(2)

Stage 2 keeps the stage-1 binning and widens its stencil by the
predictor's largest displacement, measured on the device by stage 1 (the
same neighbour sets, one binning less per step); with stage2_rebin, stage
2 re-bins the moving array by X1 and carries X, X1 and k1 through the same
permutation, as the reference rebuilds its neighbour caches after the
predictor (integrator.py:621-645).  The grid's bins are half the
reference's cells (cells_per_radius = 2).  Optional neighbour lists
(enable_lists) let the momentum pass read the neighbours its density pass
recorded.  The update order inside the stage kernels is the reference's
per-calc order (mom, gravity, step, xsph), so results match the generic
path.

step() replays the RK2 step and the particle update after it from HIP
graphs where it can (use_graph): the binning alternates every particle
property between two buffers, so two graphs, captured once from the two
buffer parities, alternate and each replay is one step.  Host read-backs
(bounds, the jump count) stay outside the graphs.
"""
import ctypes

import torch

from .. import _abi
from .._device import ptr, require_gpu
from ..base.step_graph import StepGraphs
from ..sph import _ops
from ..sph.funcs import (GravityForce, MomentumEquation, PositionStepping, SPHRho,
                         TaitEquation, XSPHCorrection)


class WCSPHPlan:
    # moving-array size from which the density passes run concurrently
    CONCURRENT_MIN = 1 << 18

    def __init__(self, particles, rho_calcs, tait, mom, grav, xsph, dim, radius_scale):
        self.particles = particles
        self.cm = particles.cell_manager
        self.rho_calcs = rho_calcs
        self.tait = tait            # TaitEquation function instance or None
        self.tait_arrays = set()
        self.mom = mom
        self.fluid = mom.dest
        self.grav = grav
        self.xsph = xsph
        self.dim = dim
        self.radius_scale = radius_scale
        f = mom.funcs[0]
        g = (grav.funcs[0].gx, grav.funcs[0].gy, grav.funcs[0].gz) if grav else (0, 0, 0)
        t = tait
        self.prm = _ops.make_params(
            dim, radius_scale=radius_scale,
            co=t.co if t else 1.0, ro=t.ro if t else 1000.0, gamma=t.gamma if t else 7.0,
            alpha=f.alpha, beta=f.beta, eta=f.eta,
            eps=xsph.funcs[0].eps if xsph else 0.0, g=g, has_gravity=grav is not None,
            has_xsph=xsph is not None,
            symmetric=bool(getattr(particles, "variable_h", False)))
        # sources of the fused momentum pass: union of momentum and XSPH
        # sources in particle-array order, with their roles
        srcs, flags = [], []
        for pa in particles.arrays:
            fl = 0
            if any(s is pa for s in mom.sources):
                fl |= _abi.SPH_SRC_MOMENTUM
            if xsph is not None and any(s is pa for s in xsph.sources):
                fl |= _abi.SPH_SRC_XSPH
            if fl:
                srcs.append(pa)
                flags.append(fl)
        self.srcs = srcs
        self.flags = (ctypes.c_uint32 * len(flags))(*flags)
        self._bufs = None
        self.steps = 0
        self._pk = {}
        self._nl = {}
        self.use_list = False
        self.list_ok = False
        self.list_max_per = 128
        # list_reuse: the stage-1 lists (radius (1 + skin) * search radius)
        # serve stage 2 too, valid while particles move less than
        # skin * radius / 2; otherwise every stage builds exact lists
        self.list_reuse = False
        self.list_skin = 0.0
        # stage2_rebin: re-bin the moving array by X1 before stage 2 (the
        # reference rebuilds its caches after the predictor).  Off, stage 2
        # keeps the stage-1 binning and the packed passes widen their
        # stencil by the predictor's largest displacement (measured on the
        # device by the stage-1 kernel), which skips one binning and
        # permutation per step; the neighbour sets are the same.
        self.stage2_rebin = False
        # cache_static (opt-in): the density of a non-moving array keeps the
        # partial sums of its non-moving sources (boundary from boundary)
        # between calls while neither is re-binned (same bits, fewer pairs
        # evaluated per step)
        self.cache_static = False
        self._caches = {}
        # concurrent_density: the density passes of the arrays other than
        # the moving one run on a second stream (default priority), beside
        # the moving array's pass; the stage kernel waits for both.  None
        # (auto): on when the moving array has at least CONCURRENT_MIN
        # particles; below that the launches are short and the cross-stream
        # waits cost more than the overlap gains
        self.concurrent_density = None
        self._side = None
        # fused_density: the density passes of two arrays over the same
        # sources run as one launch whose blocks cover both destination
        # sets (sph_density_multi); the shorter pass fills the GPU while the
        # longer one drains, as the second stream did, without cross-stream
        # waits.  Off, or where it does not apply, the passes launch one by
        # one (concurrently per concurrent_density)
        self.fused_density = True
        # the packed fast path serves FP32 storage
        self.use_packed = self.fluid.dtype == torch.float32
        # the packed passes bin on a grid DEFAULT_CELLS_PER_RADIUS times
        # finer than the reference's cells (cell_size, cell ids and the jump
        # tolerance keep the reference's cell): fewer candidates per
        # neighbour, the same neighbour sets
        self.cells_per_radius = self.cm.bins_per_cell
        if self.use_packed:
            self.set_cells_per_radius(self.DEFAULT_CELLS_PER_RADIUS)
        # use_graph: step() replays captured HIP graphs where the step allows
        # it (graph_applies()); False launches every step eagerly
        self.use_graph = True
        self._sg = StepGraphs(particles.arrays, self.cm, get_extra=self._graph_extra,
                              set_extra=self._set_graph_extra, key_extra=self._graph_xkey)

    def enable_lists(self, on=True, skin=None, reuse=None):
        """Density passes record neighbour lists that the momentum pass of
        the same stage reads (when the density and momentum sources are the
        same arrays); with reuse, stage-1 lists of radius (1 + skin) serve
        stage 2 as well."""
        self.use_list = bool(on) and self.use_packed and self.list_ok
        if skin is not None:
            self.list_skin = float(skin)
        if reuse is not None:
            self.list_reuse = bool(reuse)
        return self.use_list

    DEFAULT_CELLS_PER_RADIUS = 2

    def set_lds_rows(self, on=True):
        """Stage the density scan's stencil rows in LDS (sph_params.lds_rows;
        the same sums, slower on MI355X at the sizes measured)."""
        self.prm.lds_rows = 1 if on else 0

    def set_cells_per_radius(self, k):
        """Bins of the shared CellManager's grid per reference cell (k = 1:
        the reference's cells)."""
        self.cm.set_cells_per_radius(k)
        self.cells_per_radius = k

    # ----------------------------------------------------------------- match
    @classmethod
    def match(cls, particles, calcs):
        def kind(c):
            return type(c.funcs[0]) if c.funcs else None

        i = 0
        n = len(calcs)
        rho_calcs = []
        while i < n and kind(calcs[i]) is SPHRho and not calcs[i].integrates:
            if any(getattr(f, "exclude_self", False) for f in calcs[i].funcs):
                return None  # the per-calc path drops the self pairs
            rho_calcs.append(calcs[i])
            i += 1
        tait_calcs = []
        while i < n and kind(calcs[i]) is TaitEquation and not calcs[i].integrates:
            tait_calcs.append(calcs[i])
            i += 1

        def take(k, integ=True):
            nonlocal i
            if i < n and kind(calcs[i]) is k and calcs[i].integrates == integ:
                i += 1
                return calcs[i - 1]
            return None

        mom = take(MomentumEquation)
        grav = take(GravityForce)
        step = take(PositionStepping)
        xsph = take(XSPHCorrection)
        if i != n or mom is None or step is None or not rho_calcs:
            return None
        fluid = mom.dest
        dim = mom.kernel.dim
        vel = ["u", "v", "w"][:dim]
        pos = ["x", "y", "z"][:dim]
        if mom.updates != vel or step.updates != pos or step.dest is not fluid:
            return None
        if grav is not None and (grav.dest is not fluid or grav.updates != vel):
            return None
        if xsph is not None and (xsph.dest is not fluid or xsph.updates != pos):
            return None
        if any(c.updates != ["rho"] for c in rho_calcs):
            return None
        rho_dests = [c.dest for c in rho_calcs]
        if len({id(d) for d in rho_dests}) != len(rho_dests):
            return None
        tait_f = None
        tait_arrays = []
        for c in tait_calcs:
            f = c.funcs[0]
            if c.updates != ["p", "cs"] or not any(d is c.dest for d in rho_dests):
                return None
            if tait_f is not None and (f.co, f.ro, f.gamma) != (tait_f.co, tait_f.ro,
                                                               tait_f.gamma):
                return None
            tait_f = f
            tait_arrays.append(c.dest)
        if len({pa.dtype for pa in particles.arrays}) != 1:
            return None
        kernel = mom.kernel
        if rho_calcs[0].kernel.radius() != kernel.radius():
            return None
        if any(len(c.sources) > _abi.SPH_MAX_SRC for c in rho_calcs + [mom]):
            return None
        plan = cls(particles, rho_calcs, tait_f, mom, grav, xsph, dim, kernel.radius())
        plan.tait_arrays = {id(a) for a in tait_arrays}
        # packed (rho, p, cs) copies come from the density + Tait pass of
        # each array; every momentum source needs one
        plan.use_packed = plan.use_packed and all(id(pa) in plan.tait_arrays
                                                  for pa in plan.srcs)
        # the moving array's density pass can hand its neighbours to the
        # momentum pass when both see the same sources in the same order
        frho = [c for c in rho_calcs if c.dest is fluid]
        plan.list_ok = bool(frho) and [id(x) for x in frho[0].sources] == [
            id(x) for x in plan.srcs] and all(len(c.sources) > 0 for c in rho_calcs)
        # lists are opt-in (enable_lists): on MI355X the second cell search
        # of the momentum pass costs about what writing and reading the
        # list costs, and searching moves no list through HBM
        plan.use_list = False
        plan.list_max_per = 128 if dim == 3 else 64
        if len(plan.srcs) > _abi.SPH_MAX_SRC:
            return None
        return plan

    # --------------------------------------------------------------- buffers
    def _buffers(self):
        pa = self.fluid
        n = pa.get_number_of_particles()
        if self._bufs is None or self._bufs["n"] != n or self._bufs["dtype"] != pa.dtype:
            mk = lambda: torch.empty(n, dtype=pa.dtype, device=pa.device)  # noqa: E731
            self._bufs = {
                "n": n, "dtype": pa.dtype,
                "s1": [mk() for _ in range(6)],
                "k1": [mk() for _ in range(6)],
                # the second set X1 and k1 alternate with at a re-binning
                "alt": [mk() for _ in range(12)],
                "disp": torch.zeros(1, dtype=torch.float32, device=pa.device),
            }
        return self._bufs

    # ------------------------------------------------------ packed copies
    # FP32 storage: every array carries float4 copies (x,y,z,h), (u,v,w,m),
    # (rho,cs,p/rho^2,p) and (x,y,z,m), an (x,y,z) copy for the candidate
    # screen and its h range, in its current slot order, read by the fast
    # neighbour passes.  All but aux4 are re-packed after every binning of
    # the array (and from X1 for stage 2); aux4 is written by the density
    # pass that precedes every momentum pass.
    def _packed(self, pa):
        if not self.use_packed:
            return None
        n = pa.get_number_of_particles()
        pk = self._pk.get(id(pa))
        if pk is None or pk["n"] != n:
            mk = lambda: torch.empty((max(n, 1), 4), dtype=torch.float32,  # noqa: E731
                                     device=pa.device)
            pk = {"n": n, "pos4": mk(), "vel4": mk(), "aux4": mk(), "pm4": mk(),
                  "pos3": torch.empty(max(n, 1) * 3, dtype=torch.float32, device=pa.device),
                  "hrange": torch.empty(2, dtype=torch.int32, device=pa.device),
                  "mrange": torch.empty(2, dtype=torch.int32, device=pa.device),
                  "version": None}
            self._pk[id(pa)] = pk
        return pk

    def _pack(self, pa, fields=None, force=False):
        pk = self._packed(pa)
        if pk is None:
            return
        ver = pa.__dict__.get("_bin_version", 0)
        if not force and pk["version"] == ver:
            return
        stream = require_gpu(pk["pos4"])
        v = pa.view()
        for f, t in (fields or {}).items():
            setattr(v, f, ptr(t))
        # the stage-2 repack (fields: X1 positions and velocities) keeps h
        # and m, so the uniformity pairs of the step's first pack stay valid
        hr = None if fields else pk["hrange"].data_ptr()
        mr = None if fields else pk["mrange"].data_ptr()
        _abi.call("sph_pack", stream, v, pk["pos4"].data_ptr(), pk["vel4"].data_ptr(),
                  pk["pos3"].data_ptr(), pk["pm4"].data_ptr(), hr, mr)
        pk["version"] = ver

    def _pack_all(self):
        for pa in self.particles.arrays:
            self._pack(pa)

    def _override(self, s1=None):
        """Per-array field overrides: packed copies and, for the moving
        array at stage 2, the X1 state."""
        ov = {}
        if self.use_packed:
            for pa in self.particles.arrays:
                pk = self._pk[id(pa)]
                ov[id(pa)] = {"pos4": pk["pos4"], "vel4": pk["vel4"], "aux4": pk["aux4"],
                              "pos3": pk["pos3"], "pm4": pk["pm4"], "hrange": pk["hrange"],
                              "mrange": pk["mrange"]}
        if s1 is not None:
            ov.setdefault(id(self.fluid), {}).update(s1)
        return ov

    def _state0(self):
        pa = self.fluid
        return [pa.get_carray(f) for f in ("x", "y", "z", "u", "v", "w")]

    def _s1_fields(self, b):
        """Fields of the stage state X1; components >= dim stay those of X."""
        s0 = self._state0()
        names = ("x", "y", "z", "u", "v", "w")
        ov = {}
        for c in range(6):
            comp = c % 3
            ov[names[c]] = b["s1"][c] if comp < self.dim else s0[c]
        return ov

    def _nlist(self, pa, mode):
        """Neighbour list of array `pa` (superset radius (1+skin) * radius):
        mode 0 = built by the density pass, 1 = read by the later passes."""
        if not self.use_list:
            return None
        n = pa.get_number_of_particles()
        nl = self._nl.get(id(pa))
        if nl is None or nl["n"] != n:
            maxp = self.list_max_per
            nl = {"n": n,
                  "idx": torch.empty(max(n, 1) * maxp, dtype=torch.int32, device=pa.device),
                  "count": torch.empty(max(n, 1) * _abi.SPH_MAX_SRC, dtype=torch.int32,
                                       device=pa.device)}
            self._nl[id(pa)] = nl
        st = _abi.SphNlist()
        st.idx = nl["idx"].data_ptr()
        st.count = nl["count"].data_ptr()
        st.max_per = self.list_max_per
        st.mode = mode
        st.skin = self.list_skin
        return st

    def _src_cache(self, c):
        """SphSrcCache for a density calc whose destination and one source
        do not move (boundary from boundary): the source's partial sums are
        computed once and reused while neither array is re-binned."""
        pa = c.dest
        if not self.cache_static or pa is self.fluid:
            return None
        k = next((i for i, s in enumerate(c.sources) if s is not self.fluid), None)
        if k is None:
            return None
        src = c.sources[k]
        n = pa.get_number_of_particles()
        key = (pa.__dict__.get("_bin_version", 0), src.__dict__.get("_bin_version", 0),
               self.cm.grid_key(self.cm.grid), n)
        ent = self._caches.get(id(c))
        if ent is None or ent["sums"].numel() < max(n, 1):
            ent = {"sums": torch.empty(max(n, 1), dtype=torch.float64, device=pa.device),
                   "key": None}
            self._caches[id(c)] = ent
        st = _abi.SphSrcCache()
        st.sums = ent["sums"].data_ptr()
        st.src = k
        st.mode = _abi.SPH_CACHE_USE if ent["key"] == key else _abi.SPH_CACHE_FILL
        ent["key"] = key
        return st

    def _fuse_density(self):
        """Whether the density passes of the step's two arrays run as one
        launch (sph_density_multi): same sources in the same order, packed
        inputs, no lists, no source cache, FP32 pair arithmetic."""
        if not self.fused_density or len(self.rho_calcs) != 2 or not self.use_packed:
            return False
        if self.use_list or self.cache_static or self.prm.pair_prec == _abi.SPH_PAIR_F64:
            return False
        a, b = self.rho_calcs
        same = [id(x) for x in a.sources] == [id(x) for x in b.sources]
        tait = {id(a.dest) in self.tait_arrays, id(b.dest) in self.tait_arrays}
        return same and len(tait) == 1 and len(a.sources) <= _abi.SPH_MAX_SRC

    def _density(self, disp=None, override=None, list_mode=0):
        prm = self.prm
        if self._fuse_density():
            c0, c1 = self.rho_calcs
            tait = id(c0.dest) in self.tait_arrays
            dsts = [c0.dest, c1.dest]
            _ops.density_multi(self.cm, prm, dsts, c0.sources,
                               [d.get_carray("rho") for d in dsts],
                               [d.get_carray("p") if tait else None for d in dsts],
                               [d.get_carray("cs") if tait else None for d in dsts],
                               disp=disp, override=override)
            return
        main = torch.cuda.current_stream()
        side = None
        conc = self.concurrent_density
        if conc is None:
            conc = self.fluid.get_number_of_particles() >= self.CONCURRENT_MIN
        # a captured step stays on one stream (its launches are replayed in
        # order on the launch stream)
        if conc and len(self.rho_calcs) > 1 and not self.cm.capturing:
            if self._side is None:
                self._side = torch.cuda.Stream(device=self.fluid.device)
            side = self._side
            # the side stream starts after everything queued so far
            side.wait_stream(main)
        for c in self.rho_calcs:
            pa = c.dest
            rho = pa.get_carray("rho")
            # only the moving array has a momentum pass that reads a list
            nl = self._nlist(pa, list_mode) if pa is self.fluid else None
            cache = self._src_cache(c) if nl is None else None
            args = (self.cm, prm, pa, c.sources, rho)
            if id(pa) in self.tait_arrays:
                args += (pa.get_carray("p"), pa.get_carray("cs"))
            if side is not None and pa is not self.fluid:
                with torch.cuda.stream(side):
                    _ops.density(*args, disp=disp, override=override, nl=nl, cache=cache)
            else:
                _ops.density(*args, disp=disp, override=override, nl=nl, cache=cache)
        if side is not None:
            # the momentum pass reads every array's density
            main.wait_stream(side)

    def _stage(self, stage, dt, b, override=None, disp_in=None, disp_out=None):
        pa = self.fluid
        stream = require_gpu(pa.get_carray("x"))
        vd = _ops.views([pa], override)
        vs = _ops.views(self.srcs, override)
        s0 = _abi.ptr_array([ptr(t) for t in self._state0()])
        k1 = _abi.ptr_array([ptr(t) for t in b["k1"]])
        if stage == 2:
            out = s0
        else:
            out = _abi.ptr_array([ptr(t) for t in b["s1"]])
        _abi.call("sph_wcsph_stage", stream, stage, self.cm.grid, self.prm, vd, len(self.srcs),
                  vs, self.flags, _ops.search_radius(self.cm, self.radius_scale), ptr(disp_in),
                  float(dt), s0, k1, out, ptr(disp_out), self._nlist(pa, 1))

    # ----------------------------------------------------------------- steps
    def rebin_stage(self, b):
        """Re-bin the moving array by X1; X, X1 and k1 follow the new order
        (X1 and k1 permuted into the step buffers' second set, which the
        next re-binning takes back)."""
        s0 = self._state0()
        # components >= dim are not integrated: X1 keeps those of X
        pos = [b["s1"][c] if c < self.dim else s0[c] for c in range(3)]
        old = b["s1"] + b["k1"]
        moved = self.cm.bin_array(self.fluid, self.cm.grid, pos=pos, extra=old,
                                  extra_out=b["alt"])
        b["s1"] = moved[:6]
        b["k1"] = moved[6:]
        b["alt"] = old

    def rk2_step(self, dt, probe=None):
        """One RK2 step.  `probe` (optional) is called as probe(phase, bufs,
        override) after each phase ("density1", "stage1", "density2",
        "stage2") with the stage buffers and the field overrides the next
        pass reads, so that a test can inspect the intermediate sums."""
        if probe is None:
            probe = _no_probe
        self.cm.update()
        b = self._buffers()
        self._pack_all()
        ov = self._override()
        # lists built anew in stage 2 need the X1 binning (their validity
        # check reads the same displacement as the widened stencil)
        keep_binning = (self.use_list and self.list_reuse) or (
            self.use_packed and not self.stage2_rebin and not self.use_list)
        if keep_binning:
            # stage 2 keeps the binning: the stage-1 kernel records the
            # predictor's largest displacement on the device, and the stage-2
            # passes widen their stencil by it; with reused lists, stage 1
            # builds the neighbour lists (superset radius) that the stage-1
            # momentum pass and every stage-2 pass read, valid while that
            # displacement stays within the list skin
            stream = require_gpu(b["disp"])
            _abi.call("sph_fill_i32", stream, b["disp"].data_ptr(), 1, 0)
            self._density(override=ov, list_mode=0)
            probe("density1", b, ov)
            self._stage(1, dt, b, override=ov, disp_out=b["disp"])
            probe("stage1", b, ov)
            s1 = self._s1_fields(b)
            self._pack(self.fluid, fields=s1, force=True)
            ov = self._override(s1)
            self._density(disp=b["disp"], override=ov, list_mode=1)
            probe("density2", b, ov)
            self._stage(2, dt, b, override=ov, disp_in=b["disp"])
            probe("stage2", b, ov)
        else:
            # each stage bins its own positions; with lists, the density pass
            # of each stage records the neighbours its momentum pass reads
            self._density(override=ov, list_mode=0)
            probe("density1", b, ov)
            self._stage(1, dt, b, override=ov)
            probe("stage1", b, ov)
            self.rebin_stage(b)
            s1 = self._s1_fields(b)
            self._pack(self.fluid, fields=s1, force=True)
            ov = self._override(s1)
            self._density(override=ov, list_mode=0)
            probe("density2", b, ov)
            self._stage(2, dt, b, override=ov)
            probe("stage2", b, ov)
        self.fluid.set_dirty(True)
        self.steps += 1

    # ------------------------------------------------------------ HIP graphs
    def graph_applies(self):
        """Whether step() may replay graphs: no neighbour lists or source
        caches (their validity is decided on the host), no periodic axes and
        no functions run by Particles.update()."""
        pt = self.particles
        return bool(self.use_graph and not self.use_list and not self.cache_static and
                    self.cm.periodic_domain is None and
                    not getattr(pt, "misc_prop_update_functions", None) and
                    not getattr(pt, "correction_manager", None))

    @property
    def graph_replays(self):
        return self._sg.replays

    @property
    def graph_captures(self):
        return self._sg.captures

    @property
    def graph_error(self):
        return self._sg.error

    def drop_graphs(self):
        self._sg.drop()

    def _graph_extra(self):
        b = self._bufs
        bufs = None if b is None else {k: list(b[k]) for k in ("s1", "k1", "alt")}
        return {"pk": {k: v["version"] for k, v in self._pk.items()}, "bufs": bufs}

    def _set_graph_extra(self, ex):
        for k, v in ex["pk"].items():
            self._pk[k]["version"] = v
        if ex["bufs"] is not None and self._bufs is not None:
            for k, v in ex["bufs"].items():
                self._bufs[k] = list(v)

    def _graph_xkey(self):
        b = self._bufs
        if b is None:
            return None
        return tuple(id(t) for k in ("s1", "k1", "alt") for t in b[k])

    def _graph_tag(self, dt):
        """What a captured step bakes in besides buffers and the grid: the
        time step, the parameters and the plan's switches."""
        return (float(dt), bytes(self.prm), bytes(self.flags), self.fused_density,
                self.concurrent_density, self.use_packed, self.stage2_rebin)

    def _eager_step(self, dt, euler=False):
        if euler:
            self.euler_step(dt)
        else:
            self.rk2_step(dt)
        self.particles.update()

    def step(self, dt, euler=False):
        """One RK2 (or Euler) step and the particle update after it (what
        RK2Integrator.integrate / EulerIntegrator.integrate run), replayed
        from a captured HIP graph when graph_applies() and the state is one
        a graph was captured from, launched eagerly otherwise (the first
        steps, after a grid change or any change of the arrays' buffers)."""
        cm = self.cm
        n0 = self.steps
        if not self.graph_applies():
            self._sg.drop()
            self._eager_step(dt, euler)
        else:
            cm._poll_jumps(wait=False)
            if not cm.grid_holds_bounds():
                # the next binning makes a new grid: eager, then capture again
                self._sg.drop()
                self._eager_step(dt, euler)
            elif self._sg.step(lambda: self._eager_step(dt, euler),
                               tag=self._graph_tag(dt) + (euler,)):
                self.particles.update()
                cm.start_jump_read()
        self.steps = n0 + 1

    def euler_step(self, dt):
        self.cm.update()
        b = self._buffers()
        self._pack_all()
        ov = self._override()
        self._density(override=ov, list_mode=0)
        self._stage(0, dt, b, override=ov)
        pa = self.fluid
        names = ("x", "y", "z", "u", "v", "w")
        for c in range(6):
            if c % 3 < self.dim:
                nm = names[c]
                old = pa._props[nm]
                pa._props[nm] = b["s1"][c]
                b["s1"][c] = old
        pa.set_dirty(True)
        self.steps += 1


def _no_probe(phase, bufs, override):
    pass
