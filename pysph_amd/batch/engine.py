"""Batched multi-scene WCSPH engine: hundreds of independent scenes of uneven
size advanced together, one launch per phase, optionally as one HIP graph
per step.

The reference steps one `Particles` set at a time (solver.py:457-508); this
engine keeps that scheme per scene (summation density, Tait, momentum with
artificial viscosity, optional gravity and XSPH, RK2) and batches the
scenes through the C ABI's sph_batch_* calls.  A scene advanced here gives
the bits it gives when advanced alone by the WCSPH solver on the same grid.
"""
import ctypes
import math

import numpy as np
import torch

from .. import _abi
from .._device import require_gpu
from ..base.cell import grid_for_domain
from ..sph import _ops

FIELDS = _abi.BATCH_FIELDS


class Scene:
    """One scene: fluid (and optional boundary) particle data in user order
    and its scheme constants.

    fluid, boundary: dicts field -> numpy array (x, y, z, u, v, w, h, m, rho,
    p, cs); boundary may be None.  info: co, rho0, gamma, alpha, beta, eta,
    eps (XSPH, 0 = none), g (3-tuple or None), dt, domain ((lo), (hi))."""

    def __init__(self, fluid, boundary, info, dim=2, name=""):
        self.fluid = {f: np.asarray(fluid[f], dtype=np.float32) for f in FIELDS}
        self.boundary = None
        if boundary is not None and len(boundary["x"]) > 0:
            self.boundary = {f: np.asarray(boundary[f], dtype=np.float32) for f in FIELDS}
        self.info = dict(info)
        self.dim = dim
        self.name = name

    @property
    def nf(self):
        return len(self.fluid["x"])

    @property
    def nb(self):
        return 0 if self.boundary is None else len(self.boundary["x"])

    @property
    def max_h(self):
        h = [float(self.fluid["h"].max())]
        if self.boundary is not None:
            h.append(float(self.boundary["h"].max()))
        return max(h)


def scene_from_arrays(fluid, boundary, info, dim=2, name=""):
    """A Scene from ParticleArrays (user order)."""
    def grab(pa):
        if pa is None:
            return None
        return {f: pa.get(f).astype(np.float32) for f in FIELDS}
    return Scene(grab(fluid), grab(boundary), info, dim=dim, name=name)


def default_schedule():
    return {"mode": 0, "blocks_per_cu": 4, "grab": 1, "order": 0}


def _sched_struct(d):
    s = _abi.SphBatchSched()
    s.mode, s.blocks_per_cu, s.grab, s.order = (int(d[k]) for k in
                                                ("mode", "blocks_per_cu", "grab", "order"))
    return s


class BatchedWCSPH:
    """RK2 WCSPH of many scenes in one set of device buffers."""

    def __init__(self, scenes, cells_per_radius=2, radius_scale=2.0, device=None,
                 grid_slack=2):
        if not scenes:
            raise ValueError("no scenes")
        dim = scenes[0].dim
        if any(s.dim != dim for s in scenes):
            raise ValueError("all scenes of a batch must share their dimension")
        self.scenes = list(scenes)
        self.dim = dim
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        self.radius_scale = float(radius_scale)
        self.cells_per_radius = int(cells_per_radius)
        dev = self.device
        nf = sum(s.nf for s in scenes)
        nb = sum(s.nb for s in scenes)
        self.nf, self.nb = nf, nb
        self.f_off = np.cumsum([0] + [s.nf for s in scenes])
        self.b_off = np.cumsum([0] + [s.nb for s in scenes])

        def cat(part, f):
            arrs = [getattr(s, part)[f] for s in scenes if getattr(s, part) is not None]
            return np.concatenate(arrs) if arrs else np.zeros(0, np.float32)

        def dset(n, init=None):
            d = {}
            for f in FIELDS:
                t = torch.empty(max(n, 1), dtype=torch.float32, device=dev)
                if init is not None and n:
                    t[:n].copy_(torch.from_numpy(init[f]))
                d[f] = t
            d["uid"] = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
            return d

        fin = {f: cat("fluid", f) for f in FIELDS}
        bin_ = {f: cat("boundary", f) for f in FIELDS}
        self.fa = dset(nf, fin)
        self.fb = dset(nf)
        self.bnd = dset(nb)
        self.bnd_in = dset(nb, bin_)
        # uid: index within the scene's user order
        fu = np.concatenate([np.arange(s.nf) for s in scenes]).astype(np.int32)
        bu = np.concatenate([np.arange(s.nb) for s in scenes] + [np.zeros(0)]).astype(np.int32)
        if nf:
            self.fa["uid"][:nf].copy_(torch.from_numpy(fu))
        if nb:
            self.bnd_in["uid"][:nb].copy_(torch.from_numpy(bu))
        mk = lambda n: torch.zeros(max(n, 1), dtype=torch.float32, device=dev)  # noqa: E731
        self.x1a = [mk(nf) for _ in range(6)]
        self.x1b = [mk(nf) for _ in range(6)]
        self.k1a = [mk(nf) for _ in range(6)]
        self.k1b = [mk(nf) for _ in range(6)]
        mk4 = lambda n: torch.zeros((max(n, 1), 4), dtype=torch.float32, device=dev)  # noqa: E731
        self.fpos4, self.fvel4, self.faux4 = mk4(nf), mk4(nf), mk4(nf)
        self.bpos4, self.bvel4, self.baux4 = mk4(nb), mk4(nb), mk4(nb)
        fsc = np.concatenate([np.full(s.nf, k) for k, s in enumerate(scenes)]).astype(np.int32)
        bsc = np.concatenate([np.full(s.nb, k) for k, s in enumerate(scenes)] +
                             [np.zeros(0)]).astype(np.int32)
        self.fscene = torch.zeros(max(nf, 1), dtype=torch.int32, device=dev)
        self.bscene = torch.zeros(max(nb, 1), dtype=torch.int32, device=dev)
        if nf:
            self.fscene[:nf].copy_(torch.from_numpy(fsc))
        if nb:
            self.bscene[:nb].copy_(torch.from_numpy(bsc))

        # scene table
        arr = (_abi.SphBatchScene * len(scenes))()
        self.grids = []
        for k, s in enumerate(scenes):
            e = arr[k]
            e.f0, e.nf = int(self.f_off[k]), s.nf
            e.b0, e.nb = int(self.b_off[k]), s.nb
            cs = self.radius_scale * s.max_h / self.cells_per_radius
            lo, hi = s.info["domain"]
            e.grid = grid_for_domain(lo, hi, cs, grid_slack)
            self.grids.append(e.grid)
            inf = s.info
            g = inf.get("g") or (0.0, 0.0, 0.0)
            eps = float(inf.get("eps") or 0.0)
            e.prm = _ops.make_params(dim, radius_scale=self.radius_scale, co=inf["co"],
                                     ro=inf["rho0"], gamma=inf.get("gamma", 7.0),
                                     alpha=inf["alpha"], beta=inf.get("beta", 0.0),
                                     eta=inf.get("eta", 0.1), eps=eps, g=g,
                                     has_gravity=inf.get("g") is not None,
                                     has_xsph=eps != 0.0)
            e.rmax = self.radius_scale * s.max_h
            e.dt = float(inf["dt"])
            # cost per particle ~ neighbours within the search radius
            dx = float(inf.get("dx", s.max_h))
            e.cost = (2.0 * s.max_h / dx) ** dim
        self._scene_arr = arr
        L = _abi.load()
        self.ncells = int(L.sph_batch_ncells(len(scenes), arr))
        wsb = int(L.sph_bin_workspace_bytes(max(nf, nb), self.ncells))
        self.ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
        i32 = lambda n: torch.zeros(max(n, 1), dtype=torch.int32, device=dev)  # noqa: E731
        self.fperm, self.fcell, self.fcell_start = i32(nf), i32(nf), i32(self.ncells + 1)
        self.bperm, self.bcell, self.bcell_start = i32(nb), i32(nb), i32(self.ncells + 1)

        B = _abi.SphBatchBuffers()
        B.nf_total, B.nb_total = nf, nb
        for name, d in (("fa", self.fa), ("fb", self.fb), ("bnd", self.bnd),
                        ("bnd_in", self.bnd_in)):
            st = getattr(B, name)
            for f in FIELDS + ("uid",):
                setattr(st, f, d[f].data_ptr())
        for name in ("x1a", "x1b", "k1a", "k1b"):
            a = getattr(B, name)
            for c, t in enumerate(getattr(self, name)):
                a[c] = t.data_ptr()
        for name in ("fpos4", "fvel4", "faux4", "bpos4", "bvel4", "baux4", "fscene", "bscene",
                     "fperm", "fcell", "fcell_start", "bperm", "bcell", "bcell_start", "ws"):
            setattr(B, name, getattr(self, name).data_ptr())
        B.ws_bytes = wsb
        self._bufs = B
        h = _abi._vp()
        _abi.call("sph_batch_create", len(scenes), arr, ctypes.byref(B), ctypes.byref(h))
        self.handle = h
        self.sched = {"density": default_schedule(), "stage": default_schedule()}
        self.graph = None
        self.graph_stream = None
        self.steps = 0
        stream = self._stream()
        _abi.call("sph_batch_bin", self.handle, stream, 2)
        _abi.call("sph_batch_pack", self.handle, stream, 2)

    # ------------------------------------------------------------------
    def _stream(self):
        return require_gpu(self.fpos4)

    def close(self):
        if self.graph is not None:
            _abi.call("sph_graph_destroy", self.graph)
            self.graph = None
        if self.handle:
            _abi.call("sph_batch_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _launch_step(self, stream):
        """The launch sequence of one RK2 step of every scene."""
        h = self.handle
        sd = _sched_struct(self.sched["density"])
        ss = _sched_struct(self.sched["stage"])
        call = _abi.call
        call("sph_batch_bin", h, stream, 0)
        call("sph_batch_pack", h, stream, 0)
        call("sph_batch_density", h, stream, 0, 0, ctypes.byref(sd))
        call("sph_batch_density", h, stream, 0, 1, ctypes.byref(sd))
        call("sph_batch_stage", h, stream, 1, ctypes.byref(ss))
        call("sph_batch_pack", h, stream, 1)
        call("sph_batch_density", h, stream, 1, 0, ctypes.byref(sd))
        call("sph_batch_density", h, stream, 1, 1, ctypes.byref(sd))
        call("sph_batch_stage", h, stream, 2, ctypes.byref(ss))

    def step(self):
        """One RK2 step of every scene (replays the captured graph if any)."""
        stream = self._stream()
        if self.graph is not None:
            _abi.call("sph_graph_launch", self.graph, stream)
        else:
            self._launch_step(stream)
        self.steps += 1

    def run(self, nsteps):
        for _ in range(nsteps):
            self.step()

    def capture(self):
        """Capture one step into a HIP graph; later step() calls replay it.
        Capture needs a non-default stream: the engine makes torch's current
        stream one for the rest of its life if it is the default stream."""
        if self.graph is not None:
            _abi.call("sph_graph_destroy", self.graph)
            self.graph = None
        cur = torch.cuda.current_stream(self.device)
        if cur.cuda_stream == 0:
            raise RuntimeError("capture() needs a non-default current stream: run it inside "
                               "`with torch.cuda.stream(torch.cuda.Stream()):`")
        torch.cuda.synchronize(self.device)
        stream = cur.cuda_stream
        _abi.call("sph_graph_begin", stream)
        try:
            self._launch_step(stream)
        except Exception:
            g = _abi._vp()
            _abi.load().sph_graph_end(stream, ctypes.byref(g))
            if g:
                _abi.load().sph_graph_destroy(g)
            raise
        g = _abi._vp()
        _abi.call("sph_graph_end", stream, ctypes.byref(g))
        self.graph = g
        return g

    # ------------------------------------------------------------------
    def time_phase(self, phase, sched, reps=3):
        """Average launch time (ms) of one phase under a schedule: 'density'
        (fluid destinations, stage-1 layout) or 'stage' (stage 1).  Both
        launches are idempotent: they rewrite rho/p/cs and X1/k1 from fb."""
        stream = self._stream()
        s = _sched_struct(sched)
        if phase == "density":
            fn = lambda: _abi.call("sph_batch_density", self.handle, stream, 0, 0,  # noqa: E731
                                   ctypes.byref(s))
        else:
            fn = lambda: _abi.call("sph_batch_stage", self.handle, stream, 1,  # noqa: E731
                                   ctypes.byref(s))
        fn()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize(self.device)
        return e0.elapsed_time(e1) / reps

    def autotune(self, candidates=None, reps=3):
        """Pick the fastest schedule of the density and stage phases by
        timing candidate schedules on this batch.  Schedules never change
        results.  Returns {phase: (schedule, ms, table)}."""
        if candidates is None:
            candidates = [default_schedule()]
            for order in (0, 1):
                for bpc in (2, 4, 8):
                    candidates.append({"mode": 1, "blocks_per_cu": bpc, "grab": 1,
                                       "order": order})
                    for grab in (1, 4):
                        candidates.append({"mode": 2, "blocks_per_cu": bpc, "grab": grab,
                                           "order": order})
        # the phases read the stage-1 layout: make one exist
        stream = self._stream()
        _abi.call("sph_batch_bin", self.handle, stream, 0)
        _abi.call("sph_batch_pack", self.handle, stream, 0)
        sd = _sched_struct(default_schedule())
        _abi.call("sph_batch_density", self.handle, stream, 0, 0, ctypes.byref(sd))
        _abi.call("sph_batch_density", self.handle, stream, 0, 1, ctypes.byref(sd))
        # fa -> fb is a copy of the state under a permutation: the step that
        # follows re-bins fa again, so the timing launches change nothing
        out = {}
        for phase in ("density", "stage"):
            table = [(c, self.time_phase(phase, c, reps)) for c in candidates]
            best = min(table, key=lambda t: t[1])
            self.sched[phase] = dict(best[0])
            out[phase] = (dict(best[0]), best[1], table)
        if self.graph is not None:
            self.capture()
        return out

    # ------------------------------------------------------------------
    def scene_state(self, k, fields=("x", "y", "u", "v", "rho", "p")):
        """Fluid fields of scene k in its user order (numpy, host)."""
        a, b = int(self.f_off[k]), int(self.f_off[k + 1])
        uid = self.fa["uid"][a:b].cpu().numpy()
        out = {}
        for f in fields:
            v = self.fa[f][a:b].cpu().numpy()
            r = np.empty_like(v)
            r[uid] = v
            out[f] = r
        return out

    def boundary_state(self, k, fields=("rho", "p")):
        a, b = int(self.b_off[k]), int(self.b_off[k + 1])
        uid = self.bnd["uid"][a:b].cpu().numpy()
        out = {}
        for f in fields:
            v = self.bnd[f][a:b].cpu().numpy()
            r = np.empty_like(v)
            r[uid] = v
            out[f] = r
        return out


def pair_estimate(scene):
    """Pairs per step estimate of a scene (lattice density)."""
    dx = scene.info.get("dx", scene.max_h)
    nbr = math.pi * (2.0 * scene.max_h / dx) ** 2 if scene.dim == 2 else \
        4.0 / 3.0 * math.pi * (2.0 * scene.max_h / dx) ** 3
    return nbr * (scene.nf + scene.nb)
