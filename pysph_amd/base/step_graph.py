"""HIP graphs of a time step that re-bins particle arrays.

A binning permutes every property of an array into the buffer the previous
binning left (ParticleArray.apply_permutation alternates two buffers per
property) and writes its cells into the other of two cell arrays, so the
buffers a step reads come back after a few steps (two for an RK2 step;
six for an Euler step, which also swaps the integrated fields with its
output buffers, a cycle of three, while the other fields alternate).  StepGraphs captures consecutive steps without running them
until the buffers are back where they started, one graph per step of that
cycle, and replays them in turn: each replay is one step.  The
Python-side state a step leaves (which buffer holds each property, the
cell arrays, the grid, dirty flags, the owner's own buffers) is recorded
at capture and restored after each replay.

A replay needs the state it was captured from; any other state (new
buffers after particles were added, another grid, another dirty pattern)
runs the step eagerly and captures again after `warmup` eager steps, whose
first binnings allocate the second buffers.  While capturing, the cell
manager keeps its grid and starts no host read-back (CellManager.capturing);
the owner reads bounds and jump counts outside the graphs.  A step that
cannot be captured (a host synchronisation, an allocation) switches graphs
off for the owner and records why in `error`.

A replay costs the host one call, so without a limit the host would queue
a whole run before the GPU finished its first steps, and what the owner
reads back between steps (the jump count that raises the reference's jump
tolerance error, the bounds that size the grid) would arrive only at the
end.  The host therefore waits until the GPU has finished the replay
`max_lead` steps back before it queues the next one.
"""
import collections
import ctypes

import torch

from .. import _abi
from .._device import stream_handle


class StepGraphs:
    def __init__(self, arrays, cell_manager, get_extra=None, set_extra=None, key_extra=None,
                 warmup=2, max_period=6, max_lead=8):
        self.arrays = list(arrays)
        self.cm = cell_manager
        # the owner's state: get_extra() is restored by set_extra(); its
        # key_extra() (hashable: the identities of buffers the step reads)
        # is part of the state's key
        self.get_extra = get_extra
        self.set_extra = set_extra
        self.key_extra = key_extra
        self.warmup = int(warmup)
        self.max_period = int(max_period)
        self.max_lead = int(max_lead)
        self._events = collections.deque()
        self.enabled = True
        self.replays = 0
        self.captures = 0
        self.error = None
        self._g = None
        self._warm = 0
        self._stream = None

    # ---------------------------------------------------------------- state
    def state(self):
        arrays = []
        for pa in self.arrays:
            d = pa.__dict__
            arrays.append((pa, dict(pa._props), dict(d.get("_spare", {})), d.get("_order"),
                           d.get("_cell"), d.get("_cell_start"), d.get("_grid_key"),
                           d.get("_bin_version", 0), d.get("is_dirty"), d.get("_refcell")))
        extra = self.get_extra() if self.get_extra else None
        xkey = self.key_extra() if self.key_extra else None
        return {"grid": self.cm.grid, "cm_dirty": self.cm.is_dirty, "arrays": arrays,
                "extra": extra, "xkey": xkey}

    def set_state(self, st):
        cm = self.cm
        cm.grid, cm.is_dirty = st["grid"], st["cm_dirty"]
        for pa, props, spare, order, cell, cstart, key, ver, dirty, ref in st["arrays"]:
            pa._props.clear()
            pa._props.update(props)
            d = pa.__dict__
            d["_spare"] = dict(spare)
            d["_order"], d["_cell"], d["_cell_start"] = order, cell, cstart
            d["_grid_key"], d["_bin_version"], d["is_dirty"] = key, ver, dirty
            d["_refcell"] = ref
        if self.set_extra:
            self.set_extra(st["extra"])

    @staticmethod
    def key(st):
        """Identity of a state: the buffers, grid and dirty flags a captured
        step reads (versions and counters excluded)."""
        g = st["grid"]
        key = [None if g is None else (g.cell_size, tuple(g.lo), tuple(g.dims))]
        for pa, props, spare, order, cell, cstart, gkey, _, dirty, ref in st["arrays"]:
            key.append((id(pa), pa.get_number_of_particles(),
                        tuple(sorted((k, id(t)) for k, t in props.items())),
                        tuple(sorted((k, id(t)) for k, t in spare.items())),
                        id(order), id(cell), id(cstart), gkey, bool(dirty), id(ref)))
        key.append(st["xkey"])
        return tuple(key)

    # ---------------------------------------------------------------- graphs
    def drop(self):
        """Destroy the captured graphs; the next steps run eagerly, then the
        step is captured again."""
        if self._g is not None:
            lib = _abi.load()
            for g in self._g["exec"]:
                lib.sph_graph_destroy(g)
        self._g = None
        self._warm = 0

    def __del__(self):
        try:
            self.drop()
        except Exception:
            pass

    def _capture(self, fn, tag):
        st0 = self.state()
        dev = self.arrays[0].device
        if self._stream is None:
            self._stream = torch.cuda.Stream(device=dev)
        gs = self._stream
        gs.wait_stream(torch.cuda.current_stream(dev))
        lib = _abi.load()
        graphs, states, err = [], [st0], None
        self.cm.capturing = True
        key0 = self.key(st0)
        try:
            with torch.cuda.stream(gs):
                hs = gs.cuda_stream
                while len(graphs) < self.max_period:
                    _abi.call("sph_graph_begin", hs)
                    try:
                        fn()
                    except BaseException:
                        g = _abi._vp()
                        lib.sph_graph_end(hs, ctypes.byref(g))
                        if g:
                            lib.sph_graph_destroy(g)
                        raise
                    g = _abi._vp()
                    _abi.call("sph_graph_end", hs, ctypes.byref(g))
                    graphs.append(g)
                    states.append(self.state())
                    if self.key(states[-1]) == key0:
                        break
        except Exception as e:  # not capturable: run eagerly from now on
            err = e
        finally:
            self.cm.capturing = False
            self.set_state(st0)
        if err is None and self.key(states[-1]) != key0:
            err = RuntimeError("the step's buffers do not come back within %d steps"
                               % self.max_period)
        if err is not None:
            for g in graphs:
                lib.sph_graph_destroy(g)
            self.error = "%s: %s" % (type(err).__name__, err)
            self.enabled = False
            return False
        self._g = {"tag": tag, "exec": graphs, "states": states,
                   "keys": [self.key(s) for s in states[:-1]]}
        self.captures += 1
        return True

    def step(self, fn, tag=None):
        """Run one step: replay a graph of fn() captured from the current
        state, else call fn() eagerly (capturing after `warmup` eager calls).
        `tag` holds what the graph bakes in besides buffers (time step,
        switches); a new tag captures again.  Returns True for a replay."""
        if not self.enabled or not torch.cuda.is_available():
            self.drop()
            fn()
            return False
        if self._g is not None and self._g["tag"] != tag:
            self.drop()
        if self._g is None:
            if self._warm < self.warmup:
                self._warm += 1
                fn()
                return False
            if not self._capture(fn, tag):
                fn()
                return False
        k = self.key(self.state())
        keys = self._g["keys"]
        if k not in keys:
            self.drop()
            fn()
            return False
        i = keys.index(k)
        _abi.call("sph_graph_launch", self._g["exec"][i], stream_handle())
        self.set_state(self._g["states"][i + 1])
        self.replays += 1
        ev = torch.cuda.Event()
        ev.record()
        self._events.append(ev)
        while len(self._events) > self.max_lead:
            self._events.popleft().synchronize()
        return True
