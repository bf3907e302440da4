"""HIP graphs of a time step that re-bins particle arrays.

A binning permutes every property of an array into the buffer the previous
binning left (ParticleArray.apply_permutation alternates two buffers per
property) and writes its cells into the other of two cell arrays, so the
buffers a step reads return after two steps.  StepGraphs captures two
consecutive steps, one from each buffer parity, without running them, and
replays them alternately: each replay is one step.  The Python-side state
a step leaves (which buffer holds each property, the cell arrays, the grid,
dirty flags) is recorded at capture and restored after each replay.

A replay needs the state it was captured from; any other state (new
buffers after particles were added, another grid, another dirty pattern)
runs the step eagerly and captures again after `warmup` eager steps, whose
first binnings allocate the second buffers.  While capturing, the cell
manager keeps its grid and starts no host read-back (CellManager.capturing);
the owner reads bounds and jump counts outside the graphs.  A step that
cannot be captured (a host synchronisation, an allocation) switches graphs
off for the owner and records why in `error`.
"""
import ctypes

import torch

from .. import _abi
from .._device import stream_handle


class StepGraphs:
    def __init__(self, arrays, cell_manager, get_extra=None, set_extra=None, warmup=2):
        self.arrays = list(arrays)
        self.cm = cell_manager
        self.get_extra = get_extra
        self.set_extra = set_extra
        self.warmup = int(warmup)
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
                           d.get("_bin_version", 0), d.get("is_dirty")))
        extra = self.get_extra() if self.get_extra else None
        return {"grid": self.cm.grid, "cm_dirty": self.cm.is_dirty, "arrays": arrays,
                "extra": extra}

    def set_state(self, st):
        cm = self.cm
        cm.grid, cm.is_dirty = st["grid"], st["cm_dirty"]
        for pa, props, spare, order, cell, cstart, key, ver, dirty in st["arrays"]:
            pa._props.clear()
            pa._props.update(props)
            d = pa.__dict__
            d["_spare"] = dict(spare)
            d["_order"], d["_cell"], d["_cell_start"] = order, cell, cstart
            d["_grid_key"], d["_bin_version"], d["is_dirty"] = key, ver, dirty
        if self.set_extra:
            self.set_extra(st["extra"])

    @staticmethod
    def key(st):
        """Identity of a state: the buffers, grid and dirty flags a captured
        step reads (versions and counters excluded)."""
        g = st["grid"]
        key = [None if g is None else (g.cell_size, tuple(g.lo), tuple(g.dims))]
        for pa, props, spare, order, cell, cstart, gkey, _, dirty in st["arrays"]:
            key.append((id(pa), pa.get_number_of_particles(),
                        tuple(sorted((k, id(t)) for k, t in props.items())),
                        tuple(sorted((k, id(t)) for k, t in spare.items())),
                        id(order), id(cell), id(cstart), gkey, bool(dirty)))
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
        try:
            with torch.cuda.stream(gs):
                hs = gs.cuda_stream
                for _ in range(2):
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
        except Exception as e:  # not capturable: run eagerly from now on
            err = e
        finally:
            self.cm.capturing = False
            self.set_state(st0)
        if err is None and self.key(states[2]) != self.key(st0):
            err = RuntimeError("the step's buffers do not return after two binnings")
        if err is not None:
            for g in graphs:
                lib.sph_graph_destroy(g)
            self.error = "%s: %s" % (type(err).__name__, err)
            self.enabled = False
            return False
        self._g = {"tag": tag, "exec": graphs, "states": states,
                   "keys": [self.key(states[0]), self.key(states[1])]}
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
        return True
