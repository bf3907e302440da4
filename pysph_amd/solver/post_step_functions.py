"""Post-step functions of the solver (reference: solver/post_step_functions.py).

PrintNeighborInformation writes, every `count` steps, the neighbour lists of
every cached locator and the particles of every cell
(post_step_functions.py:9-73).  Formats:

* ``fmt="npz"`` (default), one file per locator,
  ``neighbors_<rank>_<dest>_<source>_<t>.npz``:
    idx        int64 [n]     the destination's idx property, row order
    offsets    int64 [n+1]   CSR row offsets
    neighbors  int64 [total] the source's idx of each neighbour
    cid        int32 [n,3]   cell (floor(x/cell_size), ...) of each row
    cell_size, radius_scale, time (0-d)
  and ``cells_<rank>_<t>.npz`` with ``<array>_x/_y/_z/_idx/_cid`` per array.
  Read back with `load_neighbors` / `solver.utils.load`.
* ``fmt="pickle"``: the reference's dictionaries in its file names
  (``neighbors_<rank>_<dest>_<t>``: {idx: {'neighbors': int64 array, 'cid':
  (i, j, k)}}, ``cells_<rank>_<t>``: {(i, j, k): {'cid', 'centroid',
  'positions': {array: {'x', 'y', 'z', 'idx'}}}}), built from plain Python
  and numpy types so that any reader can load them.

The reference maps neighbour indices through the destination's idx and
keys its file by destination alone; here they go through the source's idx
(the same thing when source is destination) and each locator has its own
npz file.
"""
import os
import pickle

import numpy as np


def cell_ids(x, y, z, cell_size):
    """floor(coordinate / cell_size) per axis (cell.pyx:42-82), int32 [n,3]."""
    c = np.stack([np.floor(np.asarray(a, dtype=np.float64) / cell_size) for a in (x, y, z)], 1)
    return c.astype(np.int32)


def neighbor_record(loc, cell_size):
    """The dump arrays of one locator (rows in the destination's user order)."""
    off, nb = loc.get_csr()
    dest, src = loc.dest, loc.source
    every = {"only_real_particles": False}  # one row per destination particle
    didx = dest.get("idx", **every).astype(np.int64) if dest.has_array("idx") else \
        np.arange(dest.get_number_of_particles(), dtype=np.int64)
    sidx = src.get("idx", **every).astype(np.int64) if src.has_array("idx") else \
        np.arange(src.get_number_of_particles(), dtype=np.int64)
    x, y, z = dest.get("x", "y", "z", **every)
    return {"idx": didx, "offsets": np.asarray(off, dtype=np.int64),
            "neighbors": sidx[np.asarray(nb, dtype=np.int64)],
            "cid": cell_ids(x, y, z, cell_size)}


def load_neighbors(fname):
    """(record dict, list of neighbour arrays per row) of an npz dump."""
    with np.load(fname, allow_pickle=False) as f:
        rec = {k: f[k] for k in f.files}
    off, nb = rec["offsets"], rec["neighbors"]
    rows = [nb[off[i]:off[i + 1]] for i in range(len(off) - 1)]
    return rec, rows


class PrintNeighborInformation:
    def __init__(self, rank=0, path=None, count=10, fmt="npz"):
        if fmt not in ("npz", "pickle"):
            raise ValueError("fmt must be 'npz' or 'pickle'")
        self.rank = rank
        self.count = count
        self.path = path if path else "."
        self.fmt = fmt

    def eval(self, solver, count):
        if count % self.count != 0:
            return []
        particles = solver.particles
        time = solver.t
        cm = particles.cell_manager
        cell_size = cm.cell_size
        base = os.path.join(self.path, "neighbors_" + str(self.rank))
        written = []
        by_dest = {}
        for loc in particles.nnps_manager.particle_locator_cache.values():
            rec = neighbor_record(loc, cell_size)
            if self.fmt == "npz":
                fname = "%s_%s_%s_%s.npz" % (base, loc.dest.name, loc.source.name, str(time))
                np.savez(fname, cell_size=cell_size, radius_scale=loc.radius_scale, time=time,
                         **rec)
                written.append(fname)
            else:
                d = by_dest.setdefault(loc.dest.name, {})
                off, nb = rec["offsets"], rec["neighbors"]
                for i, pid in enumerate(rec["idx"].tolist()):
                    d[pid] = {"neighbors": nb[off[i]:off[i + 1]].copy(),
                              "cid": tuple(int(c) for c in rec["cid"][i])}
        for name, d in by_dest.items():
            fname = base + "_" + name + "_" + str(time)
            with open(fname, "wb") as f:
                pickle.dump(d, f, protocol=2)
            written.append(fname)
        cells = os.path.join(self.path, "cells_" + str(self.rank)) + "_" + str(time)
        written.append(cm.get_particle_representation(cells, fmt=self.fmt))
        return written
