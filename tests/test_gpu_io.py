"""Dumps of a GPU run in the reference's formats: Solver.dump_output's
per-array .npz (solver.py:510-538) and PrintNeighborInformation's neighbour
and cell dumps (post_step_functions.py:9-73), whose neighbour lists equal
the oracle's sets on the dumped state."""
import os
import pickle

import numpy as np
import pytest

import pysph_amd.base.api as base
import pysph_amd.solver.api as solver
from oracle import oracle as O
from pysph_amd.solver import generators as gen
from pysph_amd.solver.post_step_functions import load_neighbors
from tests.util import to_oracle

pytestmark = pytest.mark.gpu


def _run(tmp_path, fmt):
    f, b, info = gen.dam_break_2d(width=0.2, height=0.3, dx=0.01, jitter=0.1, seed=3,
                                  precision="double")
    particles = base.Particles([f, b])
    s = solver.wcsph_solver(2, co=info["co"], ro=info["rho0"], alpha=0.1, gravity=info["g"],
                            xsph_eps=0.5)
    s.setup_integrator(particles)
    s.set_time_step(info["dt"])
    s.set_final_time(2.5 * info["dt"])
    s.set_output_directory(str(tmp_path))
    s.set_print_freq(1)
    s.post_step_functions.append(solver.PrintNeighborInformation(path=str(tmp_path), count=1,
                                                                 fmt=fmt))
    s.solve()
    return s, particles, f, b


def test_solution_dumps_and_npz_neighbour_dump(tmp_path):
    s, particles, f, b = _run(tmp_path, "npz")
    assert s.count == 3
    t = str(s.t)
    # Solver.dump_output: solution_<array>_<t>.npz with the print properties
    for pa in (f, b):
        d = solver.utils.load(os.path.join(str(tmp_path), "solution_%s_%s.npz" % (pa.name, t)))
        assert {"dt", "cell_size", "np", "x", "y", "u", "v", "m", "h", "p", "rho"} <= set(d)
        assert int(d["np"]) == pa.get_number_of_particles()
        for c in ("x", "y", "u", "v", "rho", "p"):
            assert np.array_equal(d[c], pa.get(c)), c
    cs = particles.cell_manager.cell_size
    fo, bo = to_oracle(f), to_oracle(b)
    ora = {"fluid": fo, "boundary": bo}
    locs = list(particles.nnps_manager.particle_locator_cache.values())
    assert len(locs) >= 3
    for loc in locs:
        fname = os.path.join(str(tmp_path), "neighbors_0_%s_%s_%s.npz"
                             % (loc.dest.name, loc.source.name, t))
        rec, rows = load_neighbors(fname)
        assert float(rec["cell_size"]) == cs and float(rec["radius_scale"]) == 2.0
        dst, src = ora[loc.dest.name], ora[loc.source.name]
        off_o, nb_o = O.nnps_csr(dst, src, radius_scale=2.0)
        didx = loc.dest.get("idx")
        sidx = loc.source.get("idx")
        assert np.array_equal(rec["idx"], didx)
        for i, row in enumerate(rows):
            assert sorted(row.tolist()) == sorted(sidx[nb_o[off_o[i]:off_o[i + 1]]].tolist())
        x, y, z = loc.dest.get("x", "y", "z")
        assert np.array_equal(rec["cid"][:, 0], np.floor(x / cs).astype(np.int32))
        assert np.array_equal(rec["cid"][:, 1], np.floor(y / cs).astype(np.int32))
    cells = solver.utils.load(os.path.join(str(tmp_path), "cells_0_%s.npz" % t))
    assert np.array_equal(cells["fluid_x"], f.get("x"))
    assert np.array_equal(cells["boundary_idx"], b.get("idx"))


def test_pickle_neighbour_dump_matches_reference_layout(tmp_path):
    s, particles, f, b = _run(tmp_path, "pickle")
    t = str(s.t)
    # a file this test's own run wrote
    with open(os.path.join(str(tmp_path), "neighbors_0_fluid_%s" % t), "rb") as fh:
        d = pickle.load(fh)
    assert set(d) == set(f.get("idx").tolist())
    fo = to_oracle(f)
    cs = particles.cell_manager.cell_size
    # the last locator of the fluid destination in the cache wins, as in
    # the reference's one-file-per-destination layout
    last = [loc for loc in particles.nnps_manager.particle_locator_cache.values()
            if loc.dest is f][-1]
    off_o, nb_o = O.nnps_csr(fo, to_oracle(last.source), radius_scale=2.0)
    sidx = last.source.get("idx")
    x, y, z = f.get("x", "y", "z")
    for i, pid in enumerate(f.get("idx").tolist()):
        assert sorted(d[pid]["neighbors"].tolist()) == \
            sorted(sidx[nb_o[off_o[i]:off_o[i + 1]]].tolist())
        assert d[pid]["cid"] == (int(np.floor(x[i] / cs)), int(np.floor(y[i] / cs)),
                                 int(np.floor(z[i] / cs)))
    with open(os.path.join(str(tmp_path), "cells_0_%s" % t), "rb") as fh:
        cells = pickle.load(fh)
    n = sum(len(c["positions"]["fluid"]["idx"]) for c in cells.values())
    assert n == f.get_number_of_particles()
    for cid, c in cells.items():
        assert c["cid"] == cid
        assert c["centroid"] == tuple((k + 0.5) * cs for k in cid)
