"""Stage 2 of the fused RK2 step keeps the stage-1 binning and widens its
search by the predictor's displacement, recorded on the device by the
stage-1 kernel.  The neighbour sets found that way must be the exact sets at
the predicted positions X1, which the reference gets by rebuilding its
caches after the predictor (integrator.py:621-645, nnps.pyx:271-356).

The hard case is a source that moved diagonally away from the cell it was
binned in: its straight-line move is sqrt(3) times its largest per-axis
move, and a pair just inside the cut-off is only found when the culled
stencil grows by the straight-line bound.
"""
import math

import numpy as np
import pytest
import torch

import pysph_amd.base.api as base
import pysph_amd.solver.api as solver
from oracle import oracle as O
from pysph_amd import _abi
from pysph_amd.sph import _ops
from tests.util import to_oracle

pytestmark = pytest.mark.gpu

H = 0.02
CS = H          # two cells per search radius
R = 2.0 * H     # search radius


def _probe_set(seed=0, shift=0.4):
    """A jittered lattice plus probe pairs: source j just inside the upper
    corner of its cell, destination i at distance R (1 - 1e-3) from j along
    the diagonal (+1, +1, +1).  Every particle moves by shift * CS along
    each axis in stage 1."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sp = 0.6 * H
    g = np.arange(10, 26) * sp
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    pts += rng.uniform(-0.05 * sp, 0.05 * sp, pts.shape)
    t = R * (1.0 - 1e-3) / math.sqrt(3.0)
    probes = []
    for c in [(8, 8, 8), (11, 9, 10), (13, 12, 9), (9, 13, 12), (12, 10, 13)]:
        j = np.array(c, dtype=np.float64) * CS - 1e-4 * CS
        probes += [j, j + t]
        # and the same pair along a face diagonal (+1, +1, 0)
        j2 = j + np.array([0.0, 0.0, 0.5 * CS])
        probes += [j2, j2 + np.array([1.0, 1.0, 0.0]) * R * (1.0 - 1e-3) / math.sqrt(2.0)]
    pts = np.concatenate([pts, np.array(probes)])
    pts = pts.astype(np.float32).astype(np.float64)
    n = len(pts)
    dt = 1e-4
    vel = np.float32(shift * CS / dt)
    pa = base.get_particle_array(
        cl_precision="single", name="fluid", x=pts[:, 0], y=pts[:, 1], z=pts[:, 2],
        u=np.full(n, vel), v=np.full(n, vel), w=np.full(n, vel), h=np.full(n, H),
        m=np.full(n, 1000.0 * sp ** 3), rho=np.full(n, 1000.0))
    return pa, dt


def _stage2_sets(pa, dt):
    """Stage 1 of the fused plan, then a list-building density pass at X1
    over the kept binning with the recorded displacement; returns the X1
    positions (user order), the recorded displacement and the lists as CSR
    rows of user indices."""
    particles = base.Particles([pa], update_particles=False)
    particles.cell_manager.set_fixed_domain((0.0, 0.0, 0.0), (0.6, 0.6, 0.6))
    s = solver.wcsph_solver(3, co=10.0, ro=1000.0, alpha=0.1, xsph_eps=0.5, with_solid=False)
    s.setup_integrator(particles)
    plan = s.integrator.plan
    assert plan is not None and plan.use_packed
    plan.set_cells_per_radius(2)
    assert abs(plan.cm.bin_size - CS) < 1e-6 * CS  # h is stored in FP32
    plan.cm.update()
    b = plan._buffers()
    plan._pack_all()
    ov = plan._override()
    stream = torch.cuda.current_stream().cuda_stream
    _abi.call("sph_fill_i32", stream, b["disp"].data_ptr(), 1, 0)
    plan._density(override=ov)
    plan._stage(1, dt, b, override=ov, disp_out=b["disp"])
    s1 = plan._s1_fields(b)
    plan._pack(pa, fields=s1, force=True)
    ov = plan._override(s1)
    n = pa.get_number_of_particles()
    maxp = 256
    idx = torch.empty(n * maxp, dtype=torch.int32, device=pa.device)
    cnt = torch.empty(n * _abi.SPH_MAX_SRC, dtype=torch.int32, device=pa.device)
    nl = _abi.SphNlist()
    nl.idx, nl.count, nl.max_per, nl.mode, nl.skin = idx.data_ptr(), cnt.data_ptr(), maxp, 0, 0.0
    rho = torch.empty(n, dtype=torch.float32, device=pa.device)
    _ops.density(plan.cm, plan.prm, pa, [pa], rho, disp=b["disp"], override=ov, nl=nl)
    torch.cuda.synchronize()
    order = pa._order.cpu().numpy() if pa._order is not None else np.arange(n)
    counts = cnt[:n].cpu().numpy()
    assert counts.max() <= maxp
    lists = idx.cpu().numpy()
    rows = [None] * n
    for slot in range(n):
        js = lists[np.arange(counts[slot]) * n + slot]
        rows[order[slot]] = np.sort(order[js])
    x1 = {}
    for c, nm in enumerate("xyz"):
        t = b["s1"][c].cpu().numpy()
        v = np.empty_like(t)
        v[order] = t
        x1[nm] = v.astype(np.float64)
    return x1, float(b["disp"].item()), rows


def test_stage2_kept_binning_finds_diagonal_pairs():
    pa, dt = _probe_set()
    x0 = {c: pa.get(c).astype(np.float64) for c in "xyz"}
    x1, disp, rows = _stage2_sets(pa, dt)
    move = np.sqrt(sum((x1[c] - x0[c]) ** 2 for c in "xyz"))
    # the recorded displacement bounds every straight-line move
    assert disp >= move.max()
    assert disp <= move.max() * (1 + 1e-5)
    ref = to_oracle(pa)
    for c in "xyz":
        ref[c] = x1[c]
    off, nbr = O.nnps_csr(ref, None, cell_size=2.0 * H)
    n = len(x1["x"])
    assert int(off[-1]) == sum(len(r) for r in rows)
    for i in range(n):
        assert np.array_equal(rows[i], np.sort(nbr[off[i]:off[i + 1]])), i
    # the probe pairs are among them (each destination sees its source)
    npr = 20
    for k in range(n - npr, n, 2):
        assert k in set(rows[k + 1].tolist())


def test_stage2_kept_binning_matches_rebinned_step():
    """A whole RK2 step with the kept binning and with the reference's
    re-binning after the predictor agree to FP32 summation order."""
    out = []
    for rebin in (False, True):
        pa, dt = _probe_set(seed=1, shift=0.3)
        particles = base.Particles([pa], update_particles=False)
        particles.cell_manager.set_fixed_domain((0.0, 0.0, 0.0), (0.6, 0.6, 0.6))
        s = solver.wcsph_solver(3, co=10.0, ro=1000.0, alpha=0.1, xsph_eps=0.5,
                                with_solid=False)
        s.setup_integrator(particles)
        plan = s.integrator.plan
        plan.set_cells_per_radius(2)
        plan.stage2_rebin = rebin
        s.integrator.integrate(dt)
        out.append({c: pa.get(c).astype(np.float64) for c in ("x", "y", "z", "u", "v", "w",
                                                              "rho")})
    for c in ("x", "y", "z"):
        assert np.max(np.abs(out[0][c] - out[1][c])) < 1e-6, c
    for c in ("u", "v", "w"):
        scale = np.max(np.abs(out[1][c]))
        assert np.max(np.abs(out[0][c] - out[1][c])) < 1e-5 * scale, c
    assert np.max(np.abs(out[0]["rho"] / out[1]["rho"] - 1.0)) < 1e-6
