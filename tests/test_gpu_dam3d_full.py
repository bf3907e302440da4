"""The full benchmarked 3D dam break (81x81x162 fluid lattice = 1,062,882
fluid particles plus 352,458 boundary particles, FP32) after 20 RK2 steps
in bench.py's kernel configuration: sampled destinations, queried against
the whole state on the GPU, have the FP64 oracle's neighbour sets
(FixedDest, nnps.pyx:271-356 and :525-620; the style of the reference's
bench/nnps_c++/test.cpp:22-60) and its summation density
(density_funcs.pyx:38-84) on that state.
"""
import numpy as np
import pytest
import torch

import bench
import pysph_amd.base.api as base
from oracle import oracle as O
from pysph_amd.base import nnps
from pysph_amd.sph import _ops
from tests.util import csr_sorted_rows, max_rel

pytestmark = pytest.mark.gpu


def _probe(pa, pick):
    x, y, z, h = (pa.get(c).astype(np.float64)[pick] for c in ("x", "y", "z", "h"))
    return base.get_particle_array(cl_precision="single", name="probe_" + pa.name, x=x, y=y,
                                   z=z, h=h)


def _orc(pa, pick=None):
    d = {c: pa.get(c).astype(np.float64) for c in ("x", "y", "z", "h", "m")}
    if pick is not None:
        d = {c: v[pick] for c, v in d.items()}
    return O.make_array(**d)


def test_dam3d_full_size_sampled_after_20_steps():
    s, particles, f, b, info, col, tank = bench.build_case(scale=1.0)
    nf, nb = f.get_number_of_particles(), b.get_number_of_particles()
    assert (nf, nb) == (1062882, 352458)
    for _ in range(20):
        s.integrator.integrate(info["dt"])
    torch.cuda.synchronize()
    cm = particles.cell_manager
    cm.verify_jumps()
    cm.update()
    rng = np.random.Generator(np.random.PCG64(20))
    srcs = [(f, _orc(f)), (b, _orc(b))]
    cs = 2.0 * info["h"]
    prm = _ops.make_params(3, radius_scale=2.0)
    for dst, nsample in ((f, 1024), (b, 512)):
        pick = np.sort(rng.choice(dst.get_number_of_particles(), nsample, replace=False))
        probe = _probe(dst, pick)
        cm.bin_array(probe, cm.grid)
        dst_o = _orc(dst, pick)
        checked = 0
        for src, src_o in srcs:
            off, nbr = nnps.neighbor_csr(cm, src, probe, 2.0)
            off_o, nb_o = O.nnps_csr(dst_o, src_o, cell_size=cs)
            assert np.array_equal(off, off_o), (dst.name, src.name)
            for a, c in zip(csr_sorted_rows(off, nbr), csr_sorted_rows(off_o, nb_o)):
                assert np.array_equal(a, c), (dst.name, src.name)
            checked += int(off[-1])
            if src is dst:
                # each sampled particle finds itself
                for k in (0, nsample // 2, nsample - 1):
                    assert pick[k] in nbr[off[k]:off[k + 1]]
        # summation density of the sample from both arrays on the packed path
        rho = torch.empty(nsample, dtype=torch.float32, device=f.device)
        _ops.density(cm, prm, probe, [f, b], rho)
        torch.cuda.synchronize()
        order = probe._order_np()
        got = rho.cpu().numpy().astype(np.float64)
        if order is not None:
            tmp = np.empty_like(got)
            tmp[order] = got
            got = tmp
        rho_o = O.sphrho(3, dst_o, [s_o for _, s_o in srcs], cell_size=cs)
        err = max_rel(got, rho_o)
        print("%s: %d sampled destinations, %d pairs bit-exact, density max rel err %.2e"
              % (dst.name, nsample, checked, err))
        assert err < 1e-5
