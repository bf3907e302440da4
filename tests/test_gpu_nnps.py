"""Neighbour search on the GPU against the oracle and the reference's own
NNPS test expectations: bit-exact sorted index sets."""
import numpy as np
import pytest

import pysph_amd.base.api as base
from oracle import oracle as O
from pysph_amd.solver import generators as gen
from tests.util import csr_sorted_rows, to_oracle

pytestmark = pytest.mark.gpu


def _gpu_csr(src, dst, radius_scale=2.0, min_cell_size=-1.0):
    arrays = [src] if src is dst else [src, dst]
    cm = base.CellManager(arrays_to_bin=arrays, min_cell_size=min_cell_size)
    nm = base.NNPSManager(cm)
    loc = nm.get_neighbor_particle_locator(src, dst, radius_scale)
    return loc.get_csr(), cm


def _check_rows(off_g, nb_g, off_o, nb_o):
    assert np.array_equal(np.diff(off_g), np.diff(off_o)), "neighbour counts differ"
    for a, b in zip(csr_sorted_rows(off_g, nb_g), csr_sorted_rows(off_o, nb_o)):
        assert np.array_equal(a, b)


def _random_array(n, seed, precision, box=1.0, h=0.05, dim=3):
    rng = np.random.Generator(np.random.PCG64(seed))
    dt = np.float32 if precision == "single" else np.float64
    pts = rng.random((3, n)) * box
    if dim < 3:
        pts[2] = 0
    if dim < 2:
        pts[1] = 0
    pts = pts.astype(dt).astype(np.float64)
    hh = (h * (0.8 + 0.4 * rng.random(n))).astype(dt).astype(np.float64)
    return base.get_particle_array(cl_precision=precision, name="a%d" % seed, x=pts[0],
                                   y=pts[1], z=pts[2], h=hh)


@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("dim", [1, 2, 3])
def test_random_sets_match_oracle(precision, dim):
    n = {1: 3000, 2: 6000, 3: 12000}[dim]
    pa = _random_array(n, 7 + dim, precision, h={1: 0.002, 2: 0.01, 3: 0.03}[dim], dim=dim)
    (off, nb), _ = _gpu_csr(pa, pa)
    ref = to_oracle(pa)
    off_o, nb_o = O.nnps_csr(ref)
    _check_rows(off, nb, off_o, nb_o)
    assert off[-1] > 5 * n  # the sets are not trivial


@pytest.mark.parametrize("precision", ["single", "double"])
def test_two_arrays_variable_h(precision):
    src = _random_array(5000, 1, precision, h=0.04)
    dst = _random_array(3000, 2, precision, h=0.06)
    (off, nb), _ = _gpu_csr(src, dst)
    off_o, nb_o = O.nnps_csr(to_oracle(dst), to_oracle(src),
                             cell_size=2.0 * max(src.get("h").max(), dst.get("h").max()))
    _check_rows(off, nb, off_o, nb_o)


def test_exact_cutoff_ties_are_excluded():
    # lattice spacing 0.5 with h = 0.5: neighbours at distance exactly 2h = 1
    # fail the strict test (nnps.pyx:337), diagonal ones inside pass
    g = np.arange(8) * 0.5
    x, y = np.meshgrid(g, g, indexing="ij")
    pa = base.get_particle_array(x=x.ravel(), y=y.ravel(), h=np.full(64, 0.5))
    (off, nb), _ = _gpu_csr(pa, pa)
    off_o, nb_o = O.nnps_csr(to_oracle(pa))
    _check_rows(off, nb, off_o, nb_o)
    # interior particle (3,3): |d| < 1 -> 9 points (itself, 4 axial at 0.5, 4 diagonal)
    k = 3 * 8 + 3
    assert off[k + 1] - off[k] == 9


def test_clustered_3d_single():
    pa = gen.cube_array(20000, kind="clustered", seed=3, neighbours=40)
    (off, nb), _ = _gpu_csr(pa, pa)
    off_o, nb_o = O.nnps_csr(to_oracle(pa))
    _check_rows(off, nb, off_o, nb_o)


def test_self_is_last_and_rows_user_order():
    pa = _random_array(2000, 11, "double", h=0.05)
    cm = base.CellManager(arrays_to_bin=[pa])
    loc = base.NNPSManager(cm).get_neighbor_particle_locator(pa, pa, 2.0)
    for i in (0, 17, 1999):
        row = loc.py_get_nearest_particles(i).get_npy_array()
        assert row[-1] == i
        ex = loc.py_get_nearest_particles(i, exclude_self=True).get_npy_array()
        assert i not in ex and len(ex) == len(row) - 1
    # binning reorders the device layout; user-order data is unchanged
    x0 = pa.get("x")
    cm.update()
    assert np.array_equal(pa.get("x"), x0)


def test_reference_dataset_1_fixed_dest():
    """base/tests/test_nnps.py:221-270 on common_data.py:8-43."""
    p1 = base.ParticleArray(name="parr1", x={"data": [0.25, 0.8, 0.5, 0.8, 0.2, 0.5, 1.5, 1.5]},
                            y={"data": [0.25, 0.1, 0.5, 0.8, 0.9, 1.5, 0.5, 1.5]},
                            z={"data": [0.0] * 8}, h={"data": [1.0] * 8})
    p2 = base.ParticleArray(name="parr2", x={"data": [0.2, 1.2, 1.5, 0.4]},
                            y={"data": [0.0] * 4}, z={"data": [1.6, 1.5, -0.5, 0.4]},
                            h={"data": [1.0] * 4})
    cm = base.CellManager(arrays_to_bin=[p1, p2], min_cell_size=1.0, max_cell_size=2.0)
    nl = base.FixedDestNbrParticleLocator(p1, p2, 1.0, cm, "h")
    nl4 = base.FixedDestNbrParticleLocator(p1, p2, 4.0, cm, "h")
    nl05 = base.FixedDestNbrParticleLocator(p1, p2, 0.5, cm, "h")
    out = nl05.py_get_nearest_particles(3)
    assert out.length == 1 and out[0] == 0
    out = nl.py_get_nearest_particles(3)
    assert sorted(out.get_npy_array().tolist()) == [0, 1, 2, 3]
    out = nl4.py_get_nearest_particles(2)
    assert sorted(out.get_npy_array().tolist()) == list(range(8))


def test_reference_dataset_1_point_queries():
    """base/tests/test_nnps.py:94-177."""
    p1 = base.ParticleArray(name="parr1", x={"data": [0.25, 0.8, 0.5, 0.8, 0.2, 0.5, 1.5, 1.5]},
                            y={"data": [0.25, 0.1, 0.5, 0.8, 0.9, 1.5, 0.5, 1.5]},
                            z={"data": [0.0] * 8}, h={"data": [1.0] * 8})
    p2 = base.ParticleArray(name="parr2", x={"data": [0.2, 1.2, 1.5, 0.4]},
                            y={"data": [0.0] * 4}, z={"data": [1.6, 1.5, -0.5, 0.4]},
                            h={"data": [1.0] * 4})
    cm = base.CellManager(arrays_to_bin=[p1, p2], min_cell_size=1.0, max_cell_size=2.0)
    l1 = base.NbrParticleLocatorBase(p1, cm)
    out = base.LongArray()
    l1.py_get_nearest_particles_to_point(base.Point(0.4, 0.0, 0.4), 0.5, out)
    assert out.length == 1 and out[0] == 0
    out.reset()
    l1.py_get_nearest_particles_to_point(base.Point(0.4, 0.0, 0.4), 1.0, out)
    assert sorted(out.get_npy_array().tolist()) == [0, 1, 2, 3]
    out.reset()
    l1.py_get_nearest_particles_to_point(base.Point(1.5, 0.0, -0.5), 4.0, out)
    assert sorted(out.get_npy_array().tolist()) == list(range(8))
    l2 = base.NbrParticleLocatorBase(p2, cm)
    c = base.Point(1, 1, 1)
    expect = [(0.1, []), (1.0, []), (1.136, [1]), (1.358, [1, 3]),
              (1.4142135623730951, [0, 1, 3]), (1.88, [0, 1, 2, 3])]
    for r, ids in expect:
        out.reset()
        l2.py_get_nearest_particles_to_point(c, r, out)
        assert sorted(out.get_npy_array().tolist()) == ids, r
    out.reset()
    l2.py_get_nearest_particles_to_point(c, 1.88, out, 3)
    assert sorted(out.get_npy_array().tolist()) == [0, 1, 2]


def test_empty_and_single():
    one = base.get_particle_array(x=np.array([0.3]), y=np.array([0.2]), h=np.array([0.1]))
    (off, nb), _ = _gpu_csr(one, one)
    assert list(off) == [0, 1] and list(nb) == [0]
    empty = base.get_particle_array(x=np.zeros(0), name="e")
    (off, nb), _ = _gpu_csr(one, empty)
    assert list(off) == [0] and len(nb) == 0


def test_dam_break_fluid_boundary_sets():
    f, b, info = gen.dam_break_3d(dx=1.0 / 30, column=(10, 10, 20), tank=(24, 10, 20),
                                  seed=5)
    particles = base.Particles([f, b])
    nm = particles.nnps_manager
    ref_f, ref_b = to_oracle(f), to_oracle(b)
    cs = particles.cell_manager.cell_size
    for src, dst, rs, rd in ((f, f, ref_f, ref_f), (b, f, ref_b, ref_f), (f, b, ref_f, ref_b)):
        off, nb = nm.get_neighbor_particle_locator(src, dst, 2.0).get_csr()
        if src is dst:
            off_o, nb_o = O.nnps_csr(rd, cell_size=cs)
        else:
            off_o, nb_o = O.nnps_csr(rd, rs, cell_size=cs)
        _check_rows(off, nb, off_o, nb_o)


def test_jump_tolerance_raises():
    """A particle moved by two cells between binnings raises, as the
    reference's CellManager does (cell.pyx:1422-1437)."""
    from pysph_amd._abi import JumpToleranceError
    pa = _random_array(500, 4, "double", h=0.05)
    cm = base.CellManager(arrays_to_bin=[pa])
    cm.update()
    x = pa.get("x")
    x[0] = x[0] + 0.35  # 3.5 cells of 0.1
    x[0] = min(x[0], 0.99)
    if abs(x[0] - pa.get("x")[0]) < 0.2:
        x[0] = pa.get("x")[0] - 0.35
    pa.set(x=x)
    # the count is made on the device and read back without a wait: the
    # first binning after it reached the host raises, or verify_jumps()
    cm.update()
    with pytest.raises(JumpToleranceError):
        cm.verify_jumps()
    cm.verify_jumps()  # reported once


def test_jump_tolerance_raises_at_next_binning():
    import torch
    from pysph_amd._abi import JumpToleranceError
    pa = _random_array(500, 4, "double", h=0.05)
    cm = base.CellManager(arrays_to_bin=[pa])
    cm.update()
    x = pa.get("x")
    x[:10] = np.where(x[:10] > 0.5, x[:10] - 0.35, x[:10] + 0.35)
    pa.set(x=x)
    cm.update()
    torch.cuda.synchronize()
    pa.set_dirty(True)
    with pytest.raises(JumpToleranceError):
        cm.update()


def test_jump_tolerance_counts_reference_cells_on_finer_bins():
    """With bins half a cell wide the jump check still compares the
    reference's cells (cell.pyx:1422-1437): a move of two bins into the next
    cell passes, a move of two cells raises."""
    from pysph_amd._abi import JumpToleranceError
    pa = _random_array(500, 4, "double", h=0.04)
    cm = base.CellManager(arrays_to_bin=[pa], min_cell_size=0.1)
    cm.set_cells_per_radius(2)
    assert cm.cell_size == 0.1 and cm.bin_size == 0.05
    x = pa.get("x")
    x[0] = 0.3005  # cell 3, bin 6
    pa.set(x=x)
    cm.update()
    cm.verify_jumps()
    x[0] = 0.4495  # cell 4, bin 8
    pa.set(x=x)
    cm.update()
    cm.verify_jumps()
    x[0] = 0.6005  # cell 6, bin 12
    pa.set(x=x)
    cm.update()
    with pytest.raises(JumpToleranceError):
        cm.verify_jumps()


def test_stale_grid_keeps_neighbour_sets_exact():
    """The bounds that size the grid are read back asynchronously, so a
    binning may use a grid older than the positions: particles beyond it are
    clamped into its border cells and the neighbour sets stay exact."""
    pa = _random_array(4000, 9, "double", h=0.05)
    cm = base.CellManager(arrays_to_bin=[pa])
    cm.check_jumps = False
    nm = base.NNPSManager(cm)
    cm.update()
    g0 = cm.grid
    pa.set(x=pa.get("x") + 0.45, y=pa.get("y") * 1.3 - 0.15)
    cm.update()
    assert cm.grid is g0  # bounds of the moved particles not read yet
    off, nb = nm.get_neighbor_particle_locator(pa, pa, 2.0).get_csr()
    assert cm.grid is g0
    off_o, nb_o = O.nnps_csr(to_oracle(pa), cell_size=cm.cell_size)
    _check_rows(off, nb, off_o, nb_o)
    assert off[-1] > 5 * pa.get_number_of_particles()


def test_capped_fill_matches_fill_and_flags_overflow():
    import torch
    from pysph_amd import _abi
    pa = _random_array(4000, 21, "single", h=0.02, dim=2)
    cm = base.CellManager(arrays_to_bin=[pa])
    cm.update()
    n = pa.get_number_of_particles()
    st = torch.cuda.current_stream().cuda_stream
    v = pa.view()
    dev = pa.device
    counts = torch.empty(n, dtype=torch.int64, device=dev)
    off = torch.empty(n + 1, dtype=torch.int64, device=dev)
    ids = torch.arange(n, dtype=torch.int64, device=dev)
    _abi.call("sph_neighbor_counts", st, cm.grid, v, v, 2.0, counts.data_ptr())
    total = _abi.ctypes.c_int64(0)
    _abi.call("sph_scan_counts", st, n, counts.data_ptr(), off.data_ptr(),
              _abi.ctypes.byref(total))
    total = int(total.value)
    ref = torch.empty(total, dtype=torch.int64, device=dev)
    needed = _abi.ctypes.c_int64(0)
    _abi.call("sph_neighbor_fill", st, cm.grid, v, v, 2.0, ids.data_ptr(), off.data_ptr(),
              ref.data_ptr(), total, total, _abi.ctypes.byref(needed))
    # ample capacity: the same CSR, no flag
    got = torch.full((total + 64,), -7, dtype=torch.int64, device=dev)
    flag = torch.zeros(1, dtype=torch.int32, device=dev)
    _abi.call("sph_neighbor_fill_capped", st, cm.grid, v, v, 2.0, ids.data_ptr(),
              off.data_ptr(), got.data_ptr(), total + 64, flag.data_ptr())
    assert int(flag.item()) == 0
    assert torch.equal(got[:total], ref)
    assert torch.all(got[total:] == -7)
    # too small: flagged, nothing written at or past the capacity
    cap = total // 2
    got = torch.full((total,), -7, dtype=torch.int64, device=dev)
    _abi.call("sph_neighbor_fill_capped", st, cm.grid, v, v, 2.0, ids.data_ptr(),
              off.data_ptr(), got.data_ptr(), cap, flag.data_ptr())
    assert int(flag.item()) == 1
    assert torch.all(got[cap:] == -7)
    o = off.cpu().numpy()
    full_rows = np.nonzero(o[1:] <= cap)[0]
    assert torch.equal(got[:int(o[full_rows[-1] + 1])], ref[:int(o[full_rows[-1] + 1])])


@pytest.mark.skipif(not __import__("os").path.exists(O.REF_NNPS2D),
                    reason="reference C++ neighbour search not built")
@pytest.mark.parametrize("precision", ["single", "double"])
def test_2d_sets_match_reference_cpp_nnps(precision):
    """HIP sets against the reference's bench/nnps_c++ search run on the
    same (precision-rounded) points, uniform h, radius 2h."""
    rng = np.random.Generator(np.random.PCG64(23))
    dt = np.float32 if precision == "single" else np.float64
    x = rng.uniform(0, 1, 8000).astype(dt).astype(np.float64)
    y = rng.uniform(0, 1, 8000).astype(dt).astype(np.float64)
    h = 0.0125
    pa = base.get_particle_array(cl_precision=precision, name="r2d", x=x, y=y,
                                 h=np.full(len(x), h))
    (off, nb), _ = _gpu_csr(pa, pa)
    ref = O.ref_nnps2d(x, y, 2.0 * h)
    assert np.array_equal(np.diff(off), [len(r) for r in ref]), "neighbour counts differ"
    for a, b in zip(csr_sorted_rows(off, nb), ref):
        assert np.array_equal(a, b)
    assert off[-1] > 5 * len(x)
