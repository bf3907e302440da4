"""ctypes front end of the FP64 CPU oracle (``liborc.so``).

TEST INFRASTRUCTURE ONLY.  This module and the library it loads are the
checker of the HIP path: ``tests/``, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` may import it, nothing under ``pysph_amd``
does.  The arithmetic lives in ``sph_oracle.c``; its header lists the
reference file:line each function restates.

Arrays are dicts of float64 numpy arrays with the reference's property names
(x, y, z, u, v, w, h, m, rho, p, cs) plus optional ``tag`` (int64) and
``type`` (0 Fluid, 1 Solid; particle_types.py).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liborc.so")
_FIELDS = ("x", "y", "z", "u", "v", "w", "h", "m", "rho", "p", "cs")

_dp = ctypes.POINTER(ctypes.c_double)
_lp = ctypes.POINTER(ctypes.c_int64)


class OrcArr(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64)] + [(f, _dp) for f in _FIELDS] + [
        ("tag", _lp), ("type", ctypes.c_int32)]


class OrcNbrs(ctypes.Structure):
    _fields_ = [("n_dst", ctypes.c_int64), ("off", _lp), ("nbr", _lp)]


class OrcScheme(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int32), ("integrator", ctypes.c_int32),
        ("dt", ctypes.c_double), ("co", ctypes.c_double), ("ro", ctypes.c_double),
        ("gamma", ctypes.c_double), ("alpha", ctypes.c_double),
        ("beta", ctypes.c_double), ("eta", ctypes.c_double),
        ("has_gravity", ctypes.c_int32), ("g", ctypes.c_double * 3),
        ("has_xsph", ctypes.c_int32), ("eps", ctypes.c_double),
        ("cell_size", ctypes.c_double), ("radius_scale", ctypes.c_double)]


def build():
    """Compile liborc.so from sph_oracle.c (gcc, no GPU needed)."""
    subprocess.check_call(["make", "-s", "-C", _HERE, "liborc.so"])


_lib = None
_REF_DIR = "/root/reference/bench/nnps_c++"
REF_NNPS2D = os.path.join(_HERE, "_ref", "nnps2d")


def build_ref():
    """Compile the reference's 2D C++ neighbour search into _ref/ when the
    reference sources are present (returns whether the binary exists)."""
    if os.path.isdir(_REF_DIR):
        subprocess.check_call(["make", "-s", "-C", _HERE, "ref"])
    return os.path.exists(REF_NNPS2D)


def ref_nnps2d(x, y, radius):
    """Neighbour sets of every point of (x, y) within `radius` from the
    reference's bench/nnps_c++ search: list of sorted int64 arrays."""
    import tempfile
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            f.write(np.int64(len(x)).tobytes())
            f.write(np.float64(radius).tobytes())
            f.write(x.tobytes())
            f.write(y.tobytes())
        subprocess.check_call([REF_NNPS2D, fin, fout])
        raw = np.fromfile(fout, dtype=np.int64)
    rows, k = [], 0
    for _ in range(len(x)):
        c = int(raw[k])
        rows.append(np.sort(raw[k + 1:k + 1 + c]))
        k += 1 + c
    return rows


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_cubic_w.restype = ctypes.c_double
        L.orc_cubic_w.argtypes = [ctypes.c_double] * 7 + [ctypes.c_int]
        L.orc_cubic_grad.argtypes = [ctypes.c_double] * 7 + [ctypes.c_int, _dp]
        L.orc_real_to_int.restype = ctypes.c_int32
        L.orc_real_to_int.argtypes = [ctypes.c_double, ctypes.c_double]
        L.orc_nnps_csr.restype = ctypes.c_int64
        L.orc_nnps_csr.argtypes = [ctypes.POINTER(OrcArr), ctypes.POINTER(OrcArr),
                                   ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                   _lp, _lp, ctypes.c_int64, ctypes.c_int]
        L.orc_point_query.restype = ctypes.c_int64
        L.orc_point_query.argtypes = [ctypes.POINTER(OrcArr), ctypes.c_double,
                                      ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_double,
                                      ctypes.c_int64, _lp, ctypes.c_int64]
        L.orc_nbrs_build.restype = ctypes.c_int
        L.orc_nbrs_build.argtypes = [ctypes.POINTER(OrcNbrs), ctypes.POINTER(OrcArr),
                                     ctypes.POINTER(OrcArr), ctypes.c_double,
                                     ctypes.c_double]
        L.orc_nbrs_build_varh.restype = ctypes.c_int
        L.orc_nbrs_build_varh.argtypes = L.orc_nbrs_build.argtypes
        L.orc_nbrs_free.argtypes = [ctypes.POINTER(OrcNbrs)]
        pa = ctypes.POINTER(ctypes.POINTER(OrcArr))
        pn = ctypes.POINTER(ctypes.POINTER(OrcNbrs))
        L.orc_sphrho.argtypes = [ctypes.c_int, ctypes.POINTER(OrcArr), ctypes.c_int,
                                 pa, pn, _dp]
        L.orc_tait.argtypes = [ctypes.POINTER(OrcArr), _dp, ctypes.c_double,
                               ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.orc_momentum.argtypes = [ctypes.c_int, ctypes.POINTER(OrcArr), ctypes.c_int,
                                   pa, pn, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_double, _dp, _dp, _dp]
        L.orc_xsph.argtypes = [ctypes.c_int, ctypes.POINTER(OrcArr), ctypes.c_int,
                               pa, pn, ctypes.c_double, _dp, _dp, _dp]
        L.orc_density_rate.argtypes = [ctypes.c_int, ctypes.POINTER(OrcArr), ctypes.c_int,
                                       pa, pn, _dp]
        L.orc_scheme_step.restype = ctypes.c_int
        L.orc_scheme_step.argtypes = [ctypes.POINTER(OrcScheme), ctypes.c_int, pa]
        L.orc_scheme_rhs.restype = ctypes.c_int
        L.orc_scheme_rhs.argtypes = [ctypes.POINTER(OrcScheme), ctypes.c_int, pa,
                                     ctypes.c_int] + [_dp] * 6
        L.orc_conservation.argtypes = [ctypes.c_int, pa, _dp, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_double, _dp]
        L.orc_set_threads.restype = ctypes.c_int
        L.orc_set_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def set_threads(n):
    """1 = the single-threaded reference-structure mode; <=0 = all cores."""
    return lib().orc_set_threads(int(n))


def _ptr(a, t=_dp):
    return a.ctypes.data_as(t)


def make_array(n=None, ptype=0, **props):
    """A float64 property dict with the reference's defaults
    (get_particle_array, particles.py:320-322)."""
    defaults = dict(x=0.0, y=0.0, z=0.0, u=0.0, v=0.0, w=0.0, m=1.0, h=1.0,
                    p=0.0, e=0.0, rho=1.0, cs=0.0)
    if n is None:
        n = len(next(iter(props.values())))
    out = {}
    for f in _FIELDS:
        if f in props:
            out[f] = np.ascontiguousarray(props[f], dtype=np.float64).copy()
        else:
            out[f] = np.full(n, defaults[f], dtype=np.float64)
    if "tag" in props and props["tag"] is not None:
        out["tag"] = np.ascontiguousarray(props["tag"], dtype=np.int64)
    out["type"] = int(ptype)
    return out


class _Bound:
    """Keeps the numpy buffers alive while the C struct points at them."""

    def __init__(self, arr):
        for f in _FIELDS:
            a = arr[f]
            assert a.dtype == np.float64 and a.flags.c_contiguous, f
        self.arr = arr
        self.s = OrcArr()
        self.s.n = len(arr["x"])
        for f in _FIELDS:
            setattr(self.s, f, _ptr(arr[f]))
        tag = arr.get("tag")
        self.s.tag = _ptr(tag, _lp) if tag is not None else _lp()
        self.s.type = int(arr.get("type", 0))


def cubic_w(pa, pb, h, dim):
    return lib().orc_cubic_w(pa[0], pa[1], pa[2], pb[0], pb[1], pb[2], h, dim)


def cubic_grad(pa, pb, h, dim):
    g = np.zeros(3)
    lib().orc_cubic_grad(pa[0], pa[1], pa[2], pb[0], pb[1], pb[2], h, dim, _ptr(g))
    return g


def real_to_int(val, step):
    return lib().orc_real_to_int(val, step)


def nnps_csr(dst, src=None, cell_size=None, radius_scale=2.0, varh=False):
    """FixedDest neighbour lists of dst from src (self last when src is dst),
    or with varh the symmetric variable-h lists of VarHNbrParticleLocator.
    Returns (offsets int64[n+1], nbrs int64[total])."""
    same = src is None or src is dst
    if cell_size is None:
        hs = [dst["h"]] + ([] if same else [src["h"]])
        cell_size = radius_scale * max(float(np.max(h)) if len(h) else 0.0 for h in hs)
        cell_size = cell_size or 1.0
    bd = _Bound(dst)
    bs = bd if same else _Bound(src)
    n = bd.s.n
    off = np.zeros(n + 1, dtype=np.int64)
    cap = 0
    nb = np.zeros(1, dtype=np.int64)
    total = lib().orc_nnps_csr(ctypes.byref(bd.s), ctypes.byref(bs.s), int(same),
                               cell_size, radius_scale, _ptr(off, _lp), _ptr(nb, _lp), cap,
                               int(varh))
    if total < 0:
        raise MemoryError("oracle neighbour build failed")
    nb = np.zeros(max(total, 1), dtype=np.int64)
    total = lib().orc_nnps_csr(ctypes.byref(bd.s), ctypes.byref(bs.s), int(same),
                               cell_size, radius_scale, _ptr(off, _lp), _ptr(nb, _lp),
                               total, int(varh))
    return off, nb[:total]


def ghost_cells(arrs, period, pmin, cell_size, dist):
    """CellManager.create_ghost_cells (cell.pyx:1277-1376) on the cells of
    `arrs` (cell = floor(x / cell_size) per axis, cell.pyx:42-82): a cell
    whose lower face lies at or below the domain's lower face, or within
    `dist` of it, is copied one period up; one whose upper face lies at or
    above the upper face, or within `dist` of it, one period down.  The
    reference does this along x only; along several periodic axes the
    copies combine (edges and corners).  Returns (real cell ids, dict ghost
    cell id -> list of (shift vector, source cell id))."""
    cells = set()
    for a in arrs:
        for c in zip(*[np.floor(np.asarray(a[k]) / cell_size).astype(np.int64)
                       for k in ("x", "y", "z")]):
            cells.add(tuple(int(v) for v in c))
    ghosts = {}
    for cid in sorted(cells):
        opts = []
        for d in range(3):
            o = [0]
            if period[d] > 0:
                centre = (cid[d] + 0.5) * cell_size
                lower, upper = centre - 0.5 * cell_size, centre + 0.5 * cell_size
                pmax = pmin[d] + period[d]
                if lower <= pmin[d] or (lower - pmin[d]) < dist:
                    o.append(1)
                if upper >= pmax or (pmax - upper) < dist:
                    o.append(-1)
            opts.append(o)
        for sx in opts[0]:
            for sy in opts[1]:
                for sz in opts[2]:
                    if sx == 0 and sy == 0 and sz == 0:
                        continue
                    sh = (sx, sy, sz)
                    centre = [(cid[d] + 0.5) * cell_size + sh[d] * period[d] for d in range(3)]
                    gid = tuple(int(np.floor(c / cell_size)) for c in centre)
                    ghosts.setdefault(gid, []).append((sh, cid))
    return cells, ghosts


def with_ghosts(arr, period, pmin, cell_size, dist):
    """`arr` extended by the ghost copies of its particles (Cell.copy,
    cell.pyx:600-660: the copied particles' coordinates + shift * period,
    added in FP64).  Returns (extended array, origin index of each entry)."""
    n = len(arr["x"])
    cid = np.stack([np.floor(np.asarray(arr[k]) / cell_size).astype(np.int64)
                    for k in ("x", "y", "z")], 1)
    _, ghosts = ghost_cells([arr], period, pmin, cell_size, dist)
    by_cell = {}
    for i in range(n):
        by_cell.setdefault(tuple(int(v) for v in cid[i]), []).append(i)
    idx, shifts = [np.arange(n)], [np.zeros((n, 3))]
    for copies in ghosts.values():
        for sh, src_cell in copies:
            members = by_cell.get(src_cell, [])
            if members:
                idx.append(np.asarray(members))
                shifts.append(np.tile(np.asarray(sh, dtype=np.float64) *
                                      np.asarray(period, dtype=np.float64), (len(members), 1)))
    origin = np.concatenate(idx)
    sh = np.concatenate(shifts)
    out = {}
    for f in _FIELDS:
        out[f] = np.ascontiguousarray(arr[f][origin])
    for d, k in enumerate(("x", "y", "z")):
        out[k] = out[k] + sh[:, d]
    out["type"] = arr.get("type", 0)
    return out, origin


def point_query(src, cell_size, pnt, radius, exclude_index=-1):
    bs = _Bound(src)
    out = np.zeros(bs.s.n + 1, dtype=np.int64)
    k = lib().orc_point_query(ctypes.byref(bs.s), cell_size, pnt[0], pnt[1], pnt[2],
                              radius, exclude_index, _ptr(out, _lp), len(out))
    return out[:k]


class _Calc:
    """Binds dst and srcs and builds the per-(src,dst) neighbour caches."""

    def __init__(self, dst, srcs, cell_size, radius_scale, varh=False):
        self.bd = _Bound(dst)
        self.bs = [self.bd if s is dst else _Bound(s) for s in srcs]
        self.nb = [OrcNbrs() for _ in srcs]
        build = lib().orc_nbrs_build_varh if varh else lib().orc_nbrs_build
        for nb, b in zip(self.nb, self.bs):
            if build(ctypes.byref(nb), ctypes.byref(self.bd.s), ctypes.byref(b.s), cell_size,
                     radius_scale):
                raise MemoryError("oracle neighbour build failed")
        k = len(srcs)
        self.pa = (ctypes.POINTER(OrcArr) * k)(*[ctypes.pointer(b.s) for b in self.bs])
        self.pn = (ctypes.POINTER(OrcNbrs) * k)(*[ctypes.pointer(nb) for nb in self.nb])
        self.k = k

    def close(self):
        for nb in self.nb:
            lib().orc_nbrs_free(ctypes.byref(nb))


def _cell_size(arrs, radius_scale, cell_size):
    if cell_size is not None and cell_size > 0:
        return cell_size
    mh = max(float(np.max(a["h"])) if len(a["h"]) else 0.0 for a in arrs)
    return radius_scale * mh or 1.0


def sphrho(dim, dst, srcs, cell_size=None, radius_scale=2.0, varh=False):
    cs = _cell_size([dst] + list(srcs), radius_scale, cell_size)
    c = _Calc(dst, srcs, cs, radius_scale, varh)
    out = np.zeros(len(dst["x"]))
    lib().orc_sphrho(dim, ctypes.byref(c.bd.s), c.k, c.pa, c.pn, _ptr(out))
    c.close()
    return out


def tait(arr, co, ro, gamma):
    b = _Bound(arr)
    p = np.zeros(b.s.n)
    cs = np.zeros(b.s.n)
    lib().orc_tait(ctypes.byref(b.s), _ptr(arr["rho"]), co, ro, gamma, _ptr(p), _ptr(cs))
    return p, cs


def momentum(dim, dst, srcs, alpha=1.0, beta=1.0, eta=0.1, cell_size=None,
             radius_scale=2.0, varh=False):
    cs = _cell_size([dst] + list(srcs), radius_scale, cell_size)
    c = _Calc(dst, srcs, cs, radius_scale, varh)
    n = len(dst["x"])
    o = [np.zeros(n) for _ in range(3)]
    lib().orc_momentum(dim, ctypes.byref(c.bd.s), c.k, c.pa, c.pn, alpha, beta, eta,
                       _ptr(o[0]), _ptr(o[1]), _ptr(o[2]))
    c.close()
    return np.stack(o)


def xsph(dim, dst, srcs, eps=0.5, cell_size=None, radius_scale=2.0, varh=False):
    cs = _cell_size([dst] + list(srcs), radius_scale, cell_size)
    c = _Calc(dst, srcs, cs, radius_scale, varh)
    n = len(dst["x"])
    o = [np.zeros(n) for _ in range(3)]
    lib().orc_xsph(dim, ctypes.byref(c.bd.s), c.k, c.pa, c.pn, eps,
                   _ptr(o[0]), _ptr(o[1]), _ptr(o[2]))
    c.close()
    return np.stack(o)


def density_rate(dim, dst, srcs, cell_size=None, radius_scale=2.0, varh=False):
    """SPHDensityRate over srcs (density_funcs.pyx:89-158)."""
    cs = _cell_size([dst] + list(srcs), radius_scale, cell_size)
    c = _Calc(dst, srcs, cs, radius_scale, varh)
    out = np.zeros(len(dst["x"]))
    lib().orc_density_rate(dim, ctypes.byref(c.bd.s), c.k, c.pa, c.pn, _ptr(out))
    c.close()
    return out


def scheme(dim, integrator="rk2", dt=1e-5, co=1.0, ro=1000.0, gamma=7.0, alpha=1.0,
           beta=1.0, eta=0.1, gravity=None, xsph_eps=None, cell_size=0.0,
           radius_scale=2.0):
    s = OrcScheme()
    s.dim = dim
    s.integrator = 1 if integrator == "rk2" else 0
    s.dt, s.co, s.ro, s.gamma = dt, co, ro, gamma
    s.alpha, s.beta, s.eta = alpha, beta, eta
    s.has_gravity = 0 if gravity is None else 1
    g = gravity or (0.0, 0.0, 0.0)
    for i in range(3):
        s.g[i] = g[i]
    s.has_xsph = 0 if xsph_eps is None else 1
    s.eps = xsph_eps or 0.0
    s.cell_size = cell_size
    s.radius_scale = radius_scale
    return s


def _bind_all(arrs):
    bs = [_Bound(a) for a in arrs]
    pa = (ctypes.POINTER(OrcArr) * len(bs))(*[ctypes.pointer(b.s) for b in bs])
    return bs, pa


def scheme_step(sch, arrs, nsteps=1):
    """Advance the arrays in place by nsteps integrator steps."""
    bs, pa = _bind_all(arrs)
    for _ in range(nsteps):
        if lib().orc_scheme_step(ctypes.byref(sch), len(arrs), pa):
            raise RuntimeError("oracle scheme step failed")
    return arrs


def scheme_rhs(sch, arrs, fluid_index=0):
    """rho/p/cs written into every array; returns (a[3,n], dv_xsph[3,n])."""
    bs, pa = _bind_all(arrs)
    n = len(arrs[fluid_index]["x"])
    o = [np.zeros(n) for _ in range(6)]
    if lib().orc_scheme_rhs(ctypes.byref(sch), len(arrs), pa, fluid_index,
                            *[_ptr(a) for a in o]):
        raise RuntimeError("oracle rhs failed")
    return np.stack(o[:3]), np.stack(o[3:])


def conservation(arrs, g=(0.0, 0.0, 0.0), co=1.0, ro=1000.0, gamma=7.0):
    bs, pa = _bind_all(arrs)
    out = np.zeros(8)
    gg = np.asarray(g, dtype=np.float64)
    lib().orc_conservation(len(arrs), pa, _ptr(gg), co, ro, gamma, _ptr(out))
    return out
