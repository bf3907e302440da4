"""Launchers of the GPU operations behind the SPH function classes.

Each launcher takes ParticleArrays (binned by their CellManager) and device
output tensors, fills the ABI structs and calls libsph_mi355x.so on the
current HIP stream.
"""
import ctypes

import torch

from .. import _abi
from .._device import dtype_code, ptr, require_gpu

# Pair arithmetic of the neighbour passes for FP32 storage.  All-FP32
# (SPH_PAIR_NATIVE) keeps the momentum sum of a 3D column at rest within
# 2e-7 normwise of the FP64 oracle (the contract is 1e-5); SPH_PAIR_F64
# evaluates each pair in FP64 (2e-8) at about 1.3x the cost.
DEFAULT_PAIR_PREC = _abi.SPH_PAIR_NATIVE


def make_params(dim, radius_scale=2.0, co=1.0, ro=1000.0, gamma=7.0, alpha=1.0, beta=1.0,
                eta=0.1, eps=0.5, g=(0.0, 0.0, 0.0), has_gravity=False, has_xsph=False,
                pair_prec=None, symmetric=False, exclude_self=0, lds_rows=False):
    """sph_params; `symmetric` selects the variable-h pair sets of
    VarHNbrParticleLocator (nnps.pyx:654-749); bit s of `exclude_self`
    drops the self pair of source s (the destination) from the density;
    `lds_rows` stages the density scan's stencil rows in LDS (opt-in)."""
    p = _abi.SphParams()
    p.dim = int(dim)
    p.pair_prec = DEFAULT_PAIR_PREC if pair_prec is None else int(pair_prec)
    p.radius_scale = float(radius_scale)
    p.co, p.ro, p.gamma = float(co), float(ro), float(gamma)
    p.alpha, p.beta, p.eta, p.eps = float(alpha), float(beta), float(eta), float(eps)
    for i in range(3):
        p.g[i] = float(g[i])
    p.has_gravity = 1 if has_gravity else 0
    p.has_xsph = 1 if has_xsph else 0
    p.symmetric = 1 if symmetric else 0
    p.exclude_self = int(exclude_self)
    p.lds_rows = 1 if lds_rows else 0
    return p


def calc_params(calc, **kw):
    """make_params for a calc: its kernel's dimension and radius, and the
    symmetric pair sets when its Particles were made with variable_h."""
    sym = getattr(calc, "symmetric", None)
    if sym is None:
        sym = getattr(getattr(calc, "particles", None), "variable_h", False)
    return make_params(calc.kernel.dim, radius_scale=calc.kernel.radius(), symmetric=bool(sym),
                       **kw)


def views(arrays, override=None):
    """A C array of sph_view for the given ParticleArrays.  `override` maps
    an array to a dict of field -> tensor replacing its own fields."""
    override = override or {}
    arr = (_abi.SphView * max(len(arrays), 1))()
    for i, pa in enumerate(arrays):
        v = pa.view()
        for f, t in override.get(id(pa), {}).items():
            setattr(v, f, ptr(t))
        arr[i] = v
    return arr


def search_radius(cm, radius_scale):
    """Largest search radius of a call: radius_scale * max h."""
    return float(radius_scale) * float(cm.max_h if cm.max_h > 0 else cm.cell_size / 2.0)


def check_same_dtype(arrays):
    dts = {pa.dtype for pa in arrays}
    if len(dts) != 1:
        raise TypeError("all arrays of one operation must share a dtype")


# FP32 density passes called without packed inputs pack (x,y,z,h) and
# (u,v,w,m) into per-array float4 scratch first and take the fast path.
AUTO_PACK = True


def packed_override(arrays, override=None):
    """`override` extended with freshly packed pos4/vel4 copies of every
    FP32 array that has none (the copies follow the current slot order)."""
    ov = dict(override or {})
    for pa in arrays:
        if pa.dtype != torch.float32 or "pos4" in ov.get(id(pa), {}):
            continue
        n = pa.get_number_of_particles()
        sc = pa.__dict__.get("_pack_scratch")
        if sc is None or sc["pos4"].shape[0] < max(n, 1):
            sc = {f: torch.empty((max(n, 1), 4), dtype=torch.float32, device=pa.device)
                  for f in ("pos4", "vel4")}
            sc["pos3"] = torch.empty(max(n, 1) * 3, dtype=torch.float32, device=pa.device)
            sc["pm4"] = torch.empty((max(n, 1), 4), dtype=torch.float32, device=pa.device)
            sc["hrange"] = torch.empty(2, dtype=torch.int32, device=pa.device)
            sc["mrange"] = torch.empty(2, dtype=torch.int32, device=pa.device)
            pa.__dict__["_pack_scratch"] = sc
        # pack what the pass would read: the array's fields as overridden
        v = pa.view()
        for f, t in ov.get(id(pa), {}).items():
            setattr(v, f, ptr(t))
        _abi.call("sph_pack", require_gpu(sc["pos4"]), v, sc["pos4"].data_ptr(),
                  sc["vel4"].data_ptr(), sc["pos3"].data_ptr(), sc["pm4"].data_ptr(),
                  sc["hrange"].data_ptr(), sc["mrange"].data_ptr())
        d = dict(ov.get(id(pa), {}))
        d.update(sc)
        ov[id(pa)] = d
    return ov


def density(cm, prm, dst, srcs, rho, p=None, cs=None, disp=None, override=None, nl=None,
            cache=None):
    """SPHRho over srcs into rho (+ Tait into p, cs); `nl` (SphNlist)
    receives the accepted neighbours for a following momentum pass;
    `cache` (SphSrcCache) keeps the partial sums of one static source."""
    check_same_dtype([dst] + list(srcs))
    stream = require_gpu(rho)
    if AUTO_PACK and dst.dtype == torch.float32 and nl is None:
        override = packed_override([dst] + [s for s in srcs if s is not dst], override)
    vd = views([dst], override)
    vs = views(srcs, override)
    if cache is not None:
        _abi.call("sph_density_cached", stream, cm.grid, prm, vd, len(srcs), vs,
                  search_radius(cm, prm.radius_scale), ptr(disp), ptr(rho), ptr(p), ptr(cs),
                  cache)
        return
    _abi.call("sph_density", stream, cm.grid, prm, vd, len(srcs), vs,
              search_radius(cm, prm.radius_scale), ptr(disp), ptr(rho), ptr(p), ptr(cs), nl)


def density_multi(cm, prm, dsts, srcs, rhos, ps, css, disp=None, override=None):
    """SPHRho (+ Tait where p, cs are given) of up to two destination arrays
    over the same sources, in one launch when both take the packed FP32
    scan (sph_density_multi)."""
    check_same_dtype(list(dsts) + list(srcs))
    stream = require_gpu(*rhos)
    vd = views(dsts, override)
    vs = views(srcs, override)
    _abi.call("sph_density_multi", stream, cm.grid, prm, len(dsts), vd, len(srcs), vs,
              search_radius(cm, prm.radius_scale), ptr(disp),
              _abi.ptr_array([ptr(t) for t in rhos]), _abi.ptr_array([ptr(t) for t in ps]),
              _abi.ptr_array([ptr(t) for t in css]))


def density_rate(cm, prm, dst, srcs, rate, disp=None, override=None):
    """SPHDensityRate over srcs into rate (density_funcs.pyx:89-158)."""
    check_same_dtype([dst] + list(srcs))
    stream = require_gpu(rate)
    vd = views([dst], override)
    vs = views(srcs, override)
    _abi.call("sph_density_rate", stream, cm.grid, prm, vd, len(srcs), vs,
              search_radius(cm, prm.radius_scale), ptr(disp), ptr(rate))


def tait(prm, dst, p, cs, rho=None):
    stream = require_gpu(p)
    v = dst.view()
    if rho is not None:
        v.rho = rho.data_ptr()
    _abi.call("sph_tait", stream, prm, v, ptr(p), ptr(cs))


def momentum_xsph(cm, prm, dst, srcs, flags, acc=None, dv=None, disp=None, override=None,
                  nl=None):
    """MomentumEquation (sources flagged SPH_SRC_MOMENTUM) and XSPHCorrection
    (flagged SPH_SRC_XSPH) into acc[3] / dv[3]."""
    check_same_dtype([dst] + list(srcs))
    stream = require_gpu(dst.get_carray("x"))
    vd = views([dst], override)
    vs = views(srcs, override)
    fl = (ctypes.c_uint32 * max(len(srcs), 1))(*flags)
    a = _abi.ptr_array([ptr(t) for t in acc]) if acc is not None else None
    d = _abi.ptr_array([ptr(t) for t in dv]) if dv is not None else None
    _abi.call("sph_momentum_xsph", stream, cm.grid, prm, vd, len(srcs), vs, fl,
              search_radius(cm, prm.radius_scale), ptr(disp), a, d, nl)


def update(dtype, n, mode, s, ops, tag=None):
    """Batched array updates (see sph_update in the header)."""
    if not ops:
        return
    stream = require_gpu(*[o[0] for o in ops])
    arr = (_abi.SphUpdateOp * len(ops))()
    for i, (out, a, b, c) in enumerate(ops):
        arr[i].out = ptr(out)
        arr[i].a = ptr(a)
        arr[i].b = ptr(b)
        arr[i].c = ptr(c)
    _abi.call("sph_update", stream, dtype_code(dtype), int(n), int(mode), float(s), len(ops),
              arr, ptr(tag))


def conservation(prm, arrays):
    """FP64 mass, momentum, kinetic, potential, elastic and total energy
    summed over the arrays (host numpy[8])."""
    out = None
    for pa in arrays:
        if pa.get_number_of_particles() == 0:
            continue
        stream = require_gpu(pa.get_carray("x"))
        if out is None:
            out = torch.zeros(8, dtype=torch.float64, device=pa.device)
        _abi.call("sph_conservation_f64", stream, prm, pa.view(), out.data_ptr())
    return out.cpu().numpy() if out is not None else torch.zeros(8).numpy()
