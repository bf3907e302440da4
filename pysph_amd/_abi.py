"""ctypes binding of libsph_mi355x.so (declared in include/sph_mi355x.h).

The library is the only compute path of this package: every neighbour
search, pair sum and integrator update runs in it on the GPU.  Loading fails
loudly when it is missing; there is no CPU fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# SPH_MI355X_LIB selects another build of the library (e.g. a tuning variant)
LIB_PATH = os.environ.get("SPH_MI355X_LIB") or os.path.join(HERE, "libsph_mi355x.so")

SPH_OK = 0
SPH_EINVAL = -1
SPH_ECAPACITY = -2
SPH_EJUMP = -3
SPH_EUNSUPPORTED = -4
SPH_F32 = 0
SPH_F64 = 1
SPH_MAX_SRC = 4
SPH_SRC_MOMENTUM = 1
SPH_SRC_XSPH = 2
SPH_PAIR_NATIVE = 0
SPH_PAIR_F64 = 1
SPH_BOUNDS_WORDS = 456
SPH_SCAN_NOALLOC_MAX_COUNTS = 1 << 27

_vp = ctypes.c_void_p


class SphView(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("dtype", ctypes.c_int32),
                ("reserved", ctypes.c_int32)] + [
        (f, _vp) for f in ("x", "y", "z", "u", "v", "w", "h", "m", "rho", "p", "cs")] + [
        ("tag", _vp), ("cell", _vp), ("cell_start", _vp), ("pos4", _vp), ("vel4", _vp),
        ("aux4", _vp), ("pos3", _vp), ("pm4", _vp), ("hrange", _vp),
        ("mrange", _vp)]


class SphGrid(ctypes.Structure):
    _fields_ = [("cell_size", ctypes.c_double), ("lo", ctypes.c_int32 * 3),
                ("dims", ctypes.c_int32 * 3), ("period", ctypes.c_double * 3),
                ("pmin", ctypes.c_double * 3), ("bins_per_cell", ctypes.c_int32)]


class SphParams(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int32), ("pair_prec", ctypes.c_int32),
                ("radius_scale", ctypes.c_double), ("co", ctypes.c_double),
                ("ro", ctypes.c_double), ("gamma", ctypes.c_double),
                ("alpha", ctypes.c_double), ("beta", ctypes.c_double),
                ("eta", ctypes.c_double), ("eps", ctypes.c_double),
                ("g", ctypes.c_double * 3), ("has_gravity", ctypes.c_int32),
                ("has_xsph", ctypes.c_int32), ("symmetric", ctypes.c_int32),
                ("exclude_self", ctypes.c_int32), ("lds_rows", ctypes.c_int32)]


class SphNlist(ctypes.Structure):
    _fields_ = [("idx", _vp), ("count", _vp), ("max_per", ctypes.c_int32),
                ("mode", ctypes.c_int32), ("skin", ctypes.c_float), ("reserved", ctypes.c_int32)]


class SphSrcCache(ctypes.Structure):
    _fields_ = [("sums", _vp), ("src", ctypes.c_int32), ("mode", ctypes.c_int32)]


SPH_CACHE_FILL = 1
SPH_CACHE_USE = 2


class SphUpdateOp(ctypes.Structure):
    _fields_ = [("out", _vp), ("a", _vp), ("b", _vp), ("c", _vp)]


class SphBatchScene(ctypes.Structure):
    _fields_ = [("f0", ctypes.c_int64), ("nf", ctypes.c_int64), ("b0", ctypes.c_int64),
                ("nb", ctypes.c_int64), ("grid", SphGrid), ("prm", SphParams),
                ("rmax", ctypes.c_double), ("dt", ctypes.c_double), ("cost", ctypes.c_double)]


BATCH_FIELDS = ("x", "y", "z", "u", "v", "w", "h", "m", "rho", "p", "cs")


class SphBatchSet(ctypes.Structure):
    _fields_ = [(f, _vp) for f in BATCH_FIELDS] + [("uid", _vp)]


class SphBatchBuffers(ctypes.Structure):
    _fields_ = [("nf_total", ctypes.c_int64), ("nb_total", ctypes.c_int64),
                ("fa", SphBatchSet), ("fb", SphBatchSet), ("bnd", SphBatchSet),
                ("bnd_in", SphBatchSet),
                ("x1a", _vp * 6), ("x1b", _vp * 6), ("k1a", _vp * 6), ("k1b", _vp * 6)] + [
        (f, _vp) for f in ("fpos4", "fvel4", "faux4", "bpos4", "bvel4", "baux4", "fscene",
                           "bscene", "fperm", "fcell", "fcell_start", "bperm", "bcell",
                           "bcell_start", "ws")] + [("ws_bytes", ctypes.c_size_t)]


class SphBatchSched(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int32), ("blocks_per_cu", ctypes.c_int32),
                ("grab", ctypes.c_int32), ("order", ctypes.c_int32)]


SPH_MLP_HIDDEN = 128
SPH_MLP_NIN = 16
SPH_MLP_NE = 16
SPH_MLP_ROUNDS = 2
SPH_MLP_INFER = 1
SPH_MLP_EDGE_FP32 = 2


class SphMlpGraph(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("nedges", ctypes.c_int64), ("n_fluid", ctypes.c_int64),
                ("dim", ctypes.c_int32), ("flags", ctypes.c_int32)] + [
        (f, _vp) for f in ("pos4", "vel4", "rho", "btype", "off", "nbr")] + [
        (f, ctypes.c_double) for f in ("rho0", "co", "gamma", "u_scale", "g", "rho_scale",
                                       "alpha", "eta")]


class SphError(RuntimeError):
    """A non-zero return code of the C ABI."""

    def __init__(self, fn, rc, msg):
        super().__init__("%s failed (rc=%d): %s" % (fn, rc, msg))
        self.rc = rc


class JumpToleranceError(SphError):
    """A particle moved more than one cell width between two binnings
    (the reference raises RuntimeError, cell.pyx:1422-1437)."""


# name -> (restype, argtypes)
_P = ctypes.POINTER
_SIGS = {
    "sph_abi_version": (ctypes.c_int32, []),
    "sph_last_error": (ctypes.c_char_p, []),
    "sph_device_count": (ctypes.c_int32, []),
    "sph_bounds": (ctypes.c_int32, [_vp, _P(SphView), _vp]),
    "sph_bin_workspace_bytes": (ctypes.c_size_t, [ctypes.c_int64, ctypes.c_int64]),
    "sph_bin": (ctypes.c_int32, [_vp, _P(SphView), _P(SphGrid), _vp, _vp, _vp, _vp, _vp,
                                 _vp, ctypes.c_size_t]),
    "sph_permute": (ctypes.c_int32, [_vp, ctypes.c_int64, _vp, ctypes.c_int32, _P(_vp),
                                     _P(_vp), _P(ctypes.c_int32)]),
    "sph_neighbor_counts": (ctypes.c_int32, [_vp, _P(SphGrid), _P(SphView), _P(SphView),
                                             ctypes.c_double, _vp]),
    "sph_scan_counts": (ctypes.c_int32, [_vp, ctypes.c_int64, _vp, _vp,
                                         _P(ctypes.c_int64)]),
    "sph_scan_counts_noalloc": (ctypes.c_int32, [_vp, ctypes.c_int64, _vp, _vp,
                                                 _P(ctypes.c_int64)]),
    "sph_jump_check": (ctypes.c_int32, [_vp, _P(SphView), ctypes.c_double, _vp, _vp]),
    "sph_neighbor_fill": (ctypes.c_int32, [_vp, _P(SphGrid), _P(SphView), _P(SphView),
                                           ctypes.c_double, _vp, _vp, _vp, ctypes.c_int64,
                                           ctypes.c_int64, _P(ctypes.c_int64)]),
    "sph_density": (ctypes.c_int32, [_vp, _P(SphGrid), _P(SphParams), _P(SphView),
                                     ctypes.c_int32, _P(SphView), ctypes.c_double, _vp, _vp,
                                     _vp, _vp, _P(SphNlist)]),
    "sph_density_multi": (ctypes.c_int32, [_vp, _P(SphGrid), _P(SphParams), ctypes.c_int32,
                                           _P(SphView), ctypes.c_int32, _P(SphView),
                                           ctypes.c_double, _vp, _P(_vp), _P(_vp), _P(_vp)]),
    "sph_neighbor_fill_capped": (ctypes.c_int32, [_vp, _P(SphGrid), _P(SphView), _P(SphView),
                                                  ctypes.c_double, _vp, _vp, _vp,
                                                  ctypes.c_int64, _vp]),
    "sph_varh_neighbor_counts": (ctypes.c_int32, [_vp, _P(SphGrid), _P(SphView), _P(SphView),
                                                  ctypes.c_double, ctypes.c_double, _vp]),
    "sph_varh_neighbor_fill": (ctypes.c_int32, [_vp, _P(SphGrid), _P(SphView), _P(SphView),
                                                ctypes.c_double, ctypes.c_double, _vp, _vp,
                                                _vp, ctypes.c_int64, ctypes.c_int64,
                                                _P(ctypes.c_int64)]),
    "sph_density_rate": (ctypes.c_int32, [_vp, _P(SphGrid), _P(SphParams), _P(SphView),
                                          ctypes.c_int32, _P(SphView), ctypes.c_double, _vp,
                                          _vp]),
    "sph_density_cached": (ctypes.c_int32, [_vp, _P(SphGrid), _P(SphParams), _P(SphView),
                                            ctypes.c_int32, _P(SphView), ctypes.c_double, _vp,
                                            _vp, _vp, _vp, _P(SphSrcCache)]),
    "sph_tait": (ctypes.c_int32, [_vp, _P(SphParams), _P(SphView), _vp, _vp]),
    "sph_periodic_wrap": (ctypes.c_int32, [_vp, _P(SphGrid), ctypes.c_int32, ctypes.c_int64,
                                           _P(_vp), ctypes.c_int32, _P(_vp)]),
    "sph_momentum_xsph": (ctypes.c_int32, [_vp, _P(SphGrid), _P(SphParams), _P(SphView),
                                           ctypes.c_int32, _P(SphView), _P(ctypes.c_uint32),
                                           ctypes.c_double, _vp, _P(_vp), _P(_vp),
                                           _P(SphNlist)]),
    "sph_wcsph_stage": (ctypes.c_int32, [_vp, ctypes.c_int32, _P(SphGrid), _P(SphParams),
                                         _P(SphView), ctypes.c_int32, _P(SphView),
                                         _P(ctypes.c_uint32), ctypes.c_double, _vp,
                                         ctypes.c_double, _P(_vp), _P(_vp), _P(_vp), _vp,
                                         _P(SphNlist)]),
    "sph_update": (ctypes.c_int32, [_vp, ctypes.c_int32, ctypes.c_int64, ctypes.c_int32,
                                    ctypes.c_double, ctypes.c_int32, _P(SphUpdateOp), _vp]),
    "sph_conservation_f64": (ctypes.c_int32, [_vp, _P(SphParams), _P(SphView), _vp]),
    "sph_move_circle": (ctypes.c_int32, [_vp, _P(SphView), ctypes.c_int32, _vp, _vp]),
    "sph_fill_i32": (ctypes.c_int32, [_vp, _vp, ctypes.c_int64, ctypes.c_int32]),
    "sph_pack": (ctypes.c_int32, [_vp, _P(SphView), _vp, _vp, _vp, _vp, _vp, _vp]),
    "sph_batch_ncells": (ctypes.c_int64, [ctypes.c_int32, _P(SphBatchScene)]),
    "sph_batch_create": (ctypes.c_int32, [ctypes.c_int32, _P(SphBatchScene),
                                          _P(SphBatchBuffers), _P(_vp)]),
    "sph_batch_destroy": (ctypes.c_int32, [_vp]),
    "sph_batch_bin": (ctypes.c_int32, [_vp, _vp, ctypes.c_int32]),
    "sph_batch_pack": (ctypes.c_int32, [_vp, _vp, ctypes.c_int32]),
    "sph_batch_density": (ctypes.c_int32, [_vp, _vp, ctypes.c_int32, ctypes.c_int32,
                                           _P(SphBatchSched)]),
    "sph_batch_stage": (ctypes.c_int32, [_vp, _vp, ctypes.c_int32, _P(SphBatchSched)]),
    "sph_mlp_param_count": (ctypes.c_int64, []),
    "sph_mlp_create": (ctypes.c_int32, [_P(_vp)]),
    "sph_mlp_destroy": (ctypes.c_int32, [_vp]),
    "sph_mlp_set_params": (ctypes.c_int32, [_vp, _vp]),
    "sph_mlp_get": (ctypes.c_int32, [_vp, ctypes.c_int32, _vp]),
    "sph_mlp_zero_grad": (ctypes.c_int32, [_vp, _vp]),
    "sph_mlp_timing": (ctypes.c_int32, [_vp, ctypes.c_int32, _vp, _vp]),
    "sph_mlp_forward": (ctypes.c_int32, [_vp, _vp, _P(SphMlpGraph), _vp]),
    "sph_mlp_loss": (ctypes.c_int32, [_vp, _vp, _P(SphMlpGraph), _vp, _vp, _vp, _vp]),
    "sph_mlp_backward": (ctypes.c_int32, [_vp, _vp, _P(SphMlpGraph), _vp]),
    "sph_mlp_adam": (ctypes.c_int32, [_vp, _vp, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                      ctypes.c_float, ctypes.c_int32]),
    "sph_mlp_integrate": (ctypes.c_int32, [_vp, ctypes.c_int64, _vp, _vp, ctypes.c_double,
                                           ctypes.c_double, _vp, _vp, _vp, _vp]),
    "sph_graph_begin": (ctypes.c_int32, [_vp]),
    "sph_graph_end": (ctypes.c_int32, [_vp, _P(_vp)]),
    "sph_graph_launch": (ctypes.c_int32, [_vp, _vp]),
    "sph_graph_destroy": (ctypes.c_int32, [_vp]),
}

EXPORTS = tuple(_SIGS)

_lib = None


def load(path=LIB_PATH):
    """Load the library (no GPU call is made).  Raises OSError when the
    library has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise OSError("libsph_mi355x.so is not built (%s); run "
                      "`python -m pysph_amd.build` first" % path)
    L = ctypes.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    if L.sph_abi_version() != 1:
        raise OSError("libsph_mi355x.so ABI version mismatch")
    _lib = L
    return L


# Optional launch tracer: an object with before(name, args) and
# after(name, args), called around every entry-point call (used to bracket
# launches with stream events when timing kernels).
tracer = None


def call(name, *args):
    """Call an ABI entry point and raise SphError on a non-zero return."""
    L = load()
    fn = getattr(L, name)
    t = tracer
    if t is None:
        rc = fn(*args)
    else:
        t.before(name, args)
        rc = fn(*args)
        t.after(name, args)
    if rc != SPH_OK:
        msg = L.sph_last_error().decode(errors="replace")
        cls = JumpToleranceError if rc == SPH_EJUMP else SphError
        raise cls(name, rc, msg)
    return rc


def ptr_array(ptrs):
    """A C array of void* from ints / None."""
    arr = (_vp * max(len(ptrs), 1))()
    for i, p in enumerate(ptrs):
        arr[i] = p if p else None
    return arr
