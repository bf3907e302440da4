"""ParticleArray: named SoA particle properties in GPU memory.

Reference: base/particle_array.pyx:96-1279 and get_particle_array
(base/particles.py:282-356).

Layout.  Every property is one contiguous device tensor.  Real properties
share one dtype chosen by `cl_precision` (the reference's device-precision
knob): 'double' -> float64, 'single' -> float32.  'int' properties are
int32, 'long' properties int64.

Ordering.  The cell binning keeps each array sorted by cell (a counting
sort, csrc/sph_bin.hip), so the device slots are a permutation of the
particles the user created.  `_order[slot]` is the user index of the
particle in that slot.  `get()` / `set()` / `properties` always speak user
order, exactly like the reference; only device-level code sees slots.

Tags.  As in the reference, LocalReal particles come first in user order:
`align_particles` (called by `initialize`, by `set(tag=...)` and after
particles are added or removed) applies the reference's own permutation
(particle_array.pyx:954-1040) to the user order, `num_real_particles`
counts them, and `get()` (with its default only_real_particles=True) and
attribute access return the first num_real_particles values.  Aligning
renumbers user indices only; no device data moves.
"""
import numpy as np
import torch

from .. import _abi
from .._device import default_device, dtype_code, ptr, require_gpu
from .particle_types import LocalReal, ParticleType

_REAL_TYPES = ("double", "float", "real")
_POS = ("x", "y", "z")


def _kind_of(t):
    if t in _REAL_TYPES or t is None:
        return "real"
    if t == "int":
        return "int"
    if t == "long":
        return "long"
    raise ValueError("property type %r not understood" % (t,))


class ParticleArray:
    """A collection of particles of one type with named properties."""

    def __init__(self, name="", default_particle_tag=LocalReal,
                 particle_type=ParticleType.Fluid, cl_precision="double", constants=None,
                 **props):
        d = self.__dict__
        d["name"] = name
        d["particle_type"] = particle_type
        d["cl_precision"] = cl_precision
        d["dtype"] = torch.float64 if cl_precision == "double" else torch.float32
        d["device"] = default_device()
        d["constants"] = dict(constants or {})
        d["is_dirty"] = True
        d["_props"] = {}
        d["_kinds"] = {}
        d["default_values"] = {"tag": default_particle_tag, "group": 0, "local": 1}
        d["_order"] = None  # None: identity
        d["_nreal"] = 0
        d["temporary_arrays"] = set()  # names of temporary (real) arrays
        d["indices_invalid"] = True
        # properties a re-binning leaves in place (set by an integrator whose
        # fused path never reads them; their values are then meaningless)
        d["_unpermuted"] = set()
        d["_n"] = 0
        # binning state written by CellManager
        d["_cell"] = None
        d["_cell_start"] = None
        d["_bin_bufs"] = None
        d["_refcell"] = None
        d["_grid_key"] = None
        self.initialize(**props)

    # ------------------------------------------------------------------ set-up
    def initialize(self, **props):
        n = 0
        for info in props.values():
            if isinstance(info, dict) and info.get("data") is not None:
                n = max(n, len(np.atleast_1d(info["data"])))
        self.__dict__["_n"] = n
        self._props.clear()
        self._kinds.clear()
        self.__dict__["_order"] = None
        for nm, kind in (("tag", "long"), ("group", "long"), ("local", "int")):
            self._props[nm] = self._full(kind, self.default_values[nm])
            self._kinds[nm] = kind
        self.temporary_arrays.clear()
        for nm, info in props.items():
            if not isinstance(info, dict):
                info = {"data": info}
            info = dict(info)
            info.setdefault("name", nm)
            self.add_property(info)
        self.__dict__["is_dirty"] = True
        self.__dict__["_h_version"] = self.__dict__.get("_h_version", 0) + 1
        self.align_particles()

    def _torch_dtype(self, kind):
        return {"real": self.dtype, "int": torch.int32, "long": torch.int64}[kind]

    def _full(self, kind, value):
        return torch.full((self._n,), value, dtype=self._torch_dtype(kind), device=self.device)

    def add_property(self, prop_info):
        """Add a property: {'name', 'type', 'default', 'data'}
        (particle_array.pyx:729-883)."""
        nm = prop_info["name"]
        kind = _kind_of(prop_info.get("type", "double"))
        if nm in self.default_values and "data" not in prop_info:
            return
        data = prop_info.get("data")
        if data is not None and len(np.atleast_1d(data)) > 0:
            arr = np.ascontiguousarray(np.atleast_1d(data))
            if len(arr) != self._n:
                if self._n == 0 and not self._props_nonempty():
                    self.__dict__["_n"] = len(arr)
                    for k in list(self._props):
                        self._props[k] = self._full(self._kinds[k], self.default_values.get(k, 0))
                else:
                    raise ValueError("property %s has %d values for %d particles"
                                     % (nm, len(arr), self._n))
            t = torch.as_tensor(arr, dtype=self._torch_dtype(kind)).to(self.device)
            if self._order is not None:
                t = t[self._order]
            self._props[nm] = t.contiguous()
        else:
            if nm in self._props:
                return
            default = prop_info.get("default", 0)
            self._props[nm] = self._full(kind, default)
            self.default_values[nm] = default
        self._kinds[nm] = kind

    def _props_nonempty(self):
        return any(t.numel() for t in self._props.values())

    # ------------------------------------------------------------------ access
    def get_number_of_particles(self):
        return self._n

    @property
    def num_real_particles(self):
        """Number of LocalReal particles; they occupy user indices
        [0, num_real_particles) (particle_array.pyx:954-1040)."""
        return self._nreal

    def has_array(self, name):
        return name in self._props

    def _order_np(self):
        return None if self._order is None else self._order.cpu().numpy()

    def _to_user(self, t):
        a = t.detach().cpu().numpy()
        o = self._order_np()
        if o is None:
            return a.copy()
        out = np.empty_like(a)
        out[o] = a
        return out

    def _check_property(self, nm):
        if nm not in self._props and nm not in self.constants:
            raise AttributeError("property %s not present" % nm)

    def get(self, *names, only_real_particles=True):
        """Host copies of the named properties (or constants) in user order
        (particle_array.pyx:617-673): the LocalReal particles only by
        default, every particle with only_real_particles=False."""
        if not names:
            return None
        res = []
        for nm in names:
            self._check_property(nm)
            if nm in self._props:
                a = self._to_user(self._props[nm])
                res.append(a[:self._nreal] if only_real_particles else a)
            else:
                res.append(self.constants[nm])
        return res[0] if len(res) == 1 else tuple(res)

    def set(self, **props):
        """Copy user-order values into the properties, in argument order
        (particle_array.pyx:675-718).  A shorter array sets the leading
        particles, a longer one raises ValueError.  Setting `tag` re-aligns
        the particles at once (LocalReal first), so properties set after it
        in the same call follow the new order; setting x, y or z marks the
        array dirty."""
        for nm in props:
            self._check_property(nm)
        for nm, val in props.items():
            if nm not in self._props:
                self.constants[nm] = val
                continue
            t = self._props[nm]
            a = np.ascontiguousarray(val)
            v = torch.as_tensor(a, dtype=t.dtype).to(self.device).reshape(-1)
            if v.numel() == 1 and t.numel() != 1:
                v = v.expand(t.numel())
            if v.numel() > t.numel():
                raise ValueError("property %s: %d values for %d particles"
                                 % (nm, v.numel(), t.numel()))
            if v.numel() < t.numel():
                cur = self._to_user(t)
                cur[:v.numel()] = a.reshape(-1)
                v = torch.as_tensor(cur, dtype=t.dtype).to(self.device)
            if self._order is not None:
                v = v[self._order]
            t.copy_(v)
            if nm == "tag":
                self.align_particles()
            if nm in _POS:
                self.__dict__["is_dirty"] = True
            if nm == "h":
                self.__dict__["_h_version"] = self.__dict__.get("_h_version", 0) + 1

    # ----------------------------------------------------------- alignment
    def align_particles(self):
        """Move the LocalReal particles to the front of the user order with
        the reference's permutation (particle_array.pyx:954-1040) and count
        them; marks the array dirty when any particle moved."""
        n = self._n
        if n == 0 or "tag" not in self._props:
            self.__dict__["_nreal"] = 0
            return 0
        tag = self._to_user(self._props["tag"])
        index, m = align_index(tag)
        self.__dict__["_nreal"] = m
        if np.array_equal(index, np.arange(n)):
            return 0
        inv = np.empty(n, dtype=np.int64)
        inv[index] = np.arange(n)
        it = torch.as_tensor(inv, device=self.device)
        self.__dict__["_order"] = it if self._order is None else it[self._order]
        self.__dict__["is_dirty"] = True
        self.__dict__["indices_invalid"] = True
        return 0

    def set_indices_invalid(self, value):
        self.__dict__["indices_invalid"] = bool(value)

    def set_name(self, name):
        self.__dict__["name"] = name

    def set_particle_type(self, particle_type):
        self.__dict__["particle_type"] = particle_type

    def get_property_index(self, prop_name):
        return self._props.get(prop_name)

    def get_carray(self, name):
        """The device tensor of a property (slot order)."""
        if name is None:
            return None
        return self._props.get(name)

    dev = get_carray

    def set_dirty(self, value):
        self.__dict__["is_dirty"] = bool(value)

    @property
    def properties(self):
        """Every property (temporary arrays excluded), all particles, user
        order."""
        return {k: self._to_user(v) for k, v in self._props.items()
                if k not in self.temporary_arrays}

    def get_property_names(self):
        return [k for k in self._props if k not in self.temporary_arrays]

    def add_temporary_array(self, arr_name):
        """A temporary real array (particle_array.pyx:345-361): carried and
        re-ordered with the particles, not pickled."""
        if arr_name in self._props and arr_name not in self.temporary_arrays:
            raise ValueError("property (%s) exists" % arr_name)
        if arr_name not in self._props:
            self._props[arr_name] = self._full("real", 0.0)
            self._kinds[arr_name] = "real"
            self.temporary_arrays.add(arr_name)

    def remove_property(self, prop_name):
        self._props.pop(prop_name, None)
        self._kinds.pop(prop_name, None)
        self.default_values.pop(prop_name, None)
        self.temporary_arrays.discard(prop_name)

    def clear(self):
        """Drop every particle and property but tag, group and local
        (particle_array.pyx:256-265)."""
        tag_default = self.default_values["tag"]
        self.default_values.clear()
        self.default_values.update({"tag": tag_default, "group": 0, "local": 1})
        self._props.clear()
        self._kinds.clear()
        self.temporary_arrays.clear()
        self._reset_layout({}, 0)
        for nm, kind in (("tag", "long"), ("group", "long"), ("local", "int")):
            self._props[nm] = self._full(kind, self.default_values[nm])
            self._kinds[nm] = kind
        self.__dict__["indices_invalid"] = True

    # ------------------------------------------------------------ pickling
    def __getstate__(self):
        """Name, type, constants, the names of the temporary arrays and every
        property with its type, default and data (all particles, user order)
        (particle_array.pyx:200-221)."""
        props = {}
        for k, v in self._props.items():
            if k in self.temporary_arrays:
                continue
            props[k] = {"name": k, "type": {"real": "double", "int": "int",
                                            "long": "long"}[self._kinds[k]],
                        "data": self._to_user(v), "default": self.default_values.get(k, 0)}
        return {"name": self.name, "particle_type": self.particle_type,
                "cl_precision": self.cl_precision, "constants": dict(self.constants),
                "temporary_arrays": sorted(self.temporary_arrays), "properties": props}

    def __setstate__(self, d):
        """Rebuild from __getstate__'s dict on the default device
        (particle_array.pyx:223-239); temporary arrays come back empty."""
        ParticleArray.__init__(self, name=d["name"], particle_type=d["particle_type"],
                               cl_precision=d.get("cl_precision", "double"),
                               constants=d["constants"])
        self.initialize(**{k: dict(v) for k, v in d["properties"].items()})
        for k, v in d["properties"].items():
            self.default_values[k] = v.get("default", 0)
        for nm in d.get("temporary_arrays", ()):
            self.add_temporary_array(nm)

    def __getattr__(self, name):
        props = self.__dict__.get("_props", {})
        if name in props:
            return self._to_user(props[name])[:self.__dict__["_nreal"]]
        if name in self.__dict__.get("constants", {}):
            return self.__dict__["constants"][name]
        raise AttributeError("property %s not found" % name)

    def __setattr__(self, name, value):
        if name in self.__dict__:
            self.__dict__[name] = value
        elif name in self._props:
            self.set(**{name: value})
        elif name in self.constants:
            self.constants[name] = value
        else:
            raise AttributeError("property %s not found" % name)

    # --------------------------------------------------------------- editing
    def _user_tensors(self):
        """All properties as device tensors in user order."""
        if self._order is None:
            return dict(self._props)
        inv = torch.empty_like(self._order)
        inv[self._order] = torch.arange(self._n, device=self.device)
        return {k: v[inv].contiguous() for k, v in self._props.items()}

    def _reset_layout(self, tensors, n):
        self._props.clear()
        self._props.update(tensors)
        d = self.__dict__
        d["_h_version"] = d.get("_h_version", 0) + 1
        d["_n"] = n
        d["_order"] = None
        d["_cell"] = None
        d["_cell_start"] = None
        d["_bin_bufs"] = None
        d["_refcell"] = None
        d["_grid_key"] = None
        d["is_dirty"] = True

    def remove_particles(self, index_list):
        """Remove the particles at the given user indices as the reference's
        arrays do (carray_pyx.src:339-376): from the largest index down,
        each removed particle is replaced by the current last one.  Indices
        past the end are skipped; more indices than particles raise
        ValueError (particle_array.pyx:363-418).  Re-aligns."""
        idx = _as_index(index_list)
        if len(idx) > self._n:
            raise ValueError("Number of particles to be removed is greater than number of "
                             "particles in array")
        keep = np.arange(self._n, dtype=np.int64)
        length = self._n
        for i in np.sort(idx)[::-1]:
            if 0 <= i < length:
                keep[i] = keep[length - 1]
                length -= 1
        if length == self._n:
            return
        kt = torch.as_tensor(keep[:length], device=self.device)
        t = {k: v[kt].contiguous() for k, v in self._user_tensors().items()}
        self._reset_layout(t, length)
        self.align_particles()

    def remove_tagged_particles(self, tag):
        """Remove every particle with the given tag (particle_array.pyx:420-439)."""
        t = self._to_user(self._props["tag"])
        self.remove_particles(np.flatnonzero(t == tag))

    def remove_flagged_particles(self, flag_name, flag_value):
        """Remove every particle whose int property flag_name equals
        flag_value (particle_array.pyx:441-459)."""
        f = self._to_user(self._props[flag_name])
        self.remove_particles(np.flatnonzero(f == flag_value))

    def extend(self, num_particles):
        """Append num_particles particles with every property at its default
        (particle_array.pyx:573-594); no re-alignment."""
        if num_particles <= 0:
            return
        t = {}
        for k, v in self._user_tensors().items():
            pad = torch.full((num_particles,), self.default_values.get(k, 0), dtype=v.dtype,
                             device=self.device)
            t[k] = torch.cat([v, pad]).contiguous()
        self._reset_layout(t, self._n + num_particles)

    def add_particles(self, **particle_props):
        """Append particles given arrays of some existing properties; the
        others take their defaults (particle_array.pyx:461-521).  Re-aligns."""
        if not particle_props:
            return 0
        for nm in particle_props:
            if nm not in self._props or nm in self.temporary_arrays:
                raise AttributeError("property %s not present" % nm)
        k = len(np.atleast_1d(next(iter(particle_props.values()))))
        old = self._n
        self.extend(k)
        if k > 0:
            for nm, val in particle_props.items():
                cur = self._to_user(self._props[nm])
                cur[old:] = np.asarray(val).reshape(-1)
                self._props[nm] = torch.as_tensor(cur, dtype=self._props[nm].dtype).to(
                    self.device)
            self.align_particles()
            self.__dict__["is_dirty"] = True
        return 0

    def extract_particles(self, index_array, props=None):
        """A new ParticleArray with the given user indices (all properties or
        the listed ones), aligned (particle_array.pyx:1042-1099)."""
        idx = _as_index(index_array)
        it = torch.as_tensor(idx, device=self.device)
        names = list(self._props) if props is None else list(props)
        out = ParticleArray(name=self.name, particle_type=self.particle_type,
                            cl_precision=self.cl_precision, constants=dict(self.constants))
        user = self._user_tensors()
        t = {k: user[k][it].contiguous() for k in names}
        for k in ("tag", "group", "local"):
            if k not in t:
                t[k] = torch.full((len(idx),), out.default_values[k],
                                  dtype=out._torch_dtype(out._kinds[k]), device=out.device)
        for k in names:
            out._kinds[k] = self._kinds[k]
            if k in self.default_values:
                out.default_values[k] = self.default_values[k]
        out._reset_layout(t, len(idx))
        out.align_particles()
        return out

    def append_parray(self, parray):
        """Append the particles of another array; properties it has and this
        one lacks are added (particle_array.pyx:523-571).  Re-aligns."""
        n2 = parray.get_number_of_particles()
        if n2 == 0:
            return 0
        mine = self._user_tensors()
        other = parray._user_tensors()
        for k, kind in parray._kinds.items():
            if k in parray.temporary_arrays:
                continue
            if k not in mine:
                self._kinds[k] = kind
                self.default_values[k] = parray.default_values.get(k, 0)
                mine[k] = torch.full((self._n,), self.default_values[k],
                                     dtype=self._torch_dtype(kind), device=self.device)
        t = {}
        for k, v in mine.items():
            o = other.get(k)
            if o is None or k in parray.temporary_arrays:
                o = torch.full((n2,), self.default_values.get(k, 0), dtype=v.dtype,
                               device=self.device)
            t[k] = torch.cat([v, o.to(v.dtype)]).contiguous()
        for c, v in parray.constants.items():
            self.constants.setdefault(c, v)
        self._reset_layout(t, self._n + n2)
        self.align_particles()
        return 0

    def copy_properties(self, source, start_index=-1, end_index=-1):
        """Copy every property both arrays have from `source` into user
        indices [start_index, end_index) (particle_array.pyx:1121-1140)."""
        n = self._n
        a = 0 if start_index < 0 else int(start_index)
        b = n if end_index < 0 else int(end_index)
        for k in source.get_property_names():
            if k not in self._props:
                continue
            cur = self._to_user(self._props[k])
            src = source._to_user(source._props[k])
            cur[a:b] = src[:b - a]
            self._props[k] = torch.as_tensor(cur, dtype=self._props[k].dtype).to(self.device)
            if self._order is not None:
                self._props[k] = self._props[k][self._order].contiguous()

    def set_flag(self, flag_name, flag_value, indices):
        """Set int property flag_name to flag_value at the given user indices."""
        self._set_at(flag_name, flag_value, indices)

    def set_tag(self, tag_value, indices):
        """Set tag to tag_value at the given user indices (no re-alignment,
        as particle_array.pyx:1110-1116)."""
        self._set_at("tag", tag_value, indices)

    def _set_at(self, nm, value, indices):
        cur = self._to_user(self._props[nm])
        cur[_as_index(indices)] = value
        t = torch.as_tensor(cur, dtype=self._props[nm].dtype).to(self.device)
        if self._order is not None:
            t = t[self._order]
        self._props[nm].copy_(t)

    # -------------------------------------------------------- device layout
    def apply_permutation(self, perm, extra=(), extra_out=None):
        """Gather every property (and the order map) through perm
        (slot p <- slot perm[p]) on the GPU.  `extra` per-particle tensors
        are gathered along and returned in a list: into `extra_out` (tensors
        of the same shapes, none of them an input) when given, else into
        new tensors."""
        stream = require_gpu(perm)
        n = self._n
        if self._order is None:
            self.__dict__["_order"] = torch.arange(n, dtype=torch.int64, device=self.device)
        extra = list(extra)
        moved = [k for k in self._props if k not in self._unpermuted]
        names = moved + ["__order__"] + ["__extra__"] * len(extra)
        srcs = [self._props[k] for k in moved] + [self._order] + extra
        # properties and the order map alternate between two buffers (the
        # previous layout's buffer receives the next one), so that a
        # captured step finds its buffers again two binnings later
        spare = self.__dict__.setdefault("_spare", {})

        def target(name, s):
            t = spare.get(name)
            if (name == "__extra__" or t is None or t is s or t.shape != s.shape or
                    t.dtype != s.dtype or t.device != s.device):
                t = torch.empty_like(s)
            return t
        dsts = [target(k, s) for k, s in zip(names, srcs)]
        if extra_out is not None:
            extra_out = list(extra_out)
            if len(extra_out) != len(extra) or any(
                    o.shape != e.shape or o.dtype != e.dtype or o.device != e.device or
                    any(o is x for x in extra) for o, e in zip(extra_out, extra)):
                raise ValueError("extra_out must hold one tensor of the same shape, dtype and "
                                 "device per extra tensor, none of them an input")
            dsts[len(dsts) - len(extra):] = extra_out
        esz = (ctypes_int32 * len(srcs))(*[s.element_size() for s in srcs])
        _abi.call("sph_permute", stream, n, perm.data_ptr(), len(srcs),
                  _abi.ptr_array([s.data_ptr() for s in srcs]),
                  _abi.ptr_array([d.data_ptr() for d in dsts]), esz)
        out = []
        for k, s, t in zip(names, srcs, dsts):
            if k == "__order__":
                self.__dict__["_order"] = t
                spare[k] = s
            elif k == "__extra__":
                out.append(t)
            else:
                self._props[k] = t
                spare[k] = s
        return out

    def view(self, fields=None):
        """The sph_view of this array for the C ABI."""
        v = _abi.SphView()
        v.n = self._n
        v.dtype = dtype_code(self.dtype)
        for f in ("x", "y", "z", "u", "v", "w", "h", "m", "rho", "p", "cs"):
            t = self._props.get(f)
            setattr(v, f, ptr(t) if (fields is None or f in fields) else None)
        tag = self._props.get("tag")
        v.tag = ptr(tag)
        v.cell = ptr(self._cell)
        v.cell_start = ptr(self._cell_start)
        return v

    def __repr__(self):
        return "ParticleArray(name=%r, n=%d, type=%d, dtype=%s)" % (
            self.name, self._n, self.particle_type, self.dtype)


import ctypes as _ct  # noqa: E402

ctypes_int32 = _ct.c_int32


def get_particle_array(cl_precision="double", **props):
    """Create a ParticleArray with the reference's default properties
    x,y,z,u,v,w,m,h,p,e,rho,cs and idx (particles.py:282-356)."""
    prop_dict = {}
    name = ""
    particle_type = ParticleType.Fluid
    default_props = {"x": 0.0, "y": 0.0, "z": 0.0, "u": 0.0, "v": 0.0, "w": 0.0, "m": 1.0,
                     "h": 1.0, "p": 0.0, "e": 0.0, "rho": 1.0, "cs": 0.0}
    np_ = 0
    for prop, val in props.items():
        if prop in ("name", "type"):
            continue
        np_ = len(val)
        if prop == "idx":
            prop_dict[prop] = {"data": np.asarray(val), "type": "long"}
        else:
            prop_dict[prop] = {"data": np.asarray(val), "type": "double"}
    for prop, dv in default_props.items():
        if prop not in props:
            prop_dict[prop] = {"name": prop, "type": "double", "default": dv}
    if "idx" not in prop_dict and np_ != 0:
        prop_dict["idx"] = {"name": "idx", "data": np.arange(np_), "type": "long"}
    if "name" in props:
        name = props["name"]
    if "type" in props:
        particle_type = props["type"]
        assert particle_type in (ParticleType.Fluid, ParticleType.Solid,
                                 ParticleType.Probe), "Type not understood!"
    return ParticleArray(name=name, particle_type=particle_type, cl_precision=cl_precision,
                         **prop_dict)


def _as_index(index_list):
    a = index_list.get_npy_array() if hasattr(index_list, "get_npy_array") else index_list
    return np.asarray(a, dtype=np.int64).reshape(-1)


def align_index(tag):
    """The permutation of the reference's align_particles
    (particle_array.pyx:954-1040): property p becomes p[index] and the first
    m = count(tag == LocalReal) entries are the LocalReal particles in order.

    The reference walks the particles once and, for the k-th LocalReal
    particle at position i_k != k, swaps index entries k and i_k.  Position
    k < m therefore ends with i_k, a non-real position p >= m keeps p, and
    a LocalReal position p = i_k >= m ends with whatever position k held
    just before step k: k itself, unless k is the j-th LocalReal particle
    with j < k, in which case it held what position j held before step j.
    That chain is followed here by pointer jumping instead of a loop."""
    tag = np.asarray(tag)
    n = len(tag)
    index = np.arange(n, dtype=np.int64)
    real = np.flatnonzero(tag == LocalReal)
    m = len(real)
    if m == 0:
        return index, 0
    index[:m] = real
    rank = np.full(n, -1, dtype=np.int64)
    rank[real] = np.arange(m)
    ks = np.arange(m, dtype=np.int64)
    rk = rank[:m]
    nxt = np.where((rk >= 0) & (rk < ks), rk, ks)
    while True:
        jumped = nxt[nxt]
        if np.array_equal(jumped, nxt):
            break
        nxt = jumped
    far = np.flatnonzero(real >= m)
    index[real[far]] = nxt[far]
    return index, m
