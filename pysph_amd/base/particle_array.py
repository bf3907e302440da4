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
        # properties a re-binning leaves in place (set by an integrator whose
        # fused path never reads them; their values are then meaningless)
        d["_unpermuted"] = set()
        d["_n"] = 0
        # binning state written by CellManager
        d["_cell"] = None
        d["_cell_start"] = None
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
        for nm, info in props.items():
            if not isinstance(info, dict):
                info = {"data": info}
            info = dict(info)
            info.setdefault("name", nm)
            self.add_property(info)
        self.__dict__["is_dirty"] = True

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
        return self._n

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

    def get(self, *names, only_real_particles=True):
        """Host copies of the named properties in user order
        (particle_array.pyx:617-673)."""
        res = []
        for nm in names:
            if nm not in self._props:
                raise AttributeError("property %s not present" % nm)
            res.append(self._to_user(self._props[nm]))
        return res[0] if len(res) == 1 else res

    def set(self, **props):
        """Copy user-order values into the properties; setting x, y or z
        marks the array dirty (particle_array.pyx:675-718)."""
        for nm, val in props.items():
            if nm not in self._props:
                raise AttributeError("property %s not present" % nm)
            t = self._props[nm]
            v = torch.as_tensor(np.ascontiguousarray(val), dtype=t.dtype).to(self.device)
            if v.numel() == 1 and t.numel() != 1:
                v = v.expand(t.numel())
            if self._order is not None:
                v = v[self._order]
            t.copy_(v)
            if nm in _POS:
                self.__dict__["is_dirty"] = True

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
        return {k: self._to_user(v) for k, v in self._props.items()}

    def get_property_names(self):
        return list(self._props)

    def __getattr__(self, name):
        props = self.__dict__.get("_props", {})
        if name in props:
            return self._to_user(props[name])
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
        d["_n"] = n
        d["_order"] = None
        d["_cell"] = None
        d["_cell_start"] = None
        d["_grid_key"] = None
        d["is_dirty"] = True

    def remove_particles(self, index_list):
        """Remove the particles at the given user indices
        (particle_array.pyx:405-458)."""
        idx = np.asarray(index_list.get_npy_array() if hasattr(index_list, "get_npy_array")
                         else index_list, dtype=np.int64)
        keep = np.ones(self._n, dtype=bool)
        keep[idx] = False
        kt = torch.as_tensor(keep, device=self.device)
        t = {k: v[kt].contiguous() for k, v in self._user_tensors().items()}
        self._reset_layout(t, int(keep.sum()))

    def extract_particles(self, index_array):
        """A new ParticleArray with the given user indices
        (particle_array.pyx:1042-1099)."""
        idx = np.asarray(index_array.get_npy_array() if hasattr(index_array, "get_npy_array")
                         else index_array, dtype=np.int64)
        it = torch.as_tensor(idx, device=self.device)
        out = ParticleArray(name=self.name, particle_type=self.particle_type,
                            cl_precision=self.cl_precision)
        t = {k: v[it].contiguous() for k, v in self._user_tensors().items()}
        out._kinds.update(self._kinds)
        out.default_values.update(self.default_values)
        out._reset_layout(t, len(idx))
        return out

    def append_parray(self, parray):
        """Append the particles of another array (particle_array.pyx:885-952)."""
        mine = self._user_tensors()
        other = parray._user_tensors()
        n2 = parray.get_number_of_particles()
        for k, kind in parray._kinds.items():
            if k not in mine:
                self._kinds[k] = kind
                mine[k] = torch.full((self._n,), self.default_values.get(k, 0),
                                     dtype=self._torch_dtype(kind), device=self.device)
        t = {}
        for k, v in mine.items():
            o = other.get(k)
            if o is None:
                o = torch.full((n2,), self.default_values.get(k, 0), dtype=v.dtype,
                               device=self.device)
            t[k] = torch.cat([v, o.to(v.dtype)]).contiguous()
        self._reset_layout(t, self._n + n2)

    # -------------------------------------------------------- device layout
    def apply_permutation(self, perm, extra=()):
        """Gather every property (and the order map) through perm
        (slot p <- slot perm[p]) on the GPU.  `extra` per-particle tensors
        are gathered along and returned in a new list."""
        stream = require_gpu(perm)
        n = self._n
        if self._order is None:
            self.__dict__["_order"] = torch.arange(n, dtype=torch.int64, device=self.device)
        extra = list(extra)
        moved = [k for k in self._props if k not in self._unpermuted]
        names = moved + ["__order__"] + ["__extra__"] * len(extra)
        srcs = [self._props[k] for k in moved] + [self._order] + extra
        dsts = [torch.empty_like(s) for s in srcs]
        esz = (ctypes_int32 * len(srcs))(*[s.element_size() for s in srcs])
        _abi.call("sph_permute", stream, n, perm.data_ptr(), len(srcs),
                  _abi.ptr_array([s.data_ptr() for s in srcs]),
                  _abi.ptr_array([d.data_ptr() for d in dsts]), esz)
        out = []
        for k, t in zip(names, dsts):
            if k == "__order__":
                self.__dict__["_order"] = t
            elif k == "__extra__":
                out.append(t)
            else:
                self._props[k] = t
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
