"""Drop-in counterpart of raypier.core.ctracer for the gfx950 tracer.

Ray generations live in HBM as structure-of-arrays (see _device.DeviceRays);
the trace functions call librpx.so, whose kernels reproduce the reference's
trace loop (raypier/core/ctracer.pyx:2062-2385).  Small-object conveniences
(Ray, vector helpers) are plain Python as in the reference.
"""
import ctypes
import itertools
import math
import time

import numpy as np

from . import _abi
from ._abi import ray_dtype, para_dtype, gausslet_dtype, REFL_RAY, GAUSSLET, PARABASAL
from ._lib import load, check, stream_handle, require_gpu

GAUSSLET_ = GAUSSLET
PARABASAL_ = PARABASAL
INF = float("inf")
NPARA = 6

# ------------------------------------------------------------ vector maths
# Python-level helpers of ctracer.pyx:83-262.


def py_set_v(O):
    return (float(O[0]), float(O[1]), float(O[2]))


def sep(a, b):
    x, y, z = (b[0] - a[0]), (b[1] - a[1]), (b[2] - a[2])
    return math.sqrt((x * x) + (y * y) + (z * z))


def invert(v):
    return (-v[0], -v[1], -v[2])


def multvv(a, b):
    return (a[0] * b[0], a[1] * b[1], a[2] * b[2])


def multvs(a, b):
    return (a[0] * b, a[1] * b, a[2] * b)


def addvv(a, b):
    return (a[0] + b[0], a[1] + b[1], a[2] + b[2])


def addvs(a, b):
    return (a[0] + b, a[1] + b, a[2] + b)


def subvv(a, b):
    return (a[0] - b[0], a[1] - b[1], a[2] - b[2])


def subvs(a, b):
    return (a[0] - b, a[1] - b, a[2] - b)


def mag(a):
    return math.sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2])


def mag_sq(a):
    return a[0] * a[0] + a[1] * a[1] + a[2] * a[2]


def dotprod(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def cross(a, b):
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def norm(a):
    m = math.sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2])
    return (a[0] / m, a[1] / m, a[2] / m)


class Transform(object):
    """Rotation (3x3) plus translation (ctracer.pyx:268-295)."""

    def __init__(self, rotation=((1, 0, 0), (0, 1, 0), (0, 0, 1)), translation=(0, 0, 0)):
        self.rotation = rotation
        self.translation = translation

    @property
    def rotation(self):
        return [list(r) for r in self._rot.reshape(3, 3)]

    @rotation.setter
    def rotation(self, rot):
        self._rot = np.array(rot, dtype=np.float64).reshape(9)

    @property
    def translation(self):
        return tuple(self._t)

    @translation.setter
    def translation(self, dt):
        self._t = np.array(dt, dtype=np.float64).reshape(3)


def transform(t, p):
    m = t._rot
    x, y, z = p
    return (x * m[0] + y * m[1] + z * m[2] + t._t[0],
            x * m[3] + y * m[4] + z * m[5] + t._t[1],
            x * m[6] + y * m[7] + z * m[8] + t._t[2])


def get_ray_size():
    return ray_dtype.itemsize


# -------------------------------------------------------------- records
class _Record(object):
    _dtype = None

    def __init__(self, **kwds):
        self._a = np.zeros(1, self._dtype)
        self.max_length = 1000.0
        for k, v in kwds.items():
            setattr(self, k, v)

    @property
    def _r(self):
        return self._a[0]

    def _vec(self, name):
        v = self._a[name][0]
        return (float(v[0]), float(v[1]), float(v[2]))

    def _set_vec(self, name, v):
        self._a[name][0] = (v[0], v[1], v[2])


def _vprop(name, doc=None):
    return property(lambda self: self._vec(name), lambda self, v: self._set_vec(name, v), doc=doc)


def _sprop(name, cast, doc=None):
    return property(lambda self: cast(self._a[name][0]),
                    lambda self, v: self._a.__setitem__(name, cast(v)), doc=doc)


def _cprop(name, doc=None):
    def get(self):
        v = self._a[name][0]
        return complex(v.real, v.imag)

    def set_(self, v):
        v = complex(v)
        self._a[name] = v
    return property(get, set_, doc=doc)


class ParabasalRay(_Record):
    """Wrapper around one para_t record (ctracer.pyx:332-393)."""
    _dtype = para_dtype
    origin = _vprop("origin")
    direction = _vprop("direction")
    normal = _vprop("normal")
    length = _sprop("length", float)

    def __repr__(self):
        return "Parabasal Ray(o=%s, d=%s)" % (str(self.origin), str(self.direction))

    @property
    def termination(self):
        length = float(np.float32(min(self.length, self.max_length)))
        return addvv(self.origin, multvs(self.direction, length))


class Ray(_Record):
    """Wrapper around one ray_t record (ctracer.pyx:396-676)."""
    _dtype = ray_dtype
    origin = _vprop("origin", "Origin coordinates of the ray")
    direction = _vprop("direction", "direction of the ray, normalised to a unit vector")
    normal = _vprop("normal", "normal vector for the face which created this ray")
    E_vector = _vprop("E_vector", "Unit vector giving the direction of E-field polarisation")
    length = _sprop("length", float)
    phase = _sprop("phase", float)
    accumulated_path = _sprop("accumulated_path", float)
    wavelength_idx = _sprop("wavelength_idx", int)
    parent_idx = _sprop("parent_idx", int)
    end_face_idx = _sprop("end_face_idx", int)
    ray_ident = _sprop("ray_ident", int)
    ray_type_id = _sprop("ray_type_id", int)
    refractive_index = _cprop("refractive_index")
    E1_amp = _cprop("E1_amp")
    E2_amp = _cprop("E2_amp")

    def __repr__(self):
        return "Ray(o=%s, d=%s)" % (str(self.origin), str(self.direction))

    @property
    def termination(self):
        # `length` is a C float in the reference (ctracer.pyx:493)
        length = float(np.float32(self.max_length if self.length > self.max_length else self.length))
        return addvv(self.origin, multvs(self.direction, length))

    @property
    def power(self):
        n = self.refractive_index.real
        e1, e2 = self.E1_amp, self.E2_amp
        return (e1.real ** 2 + e1.imag ** 2) * n + (e2.real ** 2 + e2.imag ** 2) * n

    @property
    def amplitude(self):
        e1, e2 = self.E1_amp, self.E2_amp
        return math.sqrt(e1.real ** 2 + e1.imag ** 2 + e2.real ** 2 + e2.imag ** 2)

    @property
    def jones_vector(self):
        amp = self.amplitude
        return (self.E1_amp / amp, self.E2_amp / amp)

    @property
    def E_left(self):
        return (self.E1_amp + 1.0j * self.E2_amp) / math.sqrt(2)

    @property
    def E_right(self):
        return (self.E1_amp - 1.0j * self.E2_amp) / math.sqrt(2)

    @property
    def ellipticity(self):
        alpha, beta = self.jones_vector
        R = (alpha - 1j * beta) / math.sqrt(2.0)
        L = (alpha + 1j * beta) / math.sqrt(2.0)
        PR = R.real ** 2 + R.imag ** 2
        PL = L.real ** 2 + L.imag ** 2
        return (PR - PL) / (PR + PL)

    @property
    def major_minor_axes(self):
        E2_vector = norm(cross(self.direction, self.E_vector))
        E1_vector = norm(cross(E2_vector, self.direction))
        e1, e2 = self.E1_amp, self.E2_amp
        E1 = math.sqrt(e1.real ** 2 + e1.imag ** 2)
        E2 = math.sqrt(e2.real ** 2 + e2.imag ** 2)
        phi1 = math.atan2(e1.imag, e1.real)
        phi2 = math.atan2(e2.imag, e2.real)
        A = E1 * E1 * math.sin(2 * phi1) + E2 * E2 * math.sin(2 * phi2)
        B = E1 * E1 * math.cos(2 * phi1) + E2 * E2 * math.cos(2 * phi2)
        m = math.sqrt(A * A + B * B)
        t1 = math.atan((B - m) / A)
        t2 = math.atan((B + m) / A)
        axis1 = addvv(multvs(E1_vector, E1 * math.cos(phi1 + t1)), multvs(E2_vector, E2 * math.cos(phi2 + t1)))
        axis2 = addvv(multvs(E1_vector, E1 * math.cos(phi1 + t2)), multvs(E2_vector, E2 * math.cos(phi2 + t2)))
        return (axis1, axis2)

    def project_E(self, *axis):
        """Rotate the E_vector onto the given axis, projecting E1_amp and E2_amp
        (ctracer.pyx:646-676)."""
        E1, E2 = self.E1_amp, self.E2_amp
        d = self.direction
        E_vector = norm(cross(cross(d, self.E_vector), d))
        E_vector2bar = norm(cross(d, axis))
        E_vector2 = norm(cross(E_vector2bar, d))
        A = dotprod(E_vector, E_vector2)
        B = dotprod(E_vector, E_vector2bar)
        self.E1_amp = complex(-(E2.real * B - E1.real * A), -(E2.imag * B - E1.imag * A))
        self.E2_amp = complex((E2.real * A + E1.real * B), E2.imag * A + E1.imag * B)
        self.E_vector = E_vector2


class Gausslet(object):
    """A base ray plus six parabasal rays (ctracer.pyx:679-714)."""

    def __init__(self, **kwds):
        self._a = np.zeros(1, gausslet_dtype)
        for k, v in kwds.items():
            setattr(self, k, v)

    def __repr__(self):
        b = self.base_ray
        return "Gausslet(o=%s, d=%s)" % (str(b.origin), str(b.direction))

    @property
    def base_ray(self):
        r = Ray()
        r._a[0] = self._a["base_ray"][0]
        return r

    @base_ray.setter
    def base_ray(self, r):
        self._a["base_ray"] = r._a[0]

    @property
    def parabasal_rays(self):
        out = []
        for i in range(6):
            p = ParabasalRay()
            p._a[0] = self._a["para_rays"][0][i]
            out.append(p)
        return tuple(out)

    @parabasal_rays.setter
    def parabasal_rays(self, paras):
        for i in range(6):
            self._a["para_rays"][0, i] = paras[i]._a[0]


# ------------------------------------------------------------ collections
class RayArrayView(object):
    """numpy views of ray fields (ctracer.pyx:717-969)."""

    def _view_len(self):
        raise NotImplementedError

    def _ray_field(self, name):
        raise NotImplementedError

    def get_n_rays(self):
        return self._view_len()

    def __getitem__(self, idx):
        n = self._view_len()
        if idx < 0 or idx >= n:
            raise IndexError("Requested index %d from a size %d array" % (idx, n))
        r = Ray()
        r._a[0] = self._record(idx)
        return r

    def get_ray_list(self):
        return [self[i] for i in range(self._view_len())]

    origin = property(lambda self: self._ray_field("origin"))
    direction = property(lambda self: self._ray_field("direction"))
    normal = property(lambda self: self._ray_field("normal"))
    E_vector = property(lambda self: self._ray_field("E_vector"))
    refractive_index = property(lambda self: self._ray_field("refractive_index"))
    E1_amp = property(lambda self: self._ray_field("E1_amp"))
    E2_amp = property(lambda self: self._ray_field("E2_amp"))
    length = property(lambda self: self._ray_field("length"))
    phase = property(lambda self: self._ray_field("phase"))
    accumulated_path = property(lambda self: self._ray_field("accumulated_path"))
    wavelength_idx = property(lambda self: self._ray_field("wavelength_idx"))
    parent_idx = property(lambda self: self._ray_field("parent_idx"))
    end_face_idx = property(lambda self: self._ray_field("end_face_idx"))
    ray_ident = property(lambda self: self._ray_field("ray_ident"))
    ray_type_id = property(lambda self: self._ray_field("ray_type_id"))

    @property
    def termination(self):
        return self.origin + self.direction * self.length[:, None]


_F64_OF = {"origin": (0, 3), "direction": (3, 3), "normal": (6, 3), "E_vector": (9, 3),
           "refractive_index": (12, 2), "E1_amp": (14, 2), "E2_amp": (16, 2), "length": (18, 1),
           "phase": (19, 1), "accumulated_path": (20, 1)}
_U32_OF = {"wavelength_idx": 0, "parent_idx": 1, "end_face_idx": 2, "ray_ident": 3, "ray_type_id": 4}


class _Collection(RayArrayView):
    """Common storage logic: a host AoS copy and/or an HBM SoA copy."""
    _dtype = ray_dtype
    _gausslet = False

    def __init__(self, max_size=0):
        self._host = np.zeros(int(max_size), self._dtype)
        self._n = 0
        self._dev = None
        self._where = "host"      # "host", "device" or "both"
        self._parent = None
        self._wavelengths = np.zeros(0, dtype=np.double)
        self._wl_dev = None

    # storage management ---------------------------------------------
    @classmethod
    def _wrap_device(cls, dev, n):
        obj = cls.__new__(cls)
        _Collection.__init__(obj, 0)
        obj._dev = dev
        obj._n = int(n)
        obj._where = "device"
        return obj

    def _ensure_host(self):
        if self._where == "device":
            data = self._dev.download(self._n)
            if len(self._host) < self._n:
                self._host = np.zeros(max(self._n, 1), self._dtype)
            self._host[: self._n] = data
            self._where = "both"
        return self._host

    def _ensure_device(self):
        if self._where == "host":
            if self._dev is None or self._dev.capacity < self._n:
                from ._device import DeviceRays
                self._dev = DeviceRays(max(self._n, 1), gausslet=self._gausslet)
            self._dev.upload(self._host[: self._n])
            self._where = "both"
        elif self._dev is None:
            from ._device import DeviceRays
            self._dev = DeviceRays(max(self._n, 1), gausslet=self._gausslet)
        return self._dev

    def _host_changed(self):
        self._where = "host"

    def _device_changed(self):
        self._where = "device"

    def _grow(self, n):
        self._ensure_host()
        if n > len(self._host):
            cap = max(n, 2 * len(self._host), 1)
            new = np.zeros(cap, self._dtype)
            new[: self._n] = self._host[: self._n]
            self._host = new

    # sizes ------------------------------------------------------------
    @property
    def n_rays(self):
        return self._n

    @property
    def max_size(self):
        if self._where == "device":
            return self._dev.capacity
        return len(self._host)

    def __len__(self):
        return self._n

    def _view_len(self):
        return self._n

    def clear_ray_list(self):
        self._n = 0

    # wavelengths / parent ---------------------------------------------
    @property
    def wavelengths(self):
        return np.asarray(self._wavelengths)

    @wavelengths.setter
    def wavelengths(self, wl_list):
        self._wavelengths = np.ascontiguousarray(wl_list, dtype=np.double)
        self._wl_dev = None

    def _device_wavelengths(self):
        if self._wl_dev is None:
            torch = require_gpu()
            self._wl_dev = torch.from_numpy(np.ascontiguousarray(self._wavelengths, np.float64)).to("cuda")
        return self._wl_dev

    @property
    def parent(self):
        return self._parent

    @parent.setter
    def parent(self, rc):
        self._parent = rc
        self._neighbours = None
        self._wavelengths = rc._wavelengths
        self._wl_dev = rc._wl_dev

    # array access -----------------------------------------------------
    def copy_as_array(self):
        """Contents as a numpy array of the packed dtype (always a copy)."""
        if self._where == "device":
            return self._dev.download(self._n)
        return self._host[: self._n].copy()

    def _base_field(self, name):
        """numpy copy of one ray_t field of the (base) rays."""
        n = self._n
        if self._where == "device":
            dev = self._dev
            if name in _U32_OF:
                return dev.u32()[_U32_OF[name], :n].cpu().numpy().view(np.uint32).copy()
            start, k = _F64_OF[name]
            rows = dev.f64()[start: start + k, :n].cpu().numpy()
            if name in ("refractive_index", "E1_amp", "E2_amp"):
                return rows[0] + 1j * rows[1]
            if k == 1:
                return rows[0].copy()
            return np.ascontiguousarray(rows.T)
        host = self._host[:n]
        base = host["base_ray"] if self._gausslet else host
        return np.array(base[name])

    def _record(self, idx):
        if self._where == "device":
            return self._dev.download(self._n)[idx]
        return self._host[idx]

    def reset_length(self, max_length=INF):
        """Set the length of every ray (and parabasal ray) to max_length."""
        if self._where == "host":
            if self._gausslet:
                self._host["base_ray"]["length"][: self._n] = max_length
                self._host["para_rays"]["length"][: self._n] = max_length
            else:
                self._host["length"][: self._n] = max_length
        else:
            dev = self._dev
            dev.f64()[_abi.F64_FIELDS.index("length"), : self._n].fill_(float(max_length))
            if self._gausslet:
                dev.para_view()[:, 9, : self._n].fill_(float(max_length))
            if self._where == "both":
                self._where = "device"

    def __iter__(self):
        for i in range(self._n):
            yield self[i]


class RayCollection(_Collection):
    """A growable collection of ray_t records (ctracer.pyx:973-1154).

    Generations produced by tracing stay in HBM; host copies are made only on
    access."""

    def __init__(self, max_size=0):
        super().__init__(max_size)
        self._neighbours = None     # host (n, K) int32 array, or None
        self._nb_dev = None         # device copy / result of propagation, or None
        self._mtime = 0.0

    def _ray_field(self, name):
        return self._base_field(name)

    def __setitem__(self, idx, r):
        if idx < 0 or idx >= self._n:
            raise IndexError("Attempting to set index %d from a size %d array" % (idx, self._n))
        self._ensure_host()
        self._host[idx] = r._a[0]
        self._host_changed()

    def add_ray(self, r):
        self._grow(self._n + 1)
        self._host[self._n] = r._a[0]
        self._n += 1
        self._host_changed()

    def add_ray_list(self, rays):
        for i, r in enumerate(rays):
            if not isinstance(r, Ray):
                raise TypeError("ray list contains non-Ray instance at index %d" % i)
        for r in rays:
            self.add_ray(r)

    @property
    def base_rays(self):
        return self

    # neighbours ---------------------------------------------------------
    # Neighbour indices (n_rays, K) int32 (-1 = none) are given for source
    # rays and derived for every traced generation from its parent's, on the
    # GPU, when first asked for; a derived table is re-derived when an
    # ancestor's table is newer (ctracer.pyx:1068-1131).
    @classmethod
    def _wrap_device(cls, dev, n):
        obj = super()._wrap_device(dev, n)
        obj._neighbours = None
        obj._nb_dev = None
        obj._mtime = 0.0
        return obj

    @property
    def parent(self):
        return self._parent

    @parent.setter
    def parent(self, rc):
        _Collection.parent.fset(self, rc)
        self._nb_dev = None

    def _get_mtime(self, guard):
        if self._parent is None or guard == id(self):
            return self._mtime
        pmtime = self._parent._get_mtime(guard)
        if pmtime > self._mtime:
            self._neighbours = None
            self._nb_dev = None
            return pmtime
        return self._mtime

    def _has_neighbours(self):
        return self._neighbours is not None or self._nb_dev is not None

    def _neighbours_device(self):
        """The neighbour table as an int32 (n_rays, K) cuda tensor, or None."""
        if self._parent is not None:
            pmtime = self._parent._get_mtime(id(self))
            if not (self._mtime >= pmtime and self._has_neighbours()):
                pnb = self._parent._neighbours_device()
                if pnb is not None:
                    self._nb_dev = _propagate_neighbours(self, pnb)
                    self._neighbours = None
                self._mtime = time.monotonic()
        if self._nb_dev is None and self._neighbours is not None:
            torch = require_gpu()
            nb = np.ascontiguousarray(self._neighbours, dtype=np.int32)
            self._nb_dev = torch.from_numpy(nb.reshape(nb.shape[0], -1)).to("cuda")
        return self._nb_dev

    @property
    def neighbours(self):
        if self._parent is None and self._nb_dev is None:
            return None if self._neighbours is None else np.asarray(self._neighbours)
        dev = self._neighbours_device()
        if self._neighbours is None and dev is not None:
            self._neighbours = dev.cpu().numpy()
        return np.asarray(self._neighbours)

    @neighbours.setter
    def neighbours(self, nb):
        self._neighbours = np.asarray(nb, dtype=np.int32)
        self._nb_dev = None
        self._mtime = time.monotonic()

    @classmethod
    def from_array(cls, data):
        """A new RayCollection holding a copy of a ray_dtype array."""
        data = np.ascontiguousarray(data)
        assert data.dtype == ray_dtype
        rc = cls(len(data))
        rc._host[:] = data
        rc._n = len(data)
        return rc

    @classmethod
    def from_device(cls, dev, n):
        """Wrap an HBM SoA buffer (_device.DeviceRays) without copying."""
        return cls._wrap_device(dev, n)


class GaussletBaseRayView(RayArrayView):
    def __init__(self, owner):
        self.owner = owner

    def _view_len(self):
        return self.owner._n

    def __len__(self):
        return self.owner._n

    def _ray_field(self, name):
        return self.owner._base_field(name)

    def _record(self, idx):
        return self.owner._record(idx)["base_ray"]

    def copy_as_array(self):
        return np.ascontiguousarray(self.owner.copy_as_array()["base_ray"])


class GaussletCollection(_Collection):
    """A growable collection of gausslet_t records (ctracer.pyx:1185-1570)."""
    _dtype = gausslet_dtype
    _gausslet = True

    def __getitem__(self, idx):
        if idx < 0 or idx >= self._n:
            raise IndexError("Requested index %d from a size %d array" % (idx, self._n))
        g = Gausslet()
        g._a[0] = self._record(idx)
        return g

    def __setitem__(self, idx, g):
        if idx < 0 or idx >= self._n:
            raise IndexError("Attempting to set index %d from a size %d array" % (idx, self._n))
        self._ensure_host()
        self._host[idx] = g._a[0]
        self._host_changed()

    def _ray_field(self, name):
        return self._base_field(name)

    def add_gausslet(self, g):
        self._grow(self._n + 1)
        self._host[self._n] = g._a[0]
        self._n += 1
        self._host_changed()

    def add_gausslet_list(self, rays):
        for i, r in enumerate(rays):
            if not isinstance(r, Gausslet):
                raise TypeError("ray list contains non-Gausslet instance at index %d" % i)
        for r in rays:
            self.add_gausslet(r)

    def get_gausslet_list(self):
        return [self[i] for i in range(self._n)]

    def extend(self, gc):
        """Append the gausslets of gc (ctracer.pyx:1242-1259).  Collections
        in HBM are joined on the device."""
        if self._where == "host" and gc._where == "host":
            data = gc.copy_as_array()
            self._grow(self._n + len(data))
            self._host[self._n: self._n + len(data)] = data
            self._n += len(data)
            self._host_changed()
            return
        if gc.n_rays == 0:
            return
        from ._device import DeviceRays
        n1, n2 = self._n, gc.n_rays
        out = DeviceRays(n1 + n2, gausslet=True)
        for src, lo, n in ((self, 0, n1), (gc, n1, n2)):
            if n == 0:
                continue
            d = src._ensure_device()
            out.f64()[:, lo: lo + n] = d.f64()[:, :n]
            out.u32()[:, lo: lo + n] = d.u32()[:, :n]
            out.para_view()[:, :, lo: lo + n] = d.para_view()[:, :, :n]
        self._dev = out
        self._n = n1 + n2
        self._host = np.zeros(0, self._dtype)
        self._device_changed()

    @property
    def base_rays(self):
        return GaussletBaseRayView(self)

    @classmethod
    def from_array(cls, data):
        if data.dtype != gausslet_dtype:
            raise ValueError("Array must have gausslet_dtype dtype")
        gc = cls(len(data))
        gc._host[:] = data
        gc._n = len(data)
        return gc

    @classmethod
    def from_rays(cls, data):
        """Gausslets whose six parabasal rays copy the base ray geometry
        (ctracer.pyx:1321-1345)."""
        if data.dtype != ray_dtype:
            raise ValueError("Array must have gausslet_dtype dtype")
        gc = cls(len(data))
        gc._host["base_ray"] = data
        for j in range(6):
            p = gc._host["para_rays"][:, j]
            p["origin"] = data["origin"]
            p["direction"] = data["direction"]
            p["normal"] = data["normal"]
            p["length"] = data["length"]
        gc._n = len(data)
        return gc

    @classmethod
    def from_device(cls, dev, n):
        return cls._wrap_device(dev, n)

    def _para_field(self, name):
        a = self.copy_as_array()["para_rays"][name]
        return np.array(a)

    para_origin = property(lambda self: self._para_field("origin"))
    para_direction = property(lambda self: self._para_field("direction"))
    para_normal = property(lambda self: self._para_field("normal"))

    @property
    def para_termination(self):
        a = self.copy_as_array()["para_rays"]
        return a["origin"] + a["direction"] * a["length"][..., None]

    @property
    def total_power(self):
        b = self.copy_as_array()["base_ray"]
        n = b["refractive_index"].real
        e1, e2 = b["E1_amp"], b["E2_amp"]
        return float(np.sum((e1.real * e1.real) * n + (e1.imag * e1.imag) * n +
                            (e2.real * e2.real) * n + (e2.imag * e2.imag) * n))

    def scale_amplitude(self, scale):
        self._ensure_host()
        self._host["base_ray"]["E1_amp"][: self._n] *= scale
        self._host["base_ray"]["E2_amp"][: self._n] *= scale
        self._host_changed()

    def project_to_plane(self, origin, direction):
        """Move every base and parabasal ray along its direction to the plane
        through `origin` with normal `direction` (ctracer.pyx:1391-1420), on
        the GPU."""
        lib = load()
        if self._n == 0:
            return
        s = self._ensure_device().struct()
        o = (ctypes.c_double * 3)(*[float(x) for x in origin])
        d = (ctypes.c_double * 3)(*[float(x) for x in direction])
        check(lib.rpx_project_to_plane(ctypes.byref(s), self._n, o, d, stream_handle()))
        self._device_changed()

    @property
    def lagrange_invariant(self):
        """Per-gausslet Lagrange invariant of the parabasal ray set,
        1000 sqrt(v / 6) / wavelength (ctracer.pyx:1347-1389), computed on the
        GPU; a float64 numpy array."""
        torch = require_gpu()
        lib = load()
        n = self._n
        if n == 0:
            return np.empty((0,), dtype=np.float64)
        wl = self._device_wavelengths()
        out = torch.empty(n, dtype=torch.float64, device="cuda")
        ws = torch.empty(8, dtype=torch.uint8, device="cuda")
        s = self._ensure_device().struct()
        check(lib.rpx_lagrange_invariant(ctypes.byref(s), n, ctypes.c_void_p(wl.data_ptr()), wl.numel(),
                                         ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(ws.data_ptr()), ws.numel(),
                                         stream_handle()))
        return out.cpu().numpy()

    def config_parabasal_rays(self, wavelength_list, radius, working_dist):
        """Place the parabasal rays of a circular Gaussian mode of 1/e^2 radius
        `radius` (mm) with its waist `working_dist` from the base origin
        (ctracer.pyx:1430-1481), for all gausslets at once on the GPU."""
        torch = require_gpu()
        lib = load()
        if self._n == 0:
            return
        wl = torch.from_numpy(np.ascontiguousarray(wavelength_list, dtype=np.float64)).to("cuda")
        s = self._ensure_device().struct()
        trig = []
        for j in range(0, 6, 2):
            angle = (j * 2 * math.pi / 6)
            trig += [math.cos(angle), math.sin(angle)]
        trig = (ctypes.c_double * 6)(*trig)
        ws = torch.empty(8, dtype=torch.uint8, device="cuda")
        check(lib.rpx_config_parabasal_rays(ctypes.byref(s), self._n, ctypes.c_void_p(wl.data_ptr()), wl.numel(),
                                            float(radius), float(working_dist), trig, ctypes.c_void_p(ws.data_ptr()),
                                            ws.numel(), stream_handle()))
        self._device_changed()


# ---------------------------------------------------------- plugin bases
class InterfaceMaterial(object):
    """Abstract material: produces no child rays (ctracer.pyx:1573-1656)."""

    def __init__(self, **kwds):
        self._wavelengths = np.array([], dtype=np.double)

    def _rpx_fill_material(self, rec, builder):
        rec["type"] = _abi.MAT_NONE

    def is_decomp_material(self):
        return False

    @property
    def wavelengths(self):
        return self._wavelengths

    @wavelengths.setter
    def wavelengths(self, wavelengths):
        self._wavelengths = np.ascontiguousarray(wavelengths, dtype=np.double)
        self.on_set_wavelengths()

    def on_set_wavelengths(self):
        pass

    def eval_child_ray(self, old_ray, ray_idx, point, normal, tangent, new_rays):
        """Evaluate this material for one ray on the GPU; children are
        appended to new_rays (ctracer.pyx:1620-1632)."""
        from ._queries import eval_child_rays
        for child in eval_child_rays(self, old_ray, ray_idx, point, normal, tangent):
            new_rays.add_ray(child)


class Shape(object):
    """Base aperture: every point is inside (ctracer.pyx:1659-1664)."""

    def _rpx_shape_ops(self, rec, builder):
        rec["op"] = _abi.SHAPE_TRUE

    def point_inside(self, x, y):
        from ._queries import shape_inside
        return bool(shape_inside(self, [x], [y])[0])


class Distortion(object):
    """Abstract surface distortion (ctracer.pyx:1682-1729)."""


class Face(object):
    """Base of all faces (ctracer.pyx:1732-1809)."""
    params = []
    _face_type = None

    def __init__(self, owner=None, tolerance=0.0001, max_length=100, material=None, **kwds):
        self.name = "base Face class"
        self.tolerance = tolerance
        self.owner = owner
        self.max_length = max_length
        if isinstance(material, InterfaceMaterial):
            self.material = material
        else:
            from .cmaterials import PECMaterial
            self.material = PECMaterial()
        self.invert_normal = int(kwds.get("invert_normal", 0))
        self.idx = 0
        self.count = 0

    def update(self):
        """Pull the `params` attributes from the owner (ctracer.pyx:1761-1767).
        A face without an owner keeps its own attributes (the reference
        raises AttributeError there)."""
        if self.owner is None:
            return
        for name in self.params:
            setattr(self, name, getattr(self.owner, name))

    def _rpx_fill_face(self, rec, builder):
        raise NotImplementedError

    def intersect(self, p1, p2, is_base_ray):
        from ._queries import face_intersect
        return float(face_intersect(self, [p1], [p2], is_base_ray)[0])

    def compute_normal(self, p):
        from ._queries import face_normal
        return tuple(float(v) for v in face_normal(self, [p])[0])

    def compute_tangent(self, p):
        return (1.0, 0.0, 0.0)


class FaceList(object):
    """A group of faces sharing one transform (ctracer.pyx:1813-1964)."""

    def __init__(self, owner=None):
        self.transform = Transform()
        self.inverse_transform = Transform()
        self.owner = owner
        self.faces = []

    def sync_transforms(self):
        """Set the transforms from the owner's VTK-style transform, if any."""
        trans = getattr(self.owner, "transform", None)
        if trans is None or not hasattr(trans, "matrix"):
            return
        m = trans.matrix
        rot = [[m.get_element(i, j) for j in range(3)] for i in range(3)]
        dt = [m.get_element(i, 3) for i in range(3)]
        self.transform = Transform(rotation=rot, translation=dt)
        m = trans.linear_inverse.matrix
        rot = [[m.get_element(i, j) for j in range(3)] for i in range(3)]
        dt = [m.get_element(i, 3) for i in range(3)]
        self.inverse_transform = Transform(rotation=rot, translation=dt)

    @property
    def transform(self):
        t = Transform()
        t._rot, t._t = self._trans[0].copy(), self._trans[1].copy()
        return t

    @transform.setter
    def transform(self, t):
        self._trans = (t._rot.copy(), t._t.copy())

    @property
    def inverse_transform(self):
        t = Transform()
        t._rot, t._t = self._inv_trans[0].copy(), self._inv_trans[1].copy()
        return t

    @inverse_transform.setter
    def inverse_transform(self, t):
        self._inv_trans = (t._rot.copy(), t._t.copy())

    def __getitem__(self, idx):
        return self.faces[idx]

    def intersect(self, r):
        """Nearest face hit by Ray r within its current length; updates
        r.length / r.end_face_idx and returns the face idx or -1."""
        from ._queries import face_intersect
        end = addvv(r.origin, multvs(r.direction, r.length))
        p1 = transform(self.inverse_transform, r.origin)
        p2 = transform(self.inverse_transform, end)
        out = -1
        for face in self.faces:
            dist = float(face_intersect(face, [p1], [p2], 1)[0])
            if face.tolerance < dist < r.length:
                r.length = dist
                r.end_face_idx = face.idx
                out = face.idx
        return out

    def compute_orientation(self, face, point, piece=0):
        from ._queries import face_normal
        p = transform(self.inverse_transform, point)
        n = tuple(face_normal(face, [p])[0])
        t = face.compute_tangent(p)
        if face.invert_normal:
            n, t = invert(n), invert(t)
        rt = Transform(rotation=self.transform.rotation)
        return transform(rt, n), transform(rt, t)


# ------------------------------------------------------------ tracing
class _TraceContext(object):
    """A compiled, uploaded scene plus per-face hit bookkeeping."""

    def __init__(self, face_sets, all_faces):
        from ._scene import compile_scene
        self.face_sets = list(face_sets)
        self.all_faces = all_faces
        self.tables = compile_scene(self.face_sets)
        self.handle = self.tables.handle()
        # the same count as SceneTables.max_children, read from the uploaded
        # scene (a numpy walk over the face records costs tens of us a trace)
        self.max_children = int(load().rpx_scene_max_children(self.handle))
        self._hits = None
        self._ws = None

    @property
    def hits(self):
        """Device per-face hit counters of the one-call-per-generation path."""
        if self._hits is None:
            torch = require_gpu()
            self._hits = torch.zeros(max(self.tables.n_global_faces, 1), dtype=torch.int32, device="cuda")
        return self._hits

    def workspace(self, n):
        lib = load()
        need = int(lib.rpx_trace_workspace_bytes(n))
        if self._ws is None or self._ws.numel() < need:
            torch = require_gpu()
            self._ws = torch.empty(need, dtype=torch.uint8, device="cuda")
        return self._ws

    def chain_workspace(self, nbytes):
        if self._ws is None or self._ws.numel() < nbytes:
            torch = require_gpu()
            self._ws = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        return self._ws

    def flush_counts(self):
        """Add the device hit counters to Face.count of all_faces."""
        if self._hits is None:
            return
        h = self.hits.cpu().numpy()
        for g in np.nonzero(h)[0]:
            if g < len(self.all_faces):
                self.all_faces[g].count += int(h[g])
        self.hits.zero_()


def _wavelength_table(rays, ctx):
    wl = rays.wavelengths
    if len(wl) == 0:
        for m in ctx.tables.material_objects:
            if len(getattr(m, "_wavelengths", ())):
                wl = m._wavelengths
                rays.wavelengths = wl
                break
    return rays._device_wavelengths()


def _capture(rays, face_ids, para_length, captured):
    """Append to `captured` (GaussletCollection) the gausslets of `rays`
    whose base ray ended on one of face_ids, with the parabasal lengths
    they had before the generation (rpx_capture_gausslets)."""
    from ._device import DeviceRays
    torch = require_gpu()
    lib = load()
    n = rays.n_rays
    ids = (ctypes.c_int32 * len(face_ids))(*face_ids)
    out = DeviceRays(max(n, 1), gausslet=True)
    ws = torch.empty(max(int(lib.rpx_trace_workspace_bytes(n)), 8), dtype=torch.uint8, device="cuda")
    k = ctypes.c_uint64(0)
    s_in, s_out = rays._ensure_device().struct(), out.struct()
    check(lib.rpx_capture_gausslets(ctypes.byref(s_in), n, ids, len(face_ids), ctypes.c_void_p(para_length.data_ptr()),
                                    int(para_length.shape[1]), ctypes.byref(s_out), 0, ctypes.byref(k),
                                    ctypes.c_void_p(ws.data_ptr()), ws.numel(), stream_handle()))
    if k.value:
        captured.extend(GaussletCollection.from_device(out, k.value))
    return k.value


def _decompose(ctx, rays, out, para_length, decomp_faces, max_length):
    """The decomposition step of trace_gausslet_c (ctracer.pyx:2274-2280):
    gausslets that ended on a decomposition face were captured by its
    material; every decomposition face hit (Face.count) makes its material
    append its eval_func's gausslets to the generation, and its count is
    cleared."""
    idx = [f.idx for f in decomp_faces]
    hits = ctx.hits[idx].cpu().numpy().astype(np.int64)
    ctx.hits[idx] = 0
    mats = []
    for f in decomp_faces:
        if not any(f.material is m for m in mats):
            mats.append(f.material)
    for m in mats:
        ids = [f.idx for f in decomp_faces if f.material is m]
        got = _capture(rays, ids, para_length, m.captured_rays)
        m.capture_count += got
    extended = False
    for f, h in zip(decomp_faces, hits):
        if f.count + int(h):
            f.material.eval_decomposed_rays(out)
            extended = True
        f.count = 0
    if extended:
        out.reset_length(max_length)


def _generation(ctx, rays, max_length, one_list=-1, one_face=-1, flush=True, decomp_faces=()):
    """Trace one generation on the GPU; returns the child collection."""
    from ._device import DeviceRays
    lib = load()
    gauss = isinstance(rays, GaussletCollection)
    cls = GaussletCollection if gauss else RayCollection
    n = rays.n_rays
    if n == 0:
        out = cls(0)
        out.parent = rays
        return out
    dev_in = rays._ensure_device()
    decomp = list(decomp_faces) if gauss else []
    para_length = dev_in.para_view()[:, 9, :n].clone() if decomp else None
    wl = _wavelength_table(rays, ctx)
    cap = max(ctx.max_children * n, 1)
    dev_out = DeviceRays(cap, gausslet=gauss)
    ws = ctx.workspace(n)
    s_in, s_out = dev_in.struct(), dev_out.struct()
    n_out = ctypes.c_uint64(0)
    args = (ctypes.byref(s_in), n)
    tail = (ctypes.c_void_p(wl.data_ptr()), len(rays.wavelengths), ctypes.byref(s_out), ctypes.byref(n_out),
            ctypes.c_void_p(ctx.hits.data_ptr()), ctypes.c_void_p(ws.data_ptr()), ws.numel(), stream_handle())
    if gauss:
        code = lib.rpx_trace_gausslet(ctx.handle, one_list, one_face, *args, float(max_length), *tail)
    elif one_list < 0:
        code = lib.rpx_trace_segment(ctx.handle, *args, float(max_length), *tail)
    else:
        code = lib.rpx_trace_one_face_segment(ctx.handle, one_list, one_face, *args, float(max_length), *tail)
    check(code)
    rays._device_changed()
    out = cls.from_device(dev_out, n_out.value)
    out.parent = rays
    if decomp:
        _decompose(ctx, rays, out, para_length, decomp, max_length)
    if flush:
        ctx.flush_counts()
    return out


def trace_segment(rays, face_sets, all_faces, max_length=100, decomp_faces=[]):
    """One non-sequential generation (ctracer.pyx:2173-2182)."""
    for fs in face_sets:
        fs.sync_transforms()
    ctx = _TraceContext(face_sets, all_faces)
    return _generation(ctx, rays, max_length)


def trace_one_face_segment(rays, face_set, face_idx, all_faces, max_length=100, decomp_faces=[]):
    """One generation against a single face (ctracer.pyx:2184-2191)."""
    face_set.sync_transforms()
    ctx = _TraceContext([face_set], all_faces)
    return _generation(ctx, rays, max_length, 0, face_idx)


def trace_gausslet(rays, face_sets, all_faces, max_length=100, decomp_faces=[]):
    """One gausslet generation (ctracer.pyx:2193-2202)."""
    for fs in face_sets:
        fs.sync_transforms()
    ctx = _TraceContext(face_sets, all_faces)
    return _generation(ctx, rays, max_length, decomp_faces=decomp_faces)


def trace_one_face_gausslet(rays, face_set, face_idx, all_faces, max_length=100, decomp_faces=[]):
    face_set.sync_transforms()
    ctx = _TraceContext([face_set], all_faces)
    return _generation(ctx, rays, max_length, 0, face_idx, decomp_faces=decomp_faces)


def _propagate_neighbours(children, parent_nb):
    """Neighbours of `children` from its parent generation's table (device
    int32 (n_parent, K)): RayCollection._eval_neighbours, ctracer.pyx:1084-1106."""
    torch = require_gpu()
    lib = load()
    n, K = children.n_rays, int(parent_nb.shape[1]) if parent_nb.dim() == 2 else 0
    out = torch.full((n, K), -1, dtype=torch.int32, device="cuda")
    if n == 0 or K == 0:
        return out
    parent_nb = parent_nb.contiguous()
    s = children._ensure_device().struct()
    n_par = int(parent_nb.shape[0])
    ws = torch.empty(int(lib.rpx_neighbours_workspace_bytes(n_par)), dtype=torch.uint8, device="cuda")
    check(lib.rpx_propagate_neighbours(ctypes.byref(s), n, ctypes.c_void_p(parent_nb.data_ptr()), n_par, K,
                                       ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(ws.data_ptr()), ws.numel(),
                                       stream_handle()))
    return out


# ------------------------------------------------------ capture selection
def wavelength_maps(wl_list):
    """(sorted union of the wavelength tables, one uint32 index map per
    table): wavelength j of table t becomes union[maps[t][j]] -- the
    np.unique(..., return_inverse=True) re-indexing of ctracer.pyx:2015-2021,
    split per collection."""
    wl_list = [np.asarray(w, np.float64).reshape(-1) for w in wl_list]
    if not wl_list:
        return np.zeros(0, np.float64), []
    reduced, inverse = np.unique(np.concatenate(wl_list), return_inverse=True)
    inverse = np.asarray(inverse).reshape(-1).astype(np.uint32)
    maps, off = [], 0
    for w in wl_list:
        maps.append(inverse[off: off + len(w)].copy())
        off += len(w)
    return reduced, maps


def _select_intersections(face_set, ray_col_list, gauss):
    """Rays (gausslets) of every collection whose segment meets `face_set`,
    concatenated in list order, with wavelengths re-indexed into the sorted
    union (ctracer.pyx:1981-2058)."""
    from ._device import DeviceRays
    from ._scene import compile_scene
    lib = load()
    torch = require_gpu()
    cls = GaussletCollection if gauss else RayCollection
    ray_col_list[0]   # the reference sizes its output by the first collection
    wl_list = [np.asarray(rc.wavelengths, np.float64) for rc in ray_col_list]
    reduced, maps = wavelength_maps(wl_list)
    inverse = np.concatenate(maps) if maps else np.zeros(0, np.uint32)
    face_set.sync_transforms()
    tables = compile_scene([face_set])
    handle = tables.handle()
    total = sum(rc.n_rays for rc in ray_col_list)
    dev_out = DeviceRays(max(total, 1), gausslet=gauss)
    s_out = dev_out.struct()
    wl_map = torch.from_numpy(inverse.astype(np.int32)).to("cuda") if len(inverse) else None
    need = int(lib.rpx_trace_workspace_bytes(max((rc.n_rays for rc in ray_col_list), default=0)))
    ws = torch.empty(max(need, 8), dtype=torch.uint8, device="cuda")
    n_sel, off = 0, 0
    for rc, wl in zip(ray_col_list, wl_list):
        n = rc.n_rays
        if n:
            s_in = rc._ensure_device().struct()
            m_ptr = ctypes.c_void_p(wl_map.data_ptr() + 4 * off) if wl_map is not None and len(wl) else None
            k = ctypes.c_uint64(0)
            check(lib.rpx_select_intersections(handle, 0, ctypes.byref(s_in), n, int(gauss), m_ptr, len(wl),
                                               ctypes.byref(s_out), n_sel, ctypes.byref(k),
                                               ctypes.c_void_p(ws.data_ptr()), ws.numel(), stream_handle()))
            n_sel += k.value
        off += len(wl)
    out = cls.from_device(dev_out, n_sel)
    out.wavelengths = reduced
    return out


def select_ray_intersections(face_set, ray_col_list):
    """Copies of the rays of every RayCollection in ray_col_list whose own
    segment (origin to origin + direction * length) meets a face of
    face_set, with length and end_face_idx set to the nearest such face and
    wavelength_idx re-indexed into the sorted union of the collections'
    wavelengths (ctracer.pyx:1981-2021).  The collections are not modified."""
    return _select_intersections(face_set, ray_col_list, False)


def select_gausslet_intersections(face_set, ray_col_list):
    """select_ray_intersections for GaussletCollections: the base ray decides
    and is updated; parabasal rays are copied unchanged (ctracer.pyx:2024-2058)."""
    return _select_intersections(face_set, ray_col_list, True)
