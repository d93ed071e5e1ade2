"""Interface materials (drop-in for the device-supported subset of
raypier.core.cmaterials).

Each material keeps the reference's keywords and attributes and compiles to
one rpx_material_t record.  Child rays are produced on the GPU
(rpx_materials.h).  Dispersion curves are evaluated inside the kernels from
their coefficients; the host-side `n_inside` etc. tables of the dispersive
materials are computed on demand for inspection only.
"""
import numpy as np

from . import _abi
from .ctracer import InterfaceMaterial


_extent_cache = {"key": None, "extent": None}


def wavelength_extent(wl):
    """(min, max) of a wavelength array, cached for the most recent array
    object so that repeated traces of one collection scan it once."""
    wl = np.asarray(wl, dtype=np.double)
    key = (id(wl), wl.ctypes.data if wl.size else 0, wl.size, wl.strides)
    if _extent_cache["key"] != key:
        ext = (float(wl.min()), float(wl.max())) if wl.size else (np.inf, -np.inf)
        _extent_cache["key"], _extent_cache["extent"] = key, ext
    return _extent_cache["extent"]


class BaseDispersionCurve(object):
    """n(wavelength) by formula 0/1/2/3/5 (cmaterials.pyx:152-239)."""

    def __init__(self, formula_id, coefs, absorption=0.0, wavelength_min=0.1, wavelength_max=1000000.0):
        coefs = np.ascontiguousarray(coefs, dtype=np.double)
        if formula_id not in (0, 1, 2, 3, 5):
            raise ValueError("Unknown formula id (%d)" % (formula_id,))
        if len(coefs) > _abi.MAX_DISP_COEFS:
            raise ValueError("at most %d dispersion coefficients" % _abi.MAX_DISP_COEFS)
        self.coefs = coefs
        self.formula_id = formula_id
        self.absorption = absorption
        self.wavelength_min = wavelength_min
        self.wavelength_max = wavelength_max

    def check_range(self, wavelen):
        """Raise ValueError if any wavelength is outside the curve's range
        (c_evaluate_n, cmaterials.pyx:218-224)."""
        lo, hi = wavelength_extent(wavelen)
        if lo < self.wavelength_min or hi > self.wavelength_max:
            wl = np.asarray(wavelen, dtype=np.double)
            bad = (wl < self.wavelength_min) | (wl > self.wavelength_max)
            wvl = float(wl[bad][0])
            raise ValueError("Wavelength (%f) outside range of dispersion curve (%f -> %f)" %
                             (wvl, self.wavelength_min, self.wavelength_max))

    def evaluate_n(self, wavelen):
        """Complex refractive index at the given wavelengths (microns)."""
        wl = np.ascontiguousarray(wavelen, dtype=np.double)
        self.check_range(wl)
        c = self.coefs
        if self.formula_id == 0:
            nre = np.full(wl.shape, c[0])
        elif self.formula_id in (1, 2):
            n2 = np.full(wl.shape, 1.0)
            wl2 = wl * wl
            n2 = n2 + c[0]
            for i in range((len(c) - 1) // 2):
                ci = c[2 * i + 2] * c[2 * i + 2] if self.formula_id == 1 else c[2 * i + 2]
                n2 = n2 + c[2 * i + 1] * wl2 / (wl2 - ci)
            nre = np.sqrt(n2)
        elif self.formula_id == 3:
            n2 = np.full(wl.shape, c[0])
            for i in range(1, len(c) - 1, 2):
                n2 = n2 + c[i] * (wl ** c[i + 1])
            nre = np.sqrt(n2)
        else:   # sellmeier_5 returns no value in the reference (cmaterials.pyx:143-149)
            nre = np.zeros(wl.shape)
        n_imag = 0.0001 * self.absorption / (4 * np.pi)
        return nre + 1j * (n_imag * wl)

    def _fill(self, rec):
        rec["formula_id"] = self.formula_id
        rec["n_coefs"] = len(self.coefs)
        rec["absorption"] = self.absorption
        rec["coefs"][: len(self.coefs)] = self.coefs


class NondispersiveCurve(BaseDispersionCurve):
    """Constant refractive index (raypier/dispersion.py:21-40)."""

    def __init__(self, refractive_index=1.37, absorption=0.0):
        self._refractive_index = refractive_index
        self._absorption = absorption
        super().__init__(0, np.array([refractive_index]), absorption, 0.0, 1000000.0)


vacuum = BaseDispersionCurve(0, np.array([1.0]))


def _cplx(v):
    v = complex(v)
    return (v.real, v.imag)


class OpaqueMaterial(InterfaceMaterial):
    """A perfect absorber (cmaterials.pyx:242-251)."""

    def _rpx_fill_material(self, rec, b):
        rec["type"] = _abi.MAT_OPAQUE


class ResampleGaussletMaterial(InterfaceMaterial):
    """A pseudo-material that captures the incident gausslets and, at the
    end of the generation, launches the gausslets returned by
    eval_func(captured_rays) instead of refracting or reflecting them
    (cmaterials.pyx:1766-1831).  On the device it produces no children; the
    capture is a compaction of the generation's gausslets that ended on its
    faces (rpx_capture_gausslets)."""

    def __init__(self, size=2, eval_func=None, **kwds):
        super().__init__(**kwds)
        from .ctracer import GaussletCollection
        self.capture_count = 0
        self.captured_rays = GaussletCollection(size)
        self._evaluation_func = None
        if eval_func is not None:
            self.eval_func = eval_func

    @property
    def eval_func(self):
        return self._evaluation_func

    @eval_func.setter
    def eval_func(self, obj):
        if not callable(obj):
            raise ValueError("The eval_func property must be a callable.")
        self._evaluation_func = obj

    def is_decomp_material(self):
        return True

    def _rpx_fill_material(self, rec, b):
        rec["type"] = _abi.MAT_OPAQUE

    def eval_decomposed_rays(self, new_rays):
        """eval_decomposed_rays_c (cmaterials.pyx:1824-1835)."""
        self.captured_rays._wavelengths = new_rays._wavelengths
        self.captured_rays._wl_dev = None
        gc = self._evaluation_func(self.captured_rays)
        new_rays.extend(gc)
        self.captured_rays.clear_ray_list()


class TransparentMaterial(InterfaceMaterial):
    """Passes the ray unchanged apart from the S/P projection (cmaterials.pyx:254-278)."""

    def _rpx_fill_material(self, rec, b):
        rec["type"] = _abi.MAT_TRANSPARENT


class PECMaterial(InterfaceMaterial):
    """Perfect electrical conductor mirror (cmaterials.pyx:281-319)."""

    def _rpx_fill_material(self, rec, b):
        rec["type"] = _abi.MAT_PEC


class PartiallyReflectiveMaterial(InterfaceMaterial):
    """Fixed power reflectivity (cmaterials.pyx:322-397)."""

    def __init__(self, **kwds):
        super().__init__()
        self.reflectivity = kwds.get("reflectivity", 0.5)

    @property
    def reflectivity(self):
        return self._reflectivity

    @reflectivity.setter
    def reflectivity(self, val):
        if val < 0.0 or val > 1.0:
            raise ValueError(f"Reflectivity must be in range 0.0 to 1.0 ({val} given).")
        self._reflectivity = val

    def _rpx_fill_material(self, rec, b):
        rec["type"] = _abi.MAT_PARTIAL
        rec["reflectivity"] = self._reflectivity


class DielectricMaterial(InterfaceMaterial):
    """Real-index Snell/Fresnel transmission with TIR (cmaterials.pyx:554-724)."""
    _type = _abi.MAT_DIELECTRIC

    def __init__(self, **kwds):
        super().__init__()
        self.n_inside = kwds.get('n_inside', 1.5)
        self.n_outside = kwds.get('n_outside', 1.0)

    n_inside = property(lambda s: complex(*s._n_in), lambda s, v: setattr(s, "_n_in", _cplx(v)))
    n_outside = property(lambda s: complex(*s._n_out), lambda s, v: setattr(s, "_n_out", _cplx(v)))

    def _rpx_fill_material(self, rec, b):
        rec["type"] = self._type
        rec["n_inside"] = self._n_in
        rec["n_outside"] = self._n_out


class FullDielectricMaterial(DielectricMaterial):
    """Complex-index Fresnel reflection and transmission (cmaterials.pyx:727-872)."""
    _type = _abi.MAT_FULL_DIELECTRIC

    def __init__(self, **kwds):
        super().__init__(**kwds)
        self.n_coating = kwds.get("n_coating", 1.0)
        self.thickness = kwds.get("thickness", 0.1)
        self.reflection_threshold = kwds.get('reflection_threshold', 0.1)
        self.transmission_threshold = kwds.get('transmission_threshold', 0.1)

    n_coating = property(lambda s: complex(*s._n_coat), lambda s, v: setattr(s, "_n_coat", _cplx(v)))

    def _rpx_fill_material(self, rec, b):
        super()._rpx_fill_material(rec, b)
        rec["n_coating"] = self._n_coat
        rec["thickness"] = self.thickness
        rec["reflection_threshold"] = self.reflection_threshold
        rec["transmission_threshold"] = self.transmission_threshold


class SingleLayerCoatedMaterial(FullDielectricMaterial):
    """Single-layer thin-film coating (cmaterials.pyx:1017-1182).  The coating
    thickness is the `thickness` keyword (microns)."""
    _type = _abi.MAT_SINGLE_LAYER


class _DispersiveBase(InterfaceMaterial):
    _curves = ()

    def on_set_wavelengths(self):
        for name in self._curves:
            getattr(self, name).check_range(self._wavelengths)

    def _table(self, name):
        return getattr(self, name).evaluate_n(self._wavelengths)


class FullDielectricDispersiveMaterial(_DispersiveBase):
    """Dispersive uncoated interface (cmaterials.pyx:875-1013)."""
    _curves = ("dispersion_inside", "dispersion_outside")

    def __init__(self, **kwds):
        super().__init__()
        self.dispersion_inside = kwds.get("dispersion_inside", vacuum)
        self.dispersion_outside = kwds.get("dispersion_outside", vacuum)
        self.reflection_threshold = kwds.get('reflection_threshold', 0.1)
        self.transmission_threshold = kwds.get('transmission_threshold', 0.1)

    n_inside = property(lambda s: s._table("dispersion_inside"))
    n_outside = property(lambda s: s._table("dispersion_outside"))

    def _rpx_fill_material(self, rec, b):
        rec["type"] = _abi.MAT_FULL_DISPERSIVE
        self.dispersion_inside._fill(rec["disp_inside"])
        self.dispersion_outside._fill(rec["disp_outside"])
        rec["reflection_threshold"] = self.reflection_threshold
        rec["transmission_threshold"] = self.transmission_threshold


class CoatedDispersiveMaterial(_DispersiveBase):
    """Dispersive interface with a single-layer coating (cmaterials.pyx:1187-1434)."""
    _curves = ("dispersion_inside", "dispersion_outside", "dispersion_coating")

    def __init__(self, **kwds):
        super().__init__()
        self.reflection_threshold = kwds.get('reflection_threshold', 0.1)
        self.transmission_threshold = kwds.get('transmission_threshold', 0.1)
        self.dispersion_inside = kwds.get("dispersion_inside", vacuum)
        self.dispersion_outside = kwds.get("dispersion_outside", vacuum)
        self.dispersion_coating = kwds.get("dispersion_coating", vacuum)
        self.coating_thickness = kwds.get("coating_thickness", 0.1)

    n_inside = property(lambda s: s._table("dispersion_inside"))
    n_outside = property(lambda s: s._table("dispersion_outside"))
    n_coating = property(lambda s: s._table("dispersion_coating"))

    def _rpx_fill_material(self, rec, b):
        rec["type"] = _abi.MAT_COATED_DISPERSIVE
        self.dispersion_inside._fill(rec["disp_inside"])
        self.dispersion_outside._fill(rec["disp_outside"])
        self.dispersion_coating._fill(rec["disp_coating"])
        rec["thickness"] = self.coating_thickness
        rec["reflection_threshold"] = self.reflection_threshold
        rec["transmission_threshold"] = self.transmission_threshold


class DiffractionGratingMaterial(InterfaceMaterial):
    """Reflective grating, one diffraction order (cmaterials.pyx:1437-1599).

    orders=(o1, o2[, o3]) -- not in the reference, whose material has one
    `order` -- makes one grating emit up to three orders: one child per
    propagating order, in the listed order, each exactly the child a
    one-order grating of that order emits (same efficiency for all).  An
    evanescent order emits nothing, as in the reference."""

    def __init__(self, **kwds):
        super().__init__()
        self.lines_per_mm = kwds.get("lines_per_mm", 1000)
        self.order = kwds.get("order", 1)
        self.efficiency = kwds.get("efficiency", 1.0)
        self.origin = kwds.get("origin", (0.0, 0.0, 0.0))
        orders = kwds.get("orders")
        if orders is not None:
            orders = tuple(int(o) for o in orders)
            if not 1 <= len(orders) <= 3 or any(not -128 <= o <= 127 for o in orders):
                raise ValueError("orders: one to three orders in -128..127")
            self.order = orders[0]
        self.orders = orders

    def max_children(self):
        return len(self.orders) if self.orders else 1

    @property
    def origin(self):
        return self._origin

    @origin.setter
    def origin(self, o):
        self._origin = (float(o[0]), float(o[1]), float(o[2]))

    def _rpx_fill_material(self, rec, b):
        rec["type"] = _abi.MAT_GRATING
        rec["lines_per_mm"] = self.lines_per_mm
        rec["order"] = int(self.order)
        if self.orders and len(self.orders) > 1:
            rec["n_orders"] = len(self.orders)
            more = [o & 0xFF for o in self.orders[1:]]
            rec["more_orders"] = more[0] | ((more[1] if len(more) > 1 else 0) << 8)
        rec["efficiency"] = self.efficiency
        rec["origin"] = self._origin
