"""numpy / ctypes mirrors of the structs declared in include/rpx.h.

The record layouts are plain data; the same tables feed the HIP library and
the CPU oracle used by the tests.
"""
import ctypes
import numpy as np

# ray record layouts of raypier/core/ctracer.pyx:45-76 (packed, 188 / 80 / 668 bytes)
ray_dtype = np.dtype([('origin', np.double, (3,)),
                      ('direction', np.double, (3,)),
                      ('normal', np.double, (3,)),
                      ('E_vector', np.double, (3,)),
                      ('refractive_index', np.complex128),
                      ('E1_amp', np.complex128),
                      ('E2_amp', np.complex128),
                      ('length', np.double),
                      ('phase', np.double),
                      ('accumulated_path', np.double),
                      ('wavelength_idx', np.uint32),
                      ('parent_idx', np.uint32),
                      ('end_face_idx', np.uint32),
                      ('ray_ident', np.uint32),
                      ('ray_type_id', np.uint32)])

para_dtype = np.dtype([('origin', np.double, (3,)),
                       ('direction', np.double, (3,)),
                       ('normal', np.double, (3,)),
                       ('length', np.double)])

gausslet_dtype = np.dtype([('base_ray', ray_dtype),
                           ('para_rays', para_dtype, (6,))])

assert ray_dtype.itemsize == 188 and para_dtype.itemsize == 80 and gausslet_dtype.itemsize == 668

ABI_VERSION = 11          # RPX_ABI_VERSION of include/rpx.h

# timing slots (RPX_TIMING_*)
TIMING_KERNELS = ("k_intersect", "k_scan", "k_emit", "k_generation", "k_clear", "k_generation_split",
                  "k_field_sum", "k_field_sum_split", "k_gausslet_generation")

REFL_RAY = 1 << 0
GAUSSLET = 1 << 1
PARABASAL = 1 << 2
NO_FACE = 0xFFFFFFFF

N_F64 = 21
N_U32 = 5
N_PARA = 6
PARA_FIELDS = 10

# SoA field order (rpx_ray_f64_field / rpx_ray_u32_field)
F64_FIELDS = ["ox", "oy", "oz", "dx", "dy", "dz", "nx", "ny", "nz", "ex", "ey", "ez",
              "n_re", "n_im", "e1_re", "e1_im", "e2_re", "e2_im", "length", "phase", "accum"]
U32_FIELDS = ["wavelength_idx", "parent_idx", "end_face_idx", "ray_ident", "ray_type_id"]

# --------------------------------------------------------------- enums
FACE_CIRCULAR, FACE_SHAPED_PLANAR, FACE_RECTANGULAR, FACE_POLYGON, FACE_SPHERICAL, \
    FACE_SHAPED_SPHERICAL, FACE_EXTRUDED_PLANAR, FACE_CONIC, FACE_ASPHERIC, FACE_DISTORTION = range(1, 11)

SHAPE_TRUE, SHAPE_CIRCLE, SHAPE_RECT, SHAPE_POLYGON, SHAPE_NOT, SHAPE_AND, SHAPE_OR, SHAPE_XOR = range(1, 9)

MAT_NONE, MAT_OPAQUE, MAT_TRANSPARENT, MAT_PEC, MAT_PARTIAL, MAT_DIELECTRIC, MAT_FULL_DIELECTRIC, \
    MAT_FULL_DISPERSIVE, MAT_SINGLE_LAYER, MAT_COATED_DISPERSIVE, MAT_GRATING = range(0, 11)

RPX_OK, RPX_E_INVALID, RPX_E_UNSUPPORTED, RPX_E_CAPACITY, RPX_E_HIP, RPX_E_NO_DEVICE = range(6)

RPX_FIELD_SCRATCH_BYTES = 128 << 20   # static scratch of rpx_sum_gaussian_modes_noalloc

FACE_NPARAM = 16
MAX_DISP_COEFS = 18
MAX_ZERNIKE_SLOTS = 96

# --------------------------------------------------------------- records
transform_dtype = np.dtype([('m', np.float64, (9,)), ('t', np.float64, (3,))])

face_list_dtype = np.dtype([('trans', transform_dtype), ('inv_trans', transform_dtype),
                            ('first_face', np.int32), ('n_faces', np.int32), ('pad', np.int32, (2,))])

face_dtype = np.dtype([('type', np.int32), ('global_idx', np.int32), ('face_list', np.int32),
                       ('material', np.int32), ('invert_normal', np.int32), ('invert_normals', np.int32),
                       ('shape', np.int32), ('aux0', np.int32), ('aux1', np.int32), ('pad', np.int32),
                       ('tolerance', np.float64), ('p', np.float64, (FACE_NPARAM,))])

shape_op_dtype = np.dtype([('op', np.int32), ('a', np.int32), ('b', np.int32), ('begin', np.int32),
                           ('p', np.float64, (4,))])

dispersion_dtype = np.dtype([('formula_id', np.int32), ('n_coefs', np.int32), ('absorption', np.float64),
                             ('coefs', np.float64, (MAX_DISP_COEFS,))])

material_dtype = np.dtype([('type', np.int32), ('order', np.int32), ('n_orders', np.int32), ('more_orders', np.int32),
                           ('reflection_threshold', np.float64), ('transmission_threshold', np.float64),
                           ('n_inside', np.float64, (2,)), ('n_outside', np.float64, (2,)),
                           ('n_coating', np.float64, (2,)), ('thickness', np.float64),
                           ('reflectivity', np.float64), ('lines_per_mm', np.float64),
                           ('efficiency', np.float64), ('origin', np.float64, (3,)), ('pad2', np.float64),
                           ('disp_inside', dispersion_dtype), ('disp_outside', dispersion_dtype),
                           ('disp_coating', dispersion_dtype)])

zernike_insn_dtype = np.dtype([('op', np.int32), ('dst', np.int32), ('src', np.int32, (5,)), ('pad', np.int32)])

ZPOLY = 10              # RPX_ZPOLY
ZP_SAG, ZP_R, ZP_RP, ZP_ROR = 0, 1, 2, 3
zernike_coef_dtype = np.dtype([('n', np.int32), ('m', np.int32), ('r_ref', np.int32), ('rp_ref', np.int32),
                               ('ror_ref', np.int32), ('sag_ref', np.int32), ('value', np.float64),
                               ('nval', np.float64), ('deg', np.int32, (4,)), ('poly', np.float64, (4, ZPOLY))])

zernike_dtype = np.dtype([('unit_radius', np.float64), ('n_coefs', np.int32), ('coef_begin', np.int32),
                          ('sag_begin', np.int32), ('sag_len', np.int32), ('grad_begin', np.int32),
                          ('grad_len', np.int32), ('n_slots', np.int32), ('pad', np.int32)])

STRUCT_SIZES = {
    "rpx_transform_t": transform_dtype.itemsize,
    "rpx_face_list_t": face_list_dtype.itemsize,
    "rpx_face_t": face_dtype.itemsize,
    "rpx_shape_op_t": shape_op_dtype.itemsize,
    "rpx_dispersion_t": dispersion_dtype.itemsize,
    "rpx_material_t": material_dtype.itemsize,
    "rpx_zernike_insn_t": zernike_insn_dtype.itemsize,
    "rpx_zernike_coef_t": zernike_coef_dtype.itemsize,
    "rpx_zernike_t": zernike_dtype.itemsize,
}


class SceneDesc(ctypes.Structure):
    """rpx_scene_desc_t"""
    _fields_ = [("face_lists", ctypes.c_void_p), ("n_face_lists", ctypes.c_int32),
                ("faces", ctypes.c_void_p), ("n_faces", ctypes.c_int32),
                ("materials", ctypes.c_void_p), ("n_materials", ctypes.c_int32),
                ("shape_ops", ctypes.c_void_p), ("n_shape_ops", ctypes.c_int32),
                ("vertices", ctypes.c_void_p), ("n_vertices", ctypes.c_int32),
                ("zernikes", ctypes.c_void_p), ("n_zernikes", ctypes.c_int32),
                ("zernike_coefs", ctypes.c_void_p), ("n_zernike_coefs", ctypes.c_int32),
                ("zernike_insns", ctypes.c_void_p), ("n_zernike_insns", ctypes.c_int32)]


class Rays(ctypes.Structure):
    """rpx_rays_t"""
    _fields_ = [("f64", ctypes.c_void_p), ("u32", ctypes.c_void_p), ("para", ctypes.c_void_p),
                ("capacity", ctypes.c_uint64)]


class RaysSplit(ctypes.Structure):
    """rpx_rays_split_t"""
    _fields_ = [("f32", ctypes.c_void_p), ("f64", ctypes.c_void_p), ("u32", ctypes.c_void_p),
                ("capacity", ctypes.c_uint64)]


class JobDesc(ctypes.Structure):
    """rpx_job_desc_t"""
    _fields_ = [("scene", ctypes.c_void_p), ("rays", ctypes.c_void_p), ("n_rays", ctypes.c_uint64),
                ("wavelengths", ctypes.c_void_p), ("n_wavelengths", ctypes.c_uint64),
                ("max_length", ctypes.c_float), ("recursion_limit", ctypes.c_int32),
                ("detector_face", ctypes.c_int32), ("pad", ctypes.c_int32)]


class JobResult(ctypes.Structure):
    """rpx_job_result_t"""
    _fields_ = [("job_id", ctypes.c_uint64), ("status", ctypes.c_int32), ("generations", ctypes.c_int32),
                ("interactions", ctypes.c_uint64), ("children", ctypes.c_uint64),
                ("moments", ctypes.c_double * 6), ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double),
                ("t_finish", ctypes.c_double)]


LENS_MAX_SURFACES = 16
LENS_ZTERMS = 28                 # RPX_LENS_ZTERMS: monomials x^a y^b, a + b <= 6
LENS_MAX_WFE_WAVELENGTHS = 16
LENS_MOMENTS = 2 + 2 * LENS_MAX_WFE_WAVELENGTHS
LENS_SPOT_RMS, LENS_WAVEFRONT_RMS = 0, 1


class Lens(ctypes.Structure):
    """rpx_lens_t"""
    _fields_ = [("n_surfaces", ctypes.c_int32), ("n_media", ctypes.c_int32), ("detector_z", ctypes.c_double),
                ("z", ctypes.c_double * LENS_MAX_SURFACES),
                ("param", (ctypes.c_double * 6) * LENS_MAX_SURFACES),
                ("aperture", ctypes.c_double * LENS_MAX_SURFACES),
                ("medium_before", ctypes.c_int32 * LENS_MAX_SURFACES),
                ("medium_after", ctypes.c_int32 * LENS_MAX_SURFACES),
                ("zernike_radius", ctypes.c_double * LENS_MAX_SURFACES)]


SPLIT_F32 = 19   # RPX_SPLIT_F32_FIELDS (same order as the first 19 float64 fields)
SPLIT_F64 = 2    # RPX_SPLIT_F64_FIELDS: phase, accumulated_path

STRUCT_SIZES["rpx_scene_desc_t"] = ctypes.sizeof(SceneDesc)
STRUCT_SIZES["rpx_rays_t"] = ctypes.sizeof(Rays)
STRUCT_SIZES["rpx_rays_split_t"] = ctypes.sizeof(RaysSplit)
STRUCT_SIZES["rpx_job_desc_t"] = ctypes.sizeof(JobDesc)
STRUCT_SIZES["rpx_job_result_t"] = ctypes.sizeof(JobResult)
STRUCT_SIZES["rpx_lens_t"] = ctypes.sizeof(Lens)


class Generation(ctypes.Structure):
    """rpx_generation_t"""
    _fields_ = [("rays", Rays), ("n_rays", ctypes.c_uint64)]


STRUCT_SIZES["rpx_generation_t"] = ctypes.sizeof(Generation)

LATTICE_RAY_FIELD = 0
LATTICE_POINT_SOURCE = 1
_d3 = ctypes.c_double * 3


class Lattice(ctypes.Structure):
    """rpx_lattice_t"""
    _fields_ = [("kind", ctypes.c_int32), ("nsteps", ctypes.c_int32), ("sin30", ctypes.c_double),
                ("cos30", ctypes.c_double), ("limit", ctypes.c_double), ("spacing", ctypes.c_double),
                ("origin", _d3), ("d1", _d3), ("d2", _d3), ("direction", _d3), ("E_vector", _d3),
                ("weight_num", ctypes.c_double), ("weight_den", ctypes.c_double),
                ("E1", ctypes.c_double * 2), ("E2", ctypes.c_double * 2), ("k_abs", ctypes.c_double),
                ("k_abs_sq", ctypes.c_double), ("power", ctypes.c_double), ("ray_type_id", ctypes.c_uint32),
                ("ray_ident_index", ctypes.c_int32)]


STRUCT_SIZES["rpx_lattice_t"] = ctypes.sizeof(Lattice)
