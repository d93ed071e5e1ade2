/*
 * rpx.h -- C ABI of the MI355X (gfx950) ray-tracing and gausslet field-summation
 * library (librpx.so).
 *
 * Every entry point takes plain pointers and sizes.  Ray state lives in device
 * memory (HBM) as structure-of-arrays; the caller owns every buffer (it may come
 * from any allocator, torch included).  The library owns only scene handles.
 * Functions never throw across the ABI: they return an rpx_status_t and
 * rpx_last_error() gives a message for the calling thread.
 *
 * Reference interfaces replaced (paths relative to the raypier repository):
 *   raypier/core/ctracer.pyx   trace_segment_c / trace_one_face_segment_c /
 *                              trace_gausslet_c / trace_one_face_gausslet_c,
 *                              FaceList.intersect_c, compute_orientation_c,
 *                              RayCollection.copy_as_array / from_array
 *   raypier/core/cfaces.pyx    Face.intersect_c / compute_normal_c (per face type)
 *   raypier/core/cshapes.pyx   Shape.point_inside_c
 *   raypier/core/cdistortions.pyx  ZernikeDistortion.z_offset(_and_gradient)_c
 *   raypier/core/cmaterials.pyx    InterfaceMaterial.eval_child_ray_c /
 *                                  eval_parabasal_ray_c, BaseDispersionCurve
 *   raypier/core/cfields.pyx   sum_gaussian_modes, evaluate_modes
 *   raypier/core/fields.py     evaluate_neighbours_gc
 */
#ifndef RPX_H
#define RPX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RPX_ABI_VERSION 11

/* ------------------------------------------------------------------ status */
typedef int32_t rpx_status_t;
#define RPX_OK              0
#define RPX_E_INVALID       1   /* bad argument / malformed scene table        */
#define RPX_E_UNSUPPORTED   2   /* face / material / shape type not on device  */
#define RPX_E_CAPACITY      3   /* output buffer too small; *n_out = required  */
#define RPX_E_HIP           4   /* HIP runtime error                           */
#define RPX_E_NO_DEVICE     5   /* no usable gfx950 device                     */

/* ------------------------------------------------------------ ray flags   */
/* ctracer.pyx:20-22 */
#define RPX_REFL_RAY   (1u << 0)
#define RPX_GAUSSLET   (1u << 1)
#define RPX_PARABASAL  (1u << 2)
#define RPX_NO_FACE    0xFFFFFFFFu   /* end_face_idx / parent_idx of "-1" */

/* --------------------------------------------------------- ray SoA layout */
/* A generation of rays in HBM.  Field k of ray i lives at f64[k*capacity+i]
 * (resp. u32[k*capacity+i]).  The 21 doubles and 5 uint32 are exactly the
 * fields of ctracer.pxd:40-48 ray_t (188 bytes per ray). */
enum rpx_ray_f64_field {
    RPX_OX = 0, RPX_OY, RPX_OZ,          /* origin            */
    RPX_DX, RPX_DY, RPX_DZ,              /* direction         */
    RPX_NX, RPX_NY, RPX_NZ,              /* normal            */
    RPX_EX, RPX_EY, RPX_EZ,              /* E_vector          */
    RPX_N_RE, RPX_N_IM,                  /* refractive_index  */
    RPX_E1_RE, RPX_E1_IM,                /* E1_amp            */
    RPX_E2_RE, RPX_E2_IM,                /* E2_amp            */
    RPX_LENGTH, RPX_PHASE, RPX_ACCUM,    /* length, phase, accumulated_path */
    RPX_N_F64_FIELDS                     /* = 21 */
};
enum rpx_ray_u32_field {
    RPX_WL_IDX = 0, RPX_PARENT_IDX, RPX_END_FACE_IDX, RPX_RAY_IDENT, RPX_RAY_TYPE_ID,
    RPX_N_U32_FIELDS                     /* = 5 */
};
/* Parabasal rays of a gausslet (ctracer.pxd:54-64): 6 rays x 10 doubles,
 * field (j*10 + f) of gausslet i at para[(j*10+f)*capacity + i],
 * f = ox,oy,oz,dx,dy,dz,nx,ny,nz,length. */
#define RPX_N_PARA        6
#define RPX_PARA_FIELDS   10

typedef struct rpx_rays {
    double   *f64;       /* RPX_N_F64_FIELDS * capacity doubles            */
    uint32_t *u32;       /* RPX_N_U32_FIELDS * capacity uint32             */
    double   *para;      /* gausslets only: 60 * capacity doubles, else 0  */
    uint64_t  capacity;  /* field stride, in elements                      */
} rpx_rays_t;

/* ------------------------------------------------------------ scene table */
/* Transform of a FaceList: p' = M p + t  (ctracer.pxd:67-69 transform_t). */
typedef struct rpx_transform {
    double m[9];
    double t[3];
} rpx_transform_t;

typedef struct rpx_face_list {
    rpx_transform_t trans;       /* FaceList.transform          */
    rpx_transform_t inv_trans;   /* FaceList.inverse_transform  */
    int32_t first_face;          /* index into rpx_scene_desc.faces */
    int32_t n_faces;
    int32_t pad[2];
} rpx_face_list_t;

/* Face types (cfaces.pyx). */
enum rpx_face_type {
    RPX_FACE_CIRCULAR = 1,         /* CircularFace          cfaces.pyx:140  */
    RPX_FACE_SHAPED_PLANAR = 2,    /* ShapedPlanarFace      cfaces.pyx:194  */
    RPX_FACE_RECTANGULAR = 3,      /* RectangularFace       cfaces.pyx:354  */
    RPX_FACE_POLYGON = 4,          /* PolygonFace           cfaces.pyx:1077 */
    RPX_FACE_SPHERICAL = 5,        /* SphericalFace         cfaces.pyx:410  */
    RPX_FACE_SHAPED_SPHERICAL = 6, /* ShapedSphericalFace   cfaces.pyx:501  */
    RPX_FACE_EXTRUDED_PLANAR = 7,  /* ExtrudedPlanarFace    cfaces.pyx:611  */
    RPX_FACE_CONIC = 8,            /* ConicRevolutionFace   cfaces.pyx:1751 */
    RPX_FACE_ASPHERIC = 9,         /* AsphericFace          cfaces.pyx:1885 */
    RPX_FACE_DISTORTION = 10       /* DistortionFace        cfaces.pyx:2323 */
};

#define RPX_FACE_NPARAM 16
/* Parameter slots p[] per face type:
 *   CIRCULAR          diameter, offset, z_plane
 *   SHAPED_PLANAR     z_height
 *   RECTANGULAR       length, width, offset, z_plane
 *   POLYGON           z_plane                    (aux0 = vertex offset, aux1 = count)
 *   SPHERICAL         diameter, curvature, z_height
 *   SHAPED_SPHERICAL  curvature, z_height
 *   EXTRUDED_PLANAR   x1, y1, x2, y2, z1, z2, normal.x, normal.y, normal.z
 *   CONIC             curvature, z_height, conic_const
 *   ASPHERIC          curvature, z_height, conic_const, A4, A6, A8, A10, A12, A14, A16, atol,
 *                     R^3 (R = -curvature), sag-band half-width over the aperture
 *                     (-1 = no band), bound on |dsag/dr| over the aperture
 *   DISTORTION        accuracy                   (aux0 = base face index, aux1 = zernike index)
 * A DISTORTION face's base face is an ordinary record of the face table that
 * belongs to no face list (n_faces of every list excludes it). */
typedef struct rpx_face {
    int32_t type;            /* rpx_face_type                               */
    int32_t global_idx;      /* Face.idx: position in chain(face_lists)     */
    int32_t face_list;       /* owning face list, -1 for a distortion base  */
    int32_t material;        /* index into materials, -1 = no children      */
    int32_t invert_normal;   /* Face.invert_normal (orientation flip)       */
    int32_t invert_normals;  /* CircularFace / ShapedFace.invert_normals     */
    int32_t shape;           /* root op index into shape_ops, -1 = none      */
    int32_t aux0, aux1;
    int32_t pad;
    double  tolerance;       /* Face.tolerance                              */
    double  p[RPX_FACE_NPARAM];
} rpx_face_t;

/* Aperture shapes (cshapes.pyx) as postfix programs.  A face's `shape` field
 * is the index of the program's last (root) op; the root op's `begin` field
 * is the index of its first op.  Ops begin..root run on a boolean stack. */
enum rpx_shape_opcode {
    RPX_SHAPE_TRUE = 1,     /* Shape base class: always inside   ctracer.pyx:1659 */
    RPX_SHAPE_CIRCLE = 2,   /* CircleShape     cshapes.pyx:102 p = cx, cy, radius */
    RPX_SHAPE_RECT = 3,     /* RectangleShape  cshapes.pyx:119 p = cx, cy, width, height */
    RPX_SHAPE_POLYGON = 4,  /* PolygonShape    cshapes.pyx:139 a = vertex offset, b = count */
    RPX_SHAPE_NOT = 5,      /* InvertShape     cshapes.pyx:34  */
    RPX_SHAPE_AND = 6,      /* BooleanAND      cshapes.pyx:55  */
    RPX_SHAPE_OR = 7,       /* BooleanOR       cshapes.pyx:60  */
    RPX_SHAPE_XOR = 8       /* BooleanXOR      cshapes.pyx:65  */
};
typedef struct rpx_shape_op {
    int32_t op;
    int32_t a, b;
    int32_t begin;           /* root op: index of the program's first op */
    double  p[4];
} rpx_shape_op_t;

/* Dispersion curve (cmaterials.pyx:104-239): n(wl) by formula 0/1/2/3/5 over
 * coefs[0..n_coefs), imaginary part 1e-4*absorption/(4 pi)*wl. */
#define RPX_MAX_DISP_COEFS 18
typedef struct rpx_dispersion {
    int32_t formula_id;
    int32_t n_coefs;
    double  absorption;
    double  coefs[RPX_MAX_DISP_COEFS];
} rpx_dispersion_t;

enum rpx_material_type {
    RPX_MAT_NONE = 0,               /* InterfaceMaterial base: no children     */
    RPX_MAT_OPAQUE = 1,             /* OpaqueMaterial            cmaterials.pyx:242  */
    RPX_MAT_TRANSPARENT = 2,        /* TransparentMaterial       cmaterials.pyx:254  */
    RPX_MAT_PEC = 3,                /* PECMaterial               cmaterials.pyx:281  */
    RPX_MAT_PARTIAL = 4,            /* PartiallyReflectiveMaterial cmaterials.pyx:322 */
    RPX_MAT_DIELECTRIC = 5,         /* DielectricMaterial        cmaterials.pyx:554  */
    RPX_MAT_FULL_DIELECTRIC = 6,    /* FullDielectricMaterial    cmaterials.pyx:727  */
    RPX_MAT_FULL_DISPERSIVE = 7,    /* FullDielectricDispersiveMaterial cmaterials.pyx:875 */
    RPX_MAT_SINGLE_LAYER = 8,       /* SingleLayerCoatedMaterial cmaterials.pyx:1017 */
    RPX_MAT_COATED_DISPERSIVE = 9,  /* CoatedDispersiveMaterial  cmaterials.pyx:1187 */
    RPX_MAT_GRATING = 10            /* DiffractionGratingMaterial cmaterials.pyx:1437 */
};

typedef struct rpx_material {
    int32_t type;                   /* rpx_material_type                         */
    int32_t order;                  /* grating order (the first, for several)    */
    int32_t n_orders;               /* grating: 0 or 1 = the one `order`
                                       (DiffractionGratingMaterial); 2 or 3 =
                                       `order`, then the signed bytes of
                                       more_orders, one child per propagating
                                       order in that order (this build's
                                       multi-order grating)                       */
    int32_t more_orders;            /* second order in bits 0-7, third in 8-15   */
    double  reflection_threshold;
    double  transmission_threshold;
    double  n_inside[2];            /* complex (re, im), non-dispersive types    */
    double  n_outside[2];
    double  n_coating[2];
    double  thickness;              /* coating thickness in microns              */
    double  reflectivity;           /* PartiallyReflectiveMaterial               */
    double  lines_per_mm;           /* grating                                   */
    double  efficiency;             /* grating                                   */
    double  origin[3];              /* grating origin                            */
    double  pad2;
    rpx_dispersion_t disp_inside;   /* dispersive types                          */
    rpx_dispersion_t disp_outside;
    rpx_dispersion_t disp_coating;
} rpx_material_t;

/* Zernike distortion (cdistortions.pyx:323-514) compiled to straight-line form.
 * The reference evaluates radial polynomials by a memoised recursion over a
 * 3 x k_max workspace.  At scene upload the recursion is executed symbolically,
 * preserving the workspace slot each value is cached in (including slots shared
 * by different (n,m) pairs), giving a list of instructions
 *     slot[dst] = combine(op, value(a), value(b), value(c), value(d), value(e))
 * where value(x) = slot[x] for x >= 0, 0.0 for x == -1, 1.0 for x == -2, and
 *   op 0 (R):        r*(a + b) - c
 *   op 1 (R'):       ((a + b) + r*(c + d)) - e
 *   op 2 (R/r):      (a + b) - c
 * Then, per coefficient i (sorted by ANSI j), Z terms use value(r_ref[i]),
 * value(rp_ref[i]), value(ror_ref[i]).  Two programs exist per distortion:
 * `sag` (z_offset_c: radial values only) and `grad` (z_offset_and_gradient_c).
 * Every op is an integer-coefficient polynomial operation in r, so each value a
 * coefficient uses is a polynomial in r with integer coefficients; the
 * builder expands the programs exactly and stores those polynomials in the
 * coefficient records, which is what the kernels evaluate (Horner in r).
 * The programs are validated and kept as the provenance of the polynomials. */
typedef struct rpx_zernike_insn {
    int32_t op;
    int32_t dst;
    int32_t src[5];
    int32_t pad;
} rpx_zernike_insn_t;

#define RPX_ZPOLY 10        /* polynomial coefficients per value (n <= 9) */
enum { RPX_ZP_SAG = 0, RPX_ZP_R = 1, RPX_ZP_RP = 2, RPX_ZP_ROR = 3 };

typedef struct rpx_zernike_coef {
    int32_t n, m;             /* m signed (ANSI)                          */
    int32_t r_ref, rp_ref, ror_ref;   /* value refs in the grad program   */
    int32_t sag_ref;                  /* value ref in the sag program     */
    double  value;
    double  nval;             /* sqrt(n+1) (m == 0) or sqrt(2(n+1)), times value:
                                 the normalised coefficient N of
                                 cdistortions.pyx:434-436, computed once */
    int32_t deg[4];           /* degree of each polynomial below (0..9)   */
    double  poly[4][RPX_ZPOLY];   /* value(sag_ref), value(r_ref), value(rp_ref),
                                     value(ror_ref) as polynomials in r,
                                     ascending powers                      */
} rpx_zernike_coef_t;

typedef struct rpx_zernike {
    double  unit_radius;
    int32_t n_coefs;
    int32_t coef_begin;       /* index into zernike_coefs                  */
    int32_t sag_begin, sag_len;     /* range in zernike_insns              */
    int32_t grad_begin, grad_len;
    int32_t n_slots;          /* workspace slots the programs use (<= 3 * k_max) */
    int32_t pad;
} rpx_zernike_t;

#define RPX_MAX_ZERNIKE_SLOTS 96

typedef struct rpx_scene_desc {
    const rpx_face_list_t   *face_lists;   int32_t n_face_lists;
    const rpx_face_t        *faces;        int32_t n_faces;
    const rpx_material_t    *materials;    int32_t n_materials;
    const rpx_shape_op_t    *shape_ops;    int32_t n_shape_ops;
    const double            *vertices;     int32_t n_vertices;   /* (x, y) pairs */
    const rpx_zernike_t     *zernikes;     int32_t n_zernikes;
    const rpx_zernike_coef_t *zernike_coefs; int32_t n_zernike_coefs;
    const rpx_zernike_insn_t *zernike_insns; int32_t n_zernike_insns;
} rpx_scene_desc_t;

typedef struct rpx_scene rpx_scene_t;   /* opaque, device-resident copy */

/* ----------------------------------------------------------------- library */
int32_t     rpx_abi_version(void);
const char *rpx_last_error(void);
/* Number of bytes of each host-side struct, for binding checks. */
int32_t     rpx_struct_size(const char *name);

/* Upload a scene (face lists, faces, materials, shapes, distortions) to the
 * current device.  Replaces the per-generation walk over Python FaceList /
 * Face / InterfaceMaterial objects in ctracer.pyx:2062-2118 and the
 * FaceList.sync_transforms calls of ctracer.pyx:2173-2211. */
rpx_status_t rpx_scene_create(const rpx_scene_desc_t *desc, rpx_scene_t **out);
/* No work that reads the scene may still be in flight: its device tables are
 * reused by the next scene created on that device, with no device-wide
 * synchronisation. */
rpx_status_t rpx_scene_destroy(rpx_scene_t *scene);
/* Largest number of children one ray can produce in this scene. */
int32_t      rpx_scene_max_children(const rpx_scene_t *scene);

/* Workspace bytes needed to trace n_in rays (counts, block offsets). */
uint64_t rpx_trace_workspace_bytes(uint64_t n_in);

/* One non-sequential generation: ctracer.pyx:2062-2118 trace_segment_c.
 *   in        parent generation; length and end_face_idx are updated in place
 *   wavelengths  device array of vacuum wavelengths (microns) indexed by
 *             wavelength_idx (the materials' `wavelengths`)
 *   out       children in parent-major order, reflected before transmitted
 *   face_hits device uint32[n_global_faces], incremented per hit (Face.count)
 *   n_out     host pointer; receives the number of children.  If it exceeds
 *             out->capacity, RPX_E_CAPACITY is returned and nothing is written
 *             to `out` (the in-place parent update still happened).
 * The call is stream-ordered and returns after the stream has completed. */
rpx_status_t rpx_trace_segment(const rpx_scene_t *scene,
                               const rpx_rays_t *in, uint64_t n_in,
                               float max_length,
                               const double *wavelengths, uint64_t n_wavelengths,
                               rpx_rays_t *out, uint64_t *n_out,
                               uint32_t *face_hits,
                               void *workspace, uint64_t workspace_bytes,
                               void *stream);

/* The generation loop of core/tracer.py:9-47 trace_rays for a ray
 * collection (non-sequential, ctracer.pyx:2062-2118 per generation) with no
 * host round trip between generations.  Generation 0 is `in` (n_in rays,
 * lengths and end faces updated in place, as by rpx_trace_segment); the
 * children of generation g (g >= 0) are written to `pool` (caller-owned
 * SoA, field stride pool->capacity) at element offsets[g+1], generation
 * after generation, and traced in turn, until a generation is empty or
 * max_generations generations are traced.  The children of the last of
 * max_generations traced generations are counted, not stored (trace_rays
 * discards them).
 *   counts, offsets  host arrays of max_generations + 1: counts[g] rays in
 *                    generation g (counts[0] = n_in), offsets[g] its first
 *                    ray's element in the pool (g >= 1)
 *   n_traced         generations traced: 0..n_traced-1 are trace_rays' list
 *   face_hits        HOST uint32[n_global_faces], incremented by the hits of
 *                    every traced generation (Face.count)
 *   workspace        rpx_chain_workspace_bytes(max(n_in, pool->capacity),
 *                    max_generations) bytes
 * RPX_E_CAPACITY when the pool could not be sure to hold the next
 * generation's children: generations 0..n_traced-1 are traced and generation
 * n_traced (counts[n_traced] rays at offsets[n_traced]) is stored but not
 * traced -- continue from it with a fresh pool.  The call returns after the
 * stream has completed. */
uint64_t rpx_chain_workspace_bytes(uint64_t max_rays_per_generation, int32_t max_generations);
rpx_status_t rpx_trace_rays(const rpx_scene_t *scene,
                            const rpx_rays_t *in, uint64_t n_in,
                            float max_length,
                            const double *wavelengths, uint64_t n_wavelengths,
                            int32_t max_generations, rpx_rays_t *pool,
                            uint64_t *counts, uint64_t *offsets, int32_t *n_traced,
                            uint32_t *face_hits,
                            void *workspace, uint64_t workspace_bytes,
                            void *stream);

/* The same loop for a gausslet collection (core/tracer.py:9-47 with
 * ctracer.pyx:2214-2282 per generation: trace_gausslet_c and the parabasal
 * rays), as rpx_trace_rays with gausslet buffers (`para` set on `in` and the
 * pool) and max_length in double precision (as rpx_trace_gausslet).  For scenes without a decomposition (ResampleGaussletMaterial)
 * face, whose captures need the host between generations. */
rpx_status_t rpx_trace_gausslets(const rpx_scene_t *scene,
                            const rpx_rays_t *in, uint64_t n_in,
                            double max_length,
                            const double *wavelengths, uint64_t n_wavelengths,
                            int32_t max_generations, rpx_rays_t *pool,
                            uint64_t *counts, uint64_t *offsets, int32_t *n_traced,
                            uint32_t *face_hits,
                            void *workspace, uint64_t workspace_bytes,
                            void *stream);

/* Sequential variant, one prescribed face: ctracer.pyx:2121-2170
 * trace_one_face_segment_c.  face_list / face_idx select FaceList and the
 * index of the face inside it. */
rpx_status_t rpx_trace_one_face_segment(const rpx_scene_t *scene,
                               int32_t face_list, int32_t face_idx,
                               const rpx_rays_t *in, uint64_t n_in,
                               float max_length,
                               const double *wavelengths, uint64_t n_wavelengths,
                               rpx_rays_t *out, uint64_t *n_out,
                               uint32_t *face_hits,
                               void *workspace, uint64_t workspace_bytes,
                               void *stream);

/* Gausslet generation: ctracer.pyx:2214-2281 trace_gausslet_c with
 * trace_parabasal_rays (ctracer.pyx:2350-2385).  `in`/`out` carry para[].
 * face_list < 0 selects the non-sequential search; otherwise the single face
 * (face_list, face_idx) as in trace_one_face_gausslet_c (ctracer.pyx:2284). */
rpx_status_t rpx_trace_gausslet(const rpx_scene_t *scene,
                               int32_t face_list, int32_t face_idx,
                               const rpx_rays_t *in, uint64_t n_in,
                               double max_length,
                               const double *wavelengths, uint64_t n_wavelengths,
                               rpx_rays_t *out, uint64_t *n_out,
                               uint32_t *face_hits,
                               void *workspace, uint64_t workspace_bytes,
                               void *stream);

/* Capture-face selection: select_ray_intersections (ctracer.pyx:1981-2021)
 * and select_gausslet_intersections (ctracer.pyx:2024-2058) for ONE
 * collection of ray_col_list.  Each ray (base ray of each gausslet) is tested
 * along origin -> origin + direction * length against face list `face_list`
 * as FaceList.intersect_c does (ctracer.pyx:1882-1904).  A ray whose nearest
 * accepted face has a global index >= 0 is copied to out[out_offset + k], in
 * input order, with length and end_face_idx updated and wavelength_idx
 * replaced by wl_map[wavelength_idx]; `in` is not modified.  The caller's
 * np.unique re-indexing of the wavelength union (ctracer.pyx:2015-2021) is
 * the wl_map of each collection.  *n_out = rays selected; RPX_E_CAPACITY if
 * out_offset + *n_out exceeds out->capacity, RPX_E_INVALID if a selected
 * ray's wavelength_idx is not below n_wl_map.  Workspace:
 * rpx_trace_workspace_bytes(n_in). */
rpx_status_t rpx_select_intersections(const rpx_scene_t *scene, int32_t face_list,
                               const rpx_rays_t *in, uint64_t n_in, int32_t gausslets,
                               const uint32_t *wl_map, uint64_t n_wl_map,
                               rpx_rays_t *out, uint64_t out_offset, uint64_t *n_out,
                               void *workspace, uint64_t workspace_bytes,
                               void *stream);

/* Gausslet capture of ResampleGaussletMaterial (cmaterials.pyx:1766-1831):
 * after a gausslet generation (`in`, updated in place), copies every gausslet
 * whose end_face_idx is one of face_ids (<= 16) to out[out_offset + k] in
 * input order -- the base ray as the generation left it, the parabasal rays
 * with their lengths from para_length (device, row j at
 * para_length[j * para_length_stride + i], taken before the generation, as
 * the reference captures before trace_parabasal_rays moves them; NULL: as in
 * `in`).  *n_out = gausslets captured.  Workspace:
 * rpx_trace_workspace_bytes(n_in). */
rpx_status_t rpx_capture_gausslets(const rpx_rays_t *in, uint64_t n_in, const int32_t *face_ids,
                                   int32_t n_face_ids, const double *para_length,
                                   uint64_t para_length_stride, rpx_rays_t *out, uint64_t out_offset,
                                   uint64_t *n_out, void *workspace, uint64_t workspace_bytes,
                                   void *stream);

/* ------------------------------------------------ generation consumers
 * Walks over the generations of one trace (rpx_trace_segment outputs, in
 * trace order), resident in HBM.  A generation is its SoA buffer and count. */
typedef struct rpx_generation {
    rpx_rays_t rays;
    uint64_t   n_rays;
} rpx_generation_t;

/* Workspace of rpx_path_lengths / rpx_face_histogram (n_last: rays of the
 * last generation; 0 for the histogram). */
uint64_t rpx_lineage_workspace_bytes(int32_t n_generations, uint64_t n_last);

/* MeanOpticalPathLength._calc_result (results.py:112-122) without the final
 * mean: for every ray of the last generation with end_face_idx ==
 * target_face, in ray order, totals[k] = sum over the ray and its ancestors
 * (following parent_idx back to generation 0) of length * Re(n), summed from
 * the last generation backwards, each product rounded before its addition.
 * *n_selected = number of such rays; RPX_E_CAPACITY if above capacity;
 * RPX_E_INVALID if a parent_idx on a walked chain is outside the previous
 * generation.  The reference's result is the mean of totals. */
rpx_status_t rpx_path_lengths(const rpx_generation_t *generations, int32_t n_generations,
                              uint32_t target_face, double *totals, uint64_t capacity,
                              uint64_t *n_selected, void *workspace, uint64_t workspace_bytes,
                              void *stream);

/* End-face histogram of every generation, the vote of set_face_sequence
 * (sources.py:138-155: the most common end_face_idx of each generation).
 * counts (device, n_generations * (n_bins + 2) uint64):
 * counts[g*(n_bins+2) + b] = rays of generation g with end_face_idx == b for
 * b < n_bins; b = n_bins counts end_face_idx == 0xFFFFFFFF (no face);
 * b = n_bins + 1 counts every other value.  n_bins <= 4094. */
rpx_status_t rpx_face_histogram(const rpx_generation_t *generations, int32_t n_generations,
                                uint32_t n_bins, uint64_t *counts, void *workspace,
                                uint64_t workspace_bytes, void *stream);

/* RayCollection neighbours of a child generation from those of its parent
 * generation (ctracer.pyx:1084-1106 _eval_neighbours): with rmap[p][t] the
 * last child (highest index) of parent p whose ray_type_id & REFL_RAY is t,
 * neighbours[i*K + j] = rmap[c][t_i] for c = parent_neighbours[p_i*K + j] >= 0,
 * else -1 (K = n_neighbours; row-major int32 arrays on the device).
 * RPX_E_INVALID if a parent_idx or neighbour index is outside the parent
 * generation.  Workspace: rpx_neighbours_workspace_bytes(n_parents). */
uint64_t rpx_neighbours_workspace_bytes(uint64_t n_parents);
rpx_status_t rpx_propagate_neighbours(const rpx_rays_t *children, uint64_t n_children,
                                      const int32_t *parent_neighbours, uint64_t n_parents,
                                      int32_t n_neighbours, int32_t *neighbours,
                                      void *workspace, uint64_t workspace_bytes, void *stream);

/* --------------------------------------------------------------- sources
 * Hexagonal lattices of source rays generated in HBM.  Lattice indices vi, vj
 * run over [-nsteps-1, nsteps+1] in numpy.meshgrid order (label = row * side
 * + column with side = 2*nsteps + 3, the row giving vj); selected points
 * keep label order.  The host computes the scalars and vectors exactly as the
 * reference does (numpy) and the kernels evaluate the per-point expressions
 * in the reference's order.
 *   RPX_LATTICE_RAY_FIELD (HexagonalRayFieldSource sources.py:1035-1107,
 *   CollimatedGaussletSource gausslet_sources.py:386-457):
 *     r = (vi + sin30*vj)^2 + cos30*vj^2, selected if r < limit;
 *     origin = (vi*d1 + vj*d2) * spacing + origin; direction, E_vector given.
 *   RPX_LATTICE_POINT_SOURCE (GaussianPointSource gausslet_sources.py:516-566,
 *   make_hexagonal_grid core/gausslets.py:55-80):
 *     kx = (vi + sin30*vj)*spacing, ky = (cos30*vj)*spacing,
 *     r = kx^2 + ky^2, selected if r < limit; kz = sqrt(k_abs_sq - r);
 *     direction = normalised ((kx*d1 + ky*d2) + kz*direction) / k_abs;
 *     E_vector = direction x d2; origin = origin.
 *   Both: weight g = exp(-((r * weight_num) / weight_den)); E1/E2 = the
 *   complex amplitude times g; normal (0,1,0); n = 1; other fields 0;
 *   ray_ident = output index if ray_ident_index, else 0. */
#define RPX_LATTICE_RAY_FIELD    0
#define RPX_LATTICE_POINT_SOURCE 1

typedef struct rpx_lattice {
    int32_t  kind;
    int32_t  nsteps;
    double   sin30, cos30;
    double   limit;
    double   spacing;
    double   origin[3];
    double   d1[3], d2[3];
    double   direction[3];
    double   E_vector[3];
    double   weight_num, weight_den;
    double   E1[2], E2[2];           /* (re, im) */
    double   k_abs, k_abs_sq;
    double   power;                  /* P of the weight power sum, 0 = skip */
    uint32_t ray_type_id;
    int32_t  ray_ident_index;
} rpx_lattice_t;

uint64_t rpx_lattice_workspace_bytes(int32_t nsteps);
/* Numbers the selected points in the workspace; *n_points = their count;
 * *weight_power (optional) = sum over them of power * g^2 (the beam power
 * normalisation of CollimatedGaussletSource, gausslet_sources.py:430-432,
 * summed in a fixed order). */
rpx_status_t rpx_lattice_select(const rpx_lattice_t *lattice, uint64_t *n_points, double *weight_power,
                                void *workspace, uint64_t workspace_bytes, void *stream);
/* Writes the selected points (after rpx_lattice_select on the same
 * workspace and lattice) into `out`; with out->para the six parabasal rays
 * copy the base ray (GaussletCollection.from_rays, ctracer.pyx:1321-1345);
 * with `neighbours` (device int32, n_points x 6) the lattice neighbours
 * (sources.py:1063-1085), -1 where unselected or off the lattice. */
rpx_status_t rpx_lattice_emit(const rpx_lattice_t *lattice, rpx_rays_t *out, int32_t *neighbours,
                              void *workspace, uint64_t workspace_bytes, void *stream);

/* GaussletCollection.config_parabasal_rays (ctracer.pyx:1430-1481) on
 * gausslets in HBM.  wavelengths: device table indexed by wavelength_idx;
 * angle_cos_sin (host): cos and sin of 0, 2pi/3 and 4pi/3 as the reference's
 * libm computes them, interleaved (c0, s0, c1, s1, c2, s2).  Workspace: 8 B.
 * RPX_E_INVALID if a wavelength_idx is outside the table. */
rpx_status_t rpx_config_parabasal_rays(rpx_rays_t *gausslets, uint64_t n, const double *wavelengths,
                                       uint64_t n_wavelengths, double radius, double working_dist,
                                       const double *angle_cos_sin, void *workspace,
                                       uint64_t workspace_bytes, void *stream);

/* GaussletCollection.project_to_plane (ctracer.pyx:1391-1420): every base
 * ray (and parabasal ray, when rays->para) moves along its direction to the
 * plane through `origin` with normal `direction` (host 3-vectors); the base
 * ray's accumulated_path grows by Re(n) times the distance moved. */
rpx_status_t rpx_project_to_plane(rpx_rays_t *rays, uint64_t n, const double *origin,
                                  const double *direction, void *stream);

/* GaussletCollection.lagrange_invariant (ctracer.pyx:1347-1389): per
 * gausslet, 1000 sqrt(v/6) / wavelength, where v sums the squared
 * differences h_i.u_j - h_j.u_i over the parabasal pairs (0,5) (1,2) (3,4)
 * (1,4) (0,3) (2,5), h being a parabasal ray's offset from the base origin
 * and u its direction, both in the frame (E/|E|, E/|E| x direction).
 * out: device float64[n].  Workspace: 8 B.  RPX_E_INVALID if a
 * wavelength_idx is outside the table. */
rpx_status_t rpx_lagrange_invariant(const rpx_rays_t *gausslets, uint64_t n, const double *wavelengths,
                                    uint64_t n_wavelengths, double *out, void *workspace,
                                    uint64_t workspace_bytes, void *stream);

/* ---------------------------------------------------- FP32-split form
 * A reduced-precision ray generation with no counterpart in the reference
 * (which traces in float64 only).  Ray state is stored in float32 -- vectors,
 * refractive index, amplitudes, length -- with phase and accumulated optical
 * path in float64:
 *   f32[k*capacity + i], k in rpx_split_f32_field  (19 planes)
 *   f64[k*capacity + i], k in rpx_split_f64_field  (2 planes)
 *   u32[k*capacity + i], k in rpx_ray_u32_field    (5 planes)
 * 112 bytes per ray against 188.  Arithmetic: intersection root solves,
 * surface normals and child directions (reflection, refraction, grating) in
 * float64 from the float32 inputs; polarisation basis, Fresnel / thin-film
 * amplitudes and power thresholds in float32; optical path and phase in
 * float64.  Child order, parent links and the in-place parent update follow
 * rpx_trace_segment; near thresholds (aperture edges, power thresholds,
 * total internal reflection) the topology can differ from float64. */
typedef enum {
    RPX_S_OX = 0, RPX_S_OY, RPX_S_OZ,
    RPX_S_DX, RPX_S_DY, RPX_S_DZ,
    RPX_S_NX, RPX_S_NY, RPX_S_NZ,
    RPX_S_EX, RPX_S_EY, RPX_S_EZ,
    RPX_S_N_RE, RPX_S_N_IM,
    RPX_S_E1_RE, RPX_S_E1_IM,
    RPX_S_E2_RE, RPX_S_E2_IM,
    RPX_S_LENGTH,
    RPX_SPLIT_F32_FIELDS = 19
} rpx_split_f32_field;

typedef enum {
    RPX_S_PHASE = 0,
    RPX_S_ACCUM = 1,
    RPX_SPLIT_F64_FIELDS = 2
} rpx_split_f64_field;

typedef struct {
    float *f32;
    double *f64;
    uint32_t *u32;
    uint64_t capacity;
} rpx_rays_split_t;

/* float64 SoA <-> FP32-split SoA (device, stream-ordered). */
rpx_status_t rpx_rays_to_split(const rpx_rays_t *in, uint64_t n, rpx_rays_split_t *out, void *stream);
rpx_status_t rpx_rays_from_split(const rpx_rays_split_t *in, uint64_t n, rpx_rays_t *out, void *stream);

/* One FP32-split generation; arguments as rpx_trace_segment (workspace size
 * from rpx_trace_workspace_bytes).  face_list < 0 searches every face list;
 * otherwise only face (face_list, face_idx), as rpx_trace_one_face_segment. */
rpx_status_t rpx_trace_segment_split(const rpx_scene_t *scene,
                                     int32_t face_list, int32_t face_idx,
                                     const rpx_rays_split_t *in, uint64_t n_in,
                                     float max_length,
                                     const double *wavelengths, uint64_t n_wavelengths,
                                     rpx_rays_split_t *out, uint64_t *n_out,
                                     uint32_t *face_hits,
                                     void *workspace, uint64_t workspace_bytes,
                                     void *stream);

/* The generation loop of trace_rays in the FP32-split form, device-resident
 * exactly as rpx_trace_rays (same counts / offsets / n_traced / face_hits /
 * RPX_E_CAPACITY contract; the pool is an FP32-split SoA buffer and the
 * workspace rpx_chain_workspace_bytes(max(n_in, pool->capacity),
 * max_generations) bytes).  No counterpart in the reference. */
rpx_status_t rpx_trace_rays_split(const rpx_scene_t *scene,
                                  const rpx_rays_split_t *in, uint64_t n_in,
                                  float max_length,
                                  const double *wavelengths, uint64_t n_wavelengths,
                                  int32_t max_generations, rpx_rays_split_t *pool,
                                  uint64_t *counts, uint64_t *offsets, int32_t *n_traced,
                                  uint32_t *face_hits,
                                  void *workspace, uint64_t workspace_bytes,
                                  void *stream);

/* ------------------------------------------------------------- serving
 * A continuous-batching trace server (no counterpart in the reference, whose
 * trace loop runs one collection at a time).  Jobs -- a scene, device rays,
 * a wavelength table -- are submitted as they arrive; rpx_server_step admits
 * waiting jobs first-come-first-served while their ray buffers fit the HBM
 * budget, then traces one generation of every active job in one persistent
 * launch (replayed from a captured HIP graph), and returns the jobs that
 * finished.  A job's generations follow rpx_trace_segment (float64) exactly;
 * what it returns is a summary: interactions, generations and the moments of
 * its detector hits.  Buffers are the server's own (stream-ordered
 * allocations from a pool of its own); the caller's rays are copied at
 * submission and may be freed after rpx_server_submit returns; the wavelength
 * table must stay valid until the job finishes.  The budget covers every
 * buffer the server holds: a waiting job's input copy from submission (a
 * submit that would exceed it returns RPX_E_CAPACITY), an active job's input
 * and children buffers.  A job whose buffer cannot be allocated, or whose
 * launch reports an error, finishes with that status; the server stays
 * usable. */
typedef struct rpx_server rpx_server_t;

typedef struct {
    const rpx_scene_t *scene;
    const rpx_rays_t *rays;        /* device SoA, float64 */
    uint64_t n_rays;
    const double *wavelengths;     /* device */
    uint64_t n_wavelengths;
    float max_length;
    int32_t recursion_limit;       /* as trace_rays: generations traced */
    int32_t detector_face;         /* global face index whose hits are summed, -1 = none */
    int32_t pad;
} rpx_job_desc_t;

typedef struct {
    uint64_t job_id;
    int32_t status;                /* RPX_OK, or the error that ended the job */
    int32_t generations;           /* generations traced */
    uint64_t interactions;         /* rays of all generations whose search hit a face */
    uint64_t children;             /* rays created over all generations */
    double moments[6];             /* detector hits: count, sum x, sum y, sum z,
                                      sum x^2 + y^2 + z^2 (global frame) */
    double t_submit, t_admit, t_finish;   /* seconds on the server's monotonic clock */
} rpx_job_result_t;

rpx_status_t rpx_server_create(uint64_t hbm_budget_bytes, uint32_t max_jobs, uint64_t max_rays_in_flight,
                               void *stream, rpx_server_t **out);
rpx_status_t rpx_server_destroy(rpx_server_t *server);
/* Queue a job; its id is returned in *job_id. */
rpx_status_t rpx_server_submit(rpx_server_t *server, const rpx_job_desc_t *job, uint64_t *job_id);
/* Admit, run one batched generation, and copy up to max_finished results of
 * jobs that completed into `finished` (*n_finished of them); completed jobs
 * beyond max_finished are returned by later calls. */
rpx_status_t rpx_server_step(rpx_server_t *server, rpx_job_result_t *finished, uint32_t max_finished,
                             uint32_t *n_finished);
/* Jobs waiting for admission and jobs on the device. */
rpx_status_t rpx_server_counts(const rpx_server_t *server, uint32_t *waiting, uint32_t *active);
/* Admission statistics since the server was created: jobs that were held
 * back at least once because their buffers would have exceeded the HBM
 * budget, and the most bytes the budget has counted at once. */
rpx_status_t rpx_server_stats(const rpx_server_t *server, uint64_t *jobs_held_for_budget,
                              uint64_t *peak_used_bytes);
/* Seconds on the server's clock (the clock of t_submit / t_admit / t_finish). */
double rpx_server_clock(void);
/* Generations a step may trace per job: 1 to 4 (the default is 2; no
 * counterpart in the reference).  After the first, each generation of a step
 * is traced without a host round trip: the device lays out its tiles from
 * the previous generation's child counts, and a job traces it when it has a
 * generation left before its recursion limit and its children fit, at
 * max_children each, in the buffer the previous generation came from, which
 * it writes into (a job's two buffers alternate).  A job's results are the
 * same for every setting. */
rpx_status_t rpx_server_set_generations_per_step(rpx_server_t *server, uint32_t generations);
/* How a step's work is enqueued: on != 0 (the default) replays a HIP graph
 * built once per kernel variant and generation count, 0 launches the same
 * copies and kernels one by one on the server's stream.  Results are the
 * same either way (no counterpart in the reference). */
rpx_status_t rpx_server_set_graph_launch(rpx_server_t *server, int32_t on);
/* Test hook (no counterpart in the reference): after n_allocations further
 * job-buffer allocations succeed, the next one fails once, as if the device
 * were out of memory; n_allocations < 0 cancels a pending failure. */
rpx_status_t rpx_server_fail_allocations_after(rpx_server_t *server, int64_t n_allocations);

/* ------------------------------------------------- sequential adjoint
 * Spot RMS of a coaxial sequential lens and its gradient by a hand-written
 * adjoint (no counterpart in the reference).  Surface s is
 *   z = z[s] + c r^2 / (1 + sqrt(1 - (1 + k) c^2 r^2)) + A4 r^4 + A6 r^6
 *       + A8 r^8 + A10 r^10,     r^2 = x^2 + y^2,
 * with param[s] = {c, k, A4, A6, A8, A10} (c = 1/R in mm^-1; AsphericFace
 * with curvature -R and inverted normals traces the same surface).  Each ray
 * is traced through surfaces 0..n-1 in order -- intersection by Newton's
 * method from the vertex plane, refraction by the vector form of Snell's law
 * with index ratio n_table[wl][medium_before] / n_table[wl][medium_after] --
 * to the plane z = detector_z.  A ray that leaves a surface's clear aperture,
 * misses it or is totally reflected is dropped.  The spot RMS is
 * sqrt(mean |p - mean p|^2) over the (x, y) of the kept rays' detector
 * points.
 *
 * A surface with zernike_radius[s] = Ru > 0 carries a Zernike distortion
 * (ZernikeDistortion, cdistortions.pyx:323-514, as DistortionFace adds it to
 * its base face's sag, cfaces.pyx:2323-2438): its sag gains
 *   D(x, y) = sum_k C[s][k] (x / Ru)^a_k (y / Ru)^b_k
 * over the RPX_LENS_ZTERMS monomials a + b <= 6 in the order (0,0), (1,0),
 * (0,1), (2,0), (1,1), (0,2), (3,0), ... (degree by degree, a descending).
 * C (device, n_surfaces x RPX_LENS_ZTERMS) is the Zernike sum written in
 * those monomials (the host expands sum_j c_j N_j R_n^m(r) cos/sin(m theta));
 * the gradient with respect to C maps to the coefficients c_j by the same
 * expansion.
 *
 * Objectives: RPX_LENS_SPOT_RMS as above; RPX_LENS_WAVEFRONT_RMS, the RMS
 * over the kept rays of the wavefront error W - mean_lambda(W), with the
 * piston removed per wavelength (at most RPX_LENS_MAX_WFE_WAVELENGTHS),
 * where W is a ray's optical path from its origin (plus its
 * accumulated_path) to the foot of the perpendicular from the on-axis
 * point P = (0, 0, detector_z) on its last segment:
 *   W = accumulated_path + sum_s n_s t_s + n_last (P - q_last) . d_last,
 * the wavefront error against the sphere centred on P, to first order. */
#define RPX_LENS_MAX_SURFACES 16
#define RPX_LENS_ZTERMS 28
#define RPX_LENS_MAX_WFE_WAVELENGTHS 16
#define RPX_LENS_MOMENTS (2 + 2 * RPX_LENS_MAX_WFE_WAVELENGTHS)

enum rpx_lens_objective {
    RPX_LENS_SPOT_RMS = 0,
    RPX_LENS_WAVEFRONT_RMS = 1
};

typedef struct {
    int32_t n_surfaces;
    int32_t n_media;                            /* columns of the index table */
    double detector_z;
    double z[RPX_LENS_MAX_SURFACES];
    double param[RPX_LENS_MAX_SURFACES][6];     /* c, k, A4, A6, A8, A10 */
    double aperture[RPX_LENS_MAX_SURFACES];     /* clear radius */
    int32_t medium_before[RPX_LENS_MAX_SURFACES];
    int32_t medium_after[RPX_LENS_MAX_SURFACES];
    double zernike_radius[RPX_LENS_MAX_SURFACES];   /* Ru; 0 = no distortion */
} rpx_lens_t;

uint64_t rpx_lens_workspace_bytes(uint64_t n_rays, int32_t n_surfaces);

/* Objective moments (device, RPX_LENS_MOMENTS doubles):
 *   spot RMS:      kept rays, sum x, sum y, sum x^2 + y^2;
 *   wavefront RMS: kept rays, sum (W - mean_lambda W)^2, then per wavelength
 *                  l: moments[2 + 2l] = kept rays, moments[3 + 2l] = sum W.
 * zcart may be null when no surface has a distortion. */
rpx_status_t rpx_lens_eval(const rpx_lens_t *lens, const double *zcart, const rpx_rays_t *rays, uint64_t n_rays,
                           const double *n_table, uint64_t n_wavelengths, int32_t objective, double *moments,
                           void *workspace, uint64_t workspace_bytes, void *stream);

/* grad (device, n_surfaces * (6 + RPX_LENS_ZTERMS) doubles):
 * grad[s*6 + q] = d(objective)/d(param[s][q]);
 * grad[6*n_surfaces + s*RPX_LENS_ZTERMS + k] = d(objective)/d(C[s][k]) (0
 * where the surface has no distortion), given the moments of rpx_lens_eval
 * for the same inputs.  Intersections are recomputed, not stored. */
rpx_status_t rpx_lens_grad(const rpx_lens_t *lens, const double *zcart, const rpx_rays_t *rays, uint64_t n_rays,
                           const double *n_table, uint64_t n_wavelengths, int32_t objective, const double *moments,
                           double *grad, void *workspace, uint64_t workspace_bytes, void *stream);

/* Spot RMS without distortions (rpx_lens_eval / rpx_lens_grad with
 * RPX_LENS_SPOT_RMS and no zcart; zernike_radius is ignored).
 * Forward: moments[0..3] (device) = kept rays, sum x, sum y, sum x^2 + y^2. */
rpx_status_t rpx_lens_spot(const rpx_lens_t *lens, const rpx_rays_t *rays, uint64_t n_rays,
                           const double *n_table, uint64_t n_wavelengths,
                           double *moments, void *workspace, uint64_t workspace_bytes, void *stream);

/* Adjoint: grad[s*6 + q] (device) = d(spot RMS)/d(param[s][q]), given the
 * moments of rpx_lens_spot for the same inputs.  Intersections are
 * recomputed, not stored. */
rpx_status_t rpx_lens_spot_grad(const rpx_lens_t *lens, const rpx_rays_t *rays, uint64_t n_rays,
                                const double *n_table, uint64_t n_wavelengths, const double *moments,
                                double *grad, void *workspace, uint64_t workspace_bytes, void *stream);

/* Kernel timing with HIP events recorded on the stream each generation runs
 * on (no counterpart in the reference: measurement support).  While enabled,
 * every trace call adds the duration of its kernels to running totals, one
 * slot per kernel (below).  rpx_timing_read copies RPX_TIMING_SLOTS entries
 * of: ms = total duration, launches = number of launches, bytes = bytes the
 * kernel's algorithm reads + writes (per the ray record sizes); then resets
 * the totals.  Ray generations run k_clear + k_generation (FP32-split:
 * k_clear + k_generation_split); gausslet generations run k_clear +
 * k_gausslet_generation (slots 0-2 belonged to an earlier three-launch
 * gausslet generation and stay empty).  The two field-sum slots time the pair kernel of
 * rpx_sum_gaussian_modes (k_field_sum, k_field_sum_split) and count
 * gausslet x point pairs in `bytes`; while timing is on, that call returns
 * after its pair kernel has completed. */
enum {
    RPX_TIMING_INTERSECT = 0,
    RPX_TIMING_SCAN = 1,
    RPX_TIMING_EMIT = 2,
    RPX_TIMING_GENERATION = 3,
    RPX_TIMING_CLEAR = 4,
    RPX_TIMING_SPLIT = 5,
    RPX_TIMING_FIELD_SUM = 6,
    RPX_TIMING_FIELD_SUM_SPLIT = 7,
    RPX_TIMING_GAUSSLET = 8,
    RPX_TIMING_SLOTS = 9
};
rpx_status_t rpx_timing_enable(int32_t on);
rpx_status_t rpx_timing_read(double *ms, uint64_t *launches, double *bytes);

/* Tile order of the device-resident generation chains (rpx_trace_rays,
 * rpx_trace_rays_split; no counterpart in the reference).  A chain first
 * runs with each workgroup's tile given by its launch order, every look-back
 * wait bounded in time; if a wait runs out (the hardware started a later
 * workgroup before an earlier one could run), the chain is traced again from
 * its input with tiles handed out by an atomic counter, which needs no
 * dispatch order.  Results are the same either way.
 * rpx_tile_order_fallbacks: chains traced again so far in this process.
 * rpx_debug_lookback_ticks: test hook -- the wait bound in ticks of the
 * device's real-time counter (0 restores the default of 50 ms; 1 makes the
 * first wait of every launch-order tile run out, forcing the fallback). */
uint64_t rpx_tile_order_fallbacks(void);
rpx_status_t rpx_debug_lookback_ticks(uint64_t ticks);

/* Per-object queries, batched over n device-resident inputs.
 *   rpx_query_face: what = 0 -> out[i] = Face.intersect_c(p1[i], p2[i],
 *   is_base_ray).dist (cfaces.pyx, -1 for no intersection); what = 1 ->
 *   out[3i..3i+2] = Face.compute_normal_c(p1[i]) (local frame).
 *   face_rec indexes the scene's face table. */
rpx_status_t rpx_query_face(const rpx_scene_t *scene, int32_t face_rec, int32_t what,
                            const double *p1, const double *p2, uint64_t n,
                            int32_t is_base_ray, double *out, void *stream);
/* Shape.point_inside_c (cshapes.pyx) of the program rooted at `root`;
 * xy = n (x, y) pairs, out = n int32 0/1. */
rpx_status_t rpx_query_shape(const rpx_scene_t *scene, int32_t root, const double *xy,
                             uint64_t n, int32_t *out, void *stream);
/* ZernikeDistortion.z_offset_c -> sag[i] and z_offset_and_gradient_c ->
 * grad[3i..3i+2] = (dz/dx, dz/dy, z) (cdistortions.pyx:416-514).
 * Either output may be null. */
rpx_status_t rpx_query_zernike(const rpx_scene_t *scene, int32_t zernike, const double *xy,
                               uint64_t n, double *sag, double *grad, void *stream);
/* InterfaceMaterial.eval_child_ray_c for n independent (ray, point,
 * orientation) inputs (ctracer.pyx:1620-1632).  Ray i's children go to
 * out[3i + k], k < counts[i]; idx_base + i is written as their parent_idx. */
rpx_status_t rpx_eval_child_rays(const rpx_scene_t *scene, int32_t material,
                                 const rpx_rays_t *in, uint64_t n, uint32_t idx_base,
                                 const double *points, const double *normals,
                                 const double *tangents,
                                 const double *wavelengths, uint64_t n_wavelengths,
                                 rpx_rays_t *out, uint32_t *counts, void *stream);

/* AoS <-> SoA: RayCollection.copy_as_array / from_array (ctracer.pyx:1048,
 * 1142) for the packed 188-byte ray_t and 668-byte gausslet_t layouts.  The
 * AoS buffer is read or written by a kernel: device memory, or pinned host
 * memory mapped for the device. */
rpx_status_t rpx_pack_rays(const rpx_rays_t *soa, uint64_t n, void *aos,
                           int32_t gausslet, void *stream);
rpx_status_t rpx_unpack_rays(const void *aos, uint64_t n, rpx_rays_t *soa,
                             int32_t gausslet, void *stream);

/* evaluate_neighbours_gc (core/fields.py:114-136) fused with evaluate_modes
 * (cfields.pyx:156-228): per gausslet, the 3 complex mode coefficients.
 *   modes  device complex128[n][3] (interleaved re, im) */
rpx_status_t rpx_gausslet_modes(const rpx_rays_t *gausslets, uint64_t n,
                                double blending, double *modes, void *stream);

/* evaluate_modes (cfields.pyx:218-228) on explicit x, y, dx, dy arrays of
 * shape [n][row] (device, float64). */
rpx_status_t rpx_evaluate_modes(const double *x, const double *y,
                                const double *dx, const double *dy,
                                uint64_t n, int32_t row, double blending,
                                double *modes, void *stream);

/* ------------------------------------------------ beam decomposition
 * calc_mode_curvature (cfields.pyx:231-293): for each of n points, the mode
 * basis x', y', z' (n x 3 each, row-major) from the wavefront slopes dx, dy
 * and the curvature coefficients A, B, C from the second derivatives. */
rpx_status_t rpx_calc_mode_curvature(const double *dx, const double *dy, const double *dx2,
                                     const double *dy2, const double *dxdy, uint64_t n,
                                     double *A, double *B, double *C,
                                     double *x, double *y, double *z, void *stream);

/* apply_mode_curvature (cfields.pyx:397-424): adds the curvature
 * coefficients A, B, C (device, one per gausslet) to the parabasal
 * directions of n gausslets in HBM. */
rpx_status_t rpx_apply_mode_curvature(rpx_rays_t *gausslets, uint64_t n, const double *A,
                                      const double *B, const double *C, void *stream);

/* build_interaction_matrix (cfields.pyx:295-395): the field of mode i (at
 * (rx, ry) with curvature A, B, C, basis x, y, z and width spacing/blending)
 * at test point j (tx, ty), for every pair within max_spacing, in (i, j)
 * order: data[2k], data[2k+1] = E_ij (complex), ij[k] = i, ij[capacity+k] = j.
 * *n_out = number of pairs; RPX_E_CAPACITY if above capacity.  (The
 * reference's test_spacing only sizes its output buffer.)  Workspace:
 * rpx_trace_workspace_bytes(n * l). */
rpx_status_t rpx_build_interaction_matrix(const double *rx, const double *ry, uint64_t n,
                                          const double *tx, const double *ty, uint64_t l,
                                          const double *A, const double *B, const double *C,
                                          const double *x, const double *y, const double *z,
                                          double wavelength, double spacing, double max_spacing,
                                          double blending, double *data, uint64_t *ij,
                                          uint64_t capacity, uint64_t *n_out, void *workspace,
                                          uint64_t workspace_bytes, void *stream);

/* project_to_sphere (core/fields.py:51-80): each ray whose line meets the
 * sphere (centre, radius) moves along its direction to the intersection
 * with the smaller (most negative) path parameter, its accumulated_path
 * growing by that parameter times Re(n); rays whose line misses are dropped.
 * The kept rays go to out[0 .. *n_out) in input order.  Workspace:
 * rpx_trace_workspace_bytes(n). */
rpx_status_t rpx_project_to_sphere(const rpx_rays_t *in, uint64_t n, const double *centre, double radius,
                                   rpx_rays_t *out, uint64_t *n_out, void *workspace,
                                   uint64_t workspace_bytes, void *stream);

/* evaluate_neighbours (core/fields.py:83-111) and evaluate_modes
 * (cfields.pyx:156-228) for a ray field: every ray whose six neighbours
 * (device int32, n x 6, indices into `rays`) are all >= 0 projects them onto
 * the plane through its origin normal to its direction, takes their offsets
 * and slopes in its (E, E x d) frame and fits its Gaussian mode.  Those rays
 * go to out[0 .. *n_out) in input order, their modes (complex A, B, C as 6
 * doubles) to modes[6k ..] and, when neighbour_xy is given, the neighbour
 * coordinates x[6], y[6], dx[6], dy[6] to neighbour_xy[24k ..].
 * RPX_E_INVALID if a neighbour index is >= n.
 * Workspace: rpx_trace_workspace_bytes(n). */
rpx_status_t rpx_ray_field_modes(const rpx_rays_t *rays, uint64_t n, const int32_t *neighbours,
                                 double blending, rpx_rays_t *out, double *modes, double *neighbour_xy,
                                 uint64_t *n_out, void *workspace, uint64_t workspace_bytes, void *stream);

/* sum_gaussian_modes (cfields.pyx:52-115).
 *   rays      base rays (SoA, only the ray part is read)
 *   modes     device complex128[n_rays][3]
 *   wavelengths device float64 table indexed by wavelength_idx
 *   points    device float64[n_pts][3]
 *   out       device complex128[n_pts][3], overwritten
 *   precision 0 = FP64 (the reference arithmetic),
 *             1 = FP32 amplitudes with double-single phase */
rpx_status_t rpx_sum_gaussian_modes(const rpx_rays_t *rays, uint64_t n_rays,
                                    const double *modes,
                                    const double *wavelengths, uint64_t n_wavelengths,
                                    const double *points, uint64_t n_pts,
                                    double time_ps, double *out,
                                    int32_t precision, void *stream);

/* rpx_sum_gaussian_modes without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).  The
 * per-gausslet invariants table and the chunk partial sums live in a
 * file-scope device array of RPX_FIELD_SCRATCH_BYTES per GPU, held until the
 * process ends; nothing is allocated or freed and no memset is issued.
 * Arguments and results are those of rpx_sum_gaussian_modes (the same sums,
 * in the same order, whenever its partial sums fit beside the table; else
 * fewer gausslet chunks).  n_rays above rpx_field_noalloc_max_rays() returns
 * RPX_E_CAPACITY.  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (two streams, two threads) race and must be ordered by
 * the caller. */
#define RPX_FIELD_SCRATCH_BYTES (128ull << 20)
uint64_t     rpx_field_noalloc_max_rays(void);
rpx_status_t rpx_sum_gaussian_modes_noalloc(const rpx_rays_t *rays, uint64_t n_rays,
                                            const double *modes,
                                            const double *wavelengths, uint64_t n_wavelengths,
                                            const double *points, uint64_t n_pts,
                                            double time_ps, double *out,
                                            int32_t precision, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* RPX_H */
