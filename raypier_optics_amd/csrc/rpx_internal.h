// Device-side types and arithmetic helpers for the gfx950 tracer.
//
// All FP64 routines keep the operation order of the reference Cython code
// (raypier/core/ctracer.pyx:83-262 vector maths, cmaterials.pyx complex maths)
// and the library is compiled with -ffp-contract=off, so that add/mul/div/sqrt
// chains round exactly like the reference's x86-64 build.  Only libm-level
// functions (pow, sin, cos, atan2, exp, hypot) may differ by an ulp.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/rpx.h"

#define RPX_DEV __device__ __forceinline__
#define RPX_INF (__builtin_inf())

namespace rpx {

// ------------------------------------------------------------------ vectors
struct V3 {
    double x, y, z;
};

RPX_DEV V3 v3(double x, double y, double z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }

// transform_c / rotate_c (ctracer.pyx:83-95)
RPX_DEV V3 transform(const rpx_transform_t &t, V3 p) {
    V3 o;
    o.x = p.x * t.m[0] + p.y * t.m[1] + p.z * t.m[2] + t.t[0];
    o.y = p.x * t.m[3] + p.y * t.m[4] + p.z * t.m[5] + t.t[1];
    o.z = p.x * t.m[6] + p.y * t.m[7] + p.z * t.m[8] + t.t[2];
    return o;
}
RPX_DEV V3 rotate(const rpx_transform_t &t, V3 p) {
    V3 o;
    o.x = p.x * t.m[0] + p.y * t.m[1] + p.z * t.m[2];
    o.y = p.x * t.m[3] + p.y * t.m[4] + p.z * t.m[5];
    o.z = p.x * t.m[6] + p.y * t.m[7] + p.z * t.m[8];
    return o;
}
RPX_DEV V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
RPX_DEV V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
RPX_DEV V3 mul(V3 a, double s) { return v3(a.x * s, a.y * s, a.z * s); }
RPX_DEV double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RPX_DEV double mag_sq(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
RPX_DEV double mag(V3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
RPX_DEV V3 cross(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// norm_ divides each component by the magnitude (ctracer.pyx:251-256)
RPX_DEV V3 norm(V3 a) {
    double m = sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    return v3(a.x / m, a.y / m, a.z / m);
}
RPX_DEV V3 neg(V3 a) { return v3(-a.x, -a.y, -a.z); }
// sep_ (ctracer.pyx:109-114)
RPX_DEV double sep(V3 p1, V3 p2) {
    double a = p2.x - p1.x, b = p2.y - p1.y, c = p2.z - p1.z;
    return sqrt((a * a) + (b * b) + (c * c));
}

// ------------------------------------------------------------ complex FP64
// Mixed real/complex expressions of the reference are evaluated with the real
// operand promoted to (x, 0), as the Cython-generated C does.
struct Cx {
    double re, im;
};
RPX_DEV Cx cx(double r, double i) { Cx c; c.re = r; c.im = i; return c; }
RPX_DEV Cx cadd(Cx a, Cx b) { return cx(a.re + b.re, a.im + b.im); }
RPX_DEV Cx csub(Cx a, Cx b) { return cx(a.re - b.re, a.im - b.im); }
RPX_DEV Cx cneg(Cx a) { return cx(-a.re, -a.im); }
RPX_DEV Cx cmul(Cx a, Cx b) {
    return cx(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
// Smith's method (R. L. Smith, CACM 5(8) 1962, Algorithm 116), the scheme C
// runtimes use for double-complex division in the normal range.
RPX_DEV Cx cdiv(Cx n, Cx d) {
    double a = n.re, b = n.im, c = d.re, e = d.im;
    double ratio, denom, x, y;
    if (fabs(c) < fabs(e)) {
        ratio = c / e;
        denom = (c * ratio) + e;
        x = ((a * ratio) + b) / denom;
        y = ((b * ratio) - a) / denom;
    } else {
        ratio = e / c;
        denom = (e * ratio) + c;
        x = ((b * ratio) + a) / denom;
        y = (b - (a * ratio)) / denom;
    }
    return cx(x, y);
}
RPX_DEV double cabs_(Cx a) { return hypot(a.re, a.im); }
// Principal square root (Kahan's formulation).
RPX_DEV Cx csqrt_(Cx z) {
    double x = z.re, y = z.im;
    if (x == 0.0 && y == 0.0) return cx(0.0, y);
    if (__builtin_isinf(y)) return cx(RPX_INF, y);
    double t = sqrt((fabs(x) + hypot(x, y)) * 0.5);
    if (x >= 0.0) return cx(t, y / (2.0 * t));
    return cx(fabs(y) / (2.0 * t), copysign(t, y));
}
RPX_DEV Cx cexp_(Cx z) {
    double e = exp(z.re);
    double s, c;
    sincos(z.im, &s, &c);
    return cx(e * c, e * s);
}

// ---------------------------------------------------------- feature masks
// Face and material families; generation kernels are instantiated for a few
// masks and a scene runs the smallest one that covers what it uses.
enum : uint32_t {
    F_PLANE = 1u << 0,      // circular, shaped planar, rectangular, polygon, extruded
    F_SPHERE = 1u << 1,     // spherical, shaped spherical
    F_CONIC = 1u << 2,
    F_ASPHERE = 1u << 3,
    F_DISTORT = 1u << 4,
    M_SIMPLE = 1u << 5,     // none, opaque, transparent, PEC, partial, dielectric
    M_FRESNEL = 1u << 6,    // full dielectric, full dispersive
    M_COATED = 1u << 7,     // single layer, coated dispersive
    M_GRATING = 1u << 8,
    V_ALL = 0x1ffu,
    V_LENS = F_PLANE | F_SPHERE | F_CONIC | F_ASPHERE | M_SIMPLE | M_FRESNEL | M_COATED,
    V_SPHERES = F_PLANE | F_SPHERE | M_SIMPLE | M_FRESNEL | M_COATED
};

// ------------------------------------------------------------------- scene
// Device view of an uploaded scene (pointers into one device allocation).
struct SceneView {
    const rpx_face_list_t *face_lists;
    const rpx_face_t *faces;
    const rpx_material_t *materials;
    const rpx_shape_op_t *shape_ops;
    const double *vertices;
    const rpx_zernike_t *zernikes;
    const rpx_zernike_coef_t *zcoefs;
    const rpx_zernike_insn_t *zinsns;
    int32_t n_face_lists, n_faces, n_materials, n_shape_ops, n_vertices;
    int32_t n_zernikes, n_zcoefs, n_zinsns;
    int32_t n_list_faces;   // faces that belong to face lists (chain order)
    int32_t max_children;
};

// Intersection result (ctracer.pxd:77-81, piece/uv unused by these faces)
struct Hit {
    double dist;
    int32_t face;     // face record index, -1 = none
};

}  // namespace rpx
