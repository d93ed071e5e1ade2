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

// 1/d from the hardware reciprocal estimate and two Newton-Raphson steps
// (e = 1 - d x, x += x e), within an ulp of the rounded quotient for the
// normal, finite d of the field sum.
RPX_DEV double rcp_nr(double d) {
    double x = __builtin_amdgcn_rcp(d);
    double e = fma(-d, x, 1.0);
    x = fma(x, e, x);
    e = fma(-d, x, 1.0);
    return fma(x, e, x);
}

// 1/sqrt(a) from the hardware estimate and two Newton-Raphson steps
// (y += y (1 - a y^2) / 2), for the normal, finite, positive a of the field sum.
RPX_DEV double rsq_nr(double a) {
    double y = __builtin_amdgcn_rsq(a);
    double e = fma(-(a * y), y, 1.0);
    y = fma(0.5 * y, e, y);
    e = fma(-(a * y), y, 1.0);
    return fma(0.5 * y, e, y);
}

// sin(x) and cos(x): Cody-Waite reduction r = x - n pi/2 with pi/2 split as
// C1 (33 significant bits, so n C1 is exact for |n| < 2^20, and x - n C1 is
// exact) + C2, then the Taylor series of sin to r^15 and of cos to r^16 on
// |r| <= pi/4 (truncation below 5e-17) and the quadrant of n.  Arguments
// beyond 2^20 pi/2 go to the library routine.
RPX_DEV void sincos_cw(double x, double *s, double *c) {
    const double TWO_OVER_PI = 0.6366197723675814;
    const double C1 = 1.5707963267341256;       // 0x1.921fb544p+0
    const double C2 = 6.077100506506192e-11;    // pi/2 - C1, rounded
    const double nd = rint(x * TWO_OVER_PI);
    if (!(fabs(nd) < 1048576.0)) {
        sincos(x, s, c);
        return;
    }
    const double r = fma(-nd, C2, fma(-nd, C1, x));
    const double u = r * r;
    double ps = -7.647163731819816e-13;          // -1/15!
    ps = fma(ps, u, 1.6059043836821613e-10);     // 1/13!
    ps = fma(ps, u, -2.505210838544172e-08);     // -1/11!
    ps = fma(ps, u, 2.7557319223985893e-06);     // 1/9!
    ps = fma(ps, u, -1.984126984126984e-04);     // -1/7!
    ps = fma(ps, u, 8.333333333333333e-03);      // 1/5!
    ps = fma(ps, u, -1.6666666666666666e-01);    // -1/3!
    const double sr = fma(r * u, ps, r);
    double pc = 4.779477332387385e-14;           // 1/16!
    pc = fma(pc, u, -1.1470745597729725e-11);    // -1/14!
    pc = fma(pc, u, 2.08767569878681e-09);       // 1/12!
    pc = fma(pc, u, -2.755731922398589e-07);     // -1/10!
    pc = fma(pc, u, 2.48015873015873e-05);       // 1/8!
    pc = fma(pc, u, -1.388888888888889e-03);     // -1/6!
    pc = fma(pc, u, 4.1666666666666664e-02);     // 1/4!
    pc = fma(pc, u, -0.5);                       // -1/2!
    const double cr = fma(u, pc, 1.0);
    const int q = (int)nd & 3;
    const double sv = (q & 1) ? cr : sr, cv = (q & 1) ? sr : cr;
    *s = (q & 2) ? -sv : sv;
    *c = ((q + 1) & 2) ? -cv : cv;
}

// Float32 sin(x) and cos(x) for |x| < 2^13: r = x - n pi/2 by two fused
// steps (pi/2 = P1 + P2 in float32; fmaf rounds each step once), then the
// Taylor series of sin to r^11 and of cos to r^12 on |r| <= pi/4
// (truncation below 2e-9) and the quadrant of n.  Larger arguments go to the
// library routine.
RPX_DEV void sincos_cwf(float x, float *s, float *c) {
    const float nf = rintf(x * 0.6366197466850281f);
    if (!(fabsf(nf) < 8192.0f)) {
        sincosf(x, s, c);
        return;
    }
    const float r = fmaf(-nf, -4.371138828673793e-08f, fmaf(-nf, 1.5707963705062866f, x));
    const float u = r * r;
    float ps = -2.5052107943679403e-08f;         // -1/11!
    ps = fmaf(ps, u, 2.7557318844628753e-06f);   // 1/9!
    ps = fmaf(ps, u, -1.9841270113829523e-04f);  // -1/7!
    ps = fmaf(ps, u, 8.333333767950535e-03f);    // 1/5!
    ps = fmaf(ps, u, -1.666666716337204e-01f);   // -1/3!
    const float sr = fmaf(r * u, ps, r);
    float pc = 2.0876755879584152e-09f;          // 1/12!
    pc = fmaf(pc, u, -2.755731998149713e-07f);   // -1/10!
    pc = fmaf(pc, u, 2.4801587642286904e-05f);   // 1/8!
    pc = fmaf(pc, u, -1.3888889225199819e-03f);  // -1/6!
    pc = fmaf(pc, u, 4.16666679084301e-02f);     // 1/4!
    pc = fmaf(pc, u, -0.5f);                     // -1/2!
    const float cr = fmaf(u, pc, 1.0f);
    const int q = (int)nf & 3;
    const float sv = (q & 1) ? cr : sr, cv = (q & 1) ? sr : cr;
    *s = (q & 2) ? -sv : sv;
    *c = ((q + 1) & 2) ? -cv : cv;
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
    V_SPHERES = F_PLANE | F_SPHERE | M_SIMPLE | M_FRESNEL | M_COATED,
    // flat optics (cubes, plane and distorted mirrors, gratings); the
    // FP32-split generation kernel only, other kernels run V_ALL for it
    V_FLAT = F_PLANE | F_DISTORT | M_SIMPLE | M_FRESNEL | M_GRATING
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
