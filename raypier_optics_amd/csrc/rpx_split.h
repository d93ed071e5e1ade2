// FP32-split form of the interface materials.
//
// Ray state is float32 except phase and optical path (float64).  Every
// quantity that moves a ray -- hit point, surface normal, reflected,
// refracted and diffracted directions, and the refractive-index ratio that
// bends them -- is computed in float64 from the float32 inputs, so rounding
// enters the geometry only once per generation, at the store, and is not
// biased by float32 scene parameters.  The polarisation basis, Fresnel and
// thin-film amplitudes and the power thresholds run in float32.  The physics
// and child order follow the float64 materials (rpx_materials.h), which
// follow cmaterials.pyx.
#pragma once
#include "rpx_materials.h"

namespace rpx {

// ---------------------------------------------------------------- float32
struct F3 {
    float x, y, z;
};
RPX_DEV F3 f3(float x, float y, float z) { F3 r; r.x = x; r.y = y; r.z = z; return r; }
RPX_DEV F3 f3_of(V3 a) { return f3((float)a.x, (float)a.y, (float)a.z); }
RPX_DEV V3 v3_of(F3 a) { return v3(a.x, a.y, a.z); }
RPX_DEV F3 fadd(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
RPX_DEV F3 fsub(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
RPX_DEV F3 fmulv(F3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
RPX_DEV float fdot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RPX_DEV F3 fcross(F3 a, F3 b) {
    return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
RPX_DEV F3 fnorm(F3 a) {
    float inv = 1.0f / sqrtf(a.x * a.x + a.y * a.y + a.z * a.z);
    return f3(a.x * inv, a.y * inv, a.z * inv);
}

struct Cf {
    float re, im;
};
RPX_DEV Cf cf(float r, float i) { Cf c; c.re = r; c.im = i; return c; }
RPX_DEV Cf cf_of(Cx a) { return cf((float)a.re, (float)a.im); }
RPX_DEV Cf cfadd(Cf a, Cf b) { return cf(a.re + b.re, a.im + b.im); }
RPX_DEV Cf cfsub(Cf a, Cf b) { return cf(a.re - b.re, a.im - b.im); }
RPX_DEV Cf cfneg(Cf a) { return cf(-a.re, -a.im); }
RPX_DEV Cf cfscale(Cf a, float s) { return cf(a.re * s, a.im * s); }
RPX_DEV Cf cfmul(Cf a, Cf b) { return cf(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
// Smith's division, as in the float64 path.
RPX_DEV Cf cfdiv(Cf n, Cf d) {
    float a = n.re, b = n.im, c = d.re, e = d.im;
    if (fabsf(c) < fabsf(e)) {
        float ratio = c / e;
        float inv = 1.0f / (c * ratio + e);
        return cf((a * ratio + b) * inv, (b * ratio - a) * inv);
    }
    float ratio = e / c;
    float inv = 1.0f / (e * ratio + c);
    return cf((b * ratio + a) * inv, (b - a * ratio) * inv);
}
RPX_DEV float cfabs2(Cf a) { return a.re * a.re + a.im * a.im; }
// Principal square root (Kahan's formulation).
RPX_DEV Cf cfsqrt(Cf z) {
    float x = z.re, y = z.im;
    if (x == 0.0f && y == 0.0f) return cf(0.0f, y);
    float t = sqrtf((fabsf(x) + hypotf(x, y)) * 0.5f);
    if (x >= 0.0f) return cf(t, y / (2.0f * t));
    return cf(fabsf(y) / (2.0f * t), copysignf(t, y));
}
RPX_DEV Cf cfexp(Cf z) {
    float e = expf(z.re), s, c;
    sincosf(z.im, &s, &c);
    return cf(e * c, e * s);
}

// ------------------------------------------------------------- ray record
struct RayS {
    F3 origin, direction, normal, E_vector;
    Cf n, E1, E2;
    float length;
    double phase, accum;
    uint32_t wl_idx, parent_idx, end_face_idx, ray_ident, ray_type_id;
};

RPX_DEV RayS load_ray_s(const rpx_rays_split_t &r, uint64_t i) {
    const float *f = r.f32;
    const uint64_t c = r.capacity;
    RayS o;
    o.origin = f3(f[RPX_S_OX * c + i], f[RPX_S_OY * c + i], f[RPX_S_OZ * c + i]);
    o.direction = f3(f[RPX_S_DX * c + i], f[RPX_S_DY * c + i], f[RPX_S_DZ * c + i]);
    o.normal = f3(f[RPX_S_NX * c + i], f[RPX_S_NY * c + i], f[RPX_S_NZ * c + i]);
    o.E_vector = f3(f[RPX_S_EX * c + i], f[RPX_S_EY * c + i], f[RPX_S_EZ * c + i]);
    o.n = cf(f[RPX_S_N_RE * c + i], f[RPX_S_N_IM * c + i]);
    o.E1 = cf(f[RPX_S_E1_RE * c + i], f[RPX_S_E1_IM * c + i]);
    o.E2 = cf(f[RPX_S_E2_RE * c + i], f[RPX_S_E2_IM * c + i]);
    o.length = f[RPX_S_LENGTH * c + i];
    o.phase = r.f64[RPX_S_PHASE * c + i];
    o.accum = r.f64[RPX_S_ACCUM * c + i];
    const uint32_t *u = r.u32;
    o.wl_idx = u[RPX_WL_IDX * c + i];
    o.parent_idx = u[RPX_PARENT_IDX * c + i];
    o.end_face_idx = u[RPX_END_FACE_IDX * c + i];
    o.ray_ident = u[RPX_RAY_IDENT * c + i];
    o.ray_type_id = u[RPX_RAY_TYPE_ID * c + i];
    return o;
}

RPX_DEV void store_ray_s(const rpx_rays_split_t &r, uint64_t i, const RayS &o) {
    float *f = r.f32;
    const uint64_t c = r.capacity;
    f[RPX_S_OX * c + i] = o.origin.x;  f[RPX_S_OY * c + i] = o.origin.y;  f[RPX_S_OZ * c + i] = o.origin.z;
    f[RPX_S_DX * c + i] = o.direction.x; f[RPX_S_DY * c + i] = o.direction.y; f[RPX_S_DZ * c + i] = o.direction.z;
    f[RPX_S_NX * c + i] = o.normal.x;  f[RPX_S_NY * c + i] = o.normal.y;  f[RPX_S_NZ * c + i] = o.normal.z;
    f[RPX_S_EX * c + i] = o.E_vector.x; f[RPX_S_EY * c + i] = o.E_vector.y; f[RPX_S_EZ * c + i] = o.E_vector.z;
    f[RPX_S_N_RE * c + i] = o.n.re;    f[RPX_S_N_IM * c + i] = o.n.im;
    f[RPX_S_E1_RE * c + i] = o.E1.re;  f[RPX_S_E1_IM * c + i] = o.E1.im;
    f[RPX_S_E2_RE * c + i] = o.E2.re;  f[RPX_S_E2_IM * c + i] = o.E2.im;
    f[RPX_S_LENGTH * c + i] = o.length;
    r.f64[RPX_S_PHASE * c + i] = o.phase;
    r.f64[RPX_S_ACCUM * c + i] = o.accum;
    uint32_t *u = r.u32;
    u[RPX_WL_IDX * c + i] = o.wl_idx;
    u[RPX_PARENT_IDX * c + i] = o.parent_idx;
    u[RPX_END_FACE_IDX * c + i] = o.end_face_idx;
    u[RPX_RAY_IDENT * c + i] = o.ray_ident;
    u[RPX_RAY_TYPE_ID * c + i] = o.ray_type_id;
}

// convert_to_sp (cmaterials.pyx:49-91) in float32.  At exactly normal
// incidence the basis is left unchanged, as in float64; elsewhere the S/P
// basis is the same physical basis, so the field E1*E + E2*(d x E) agrees
// with the float64 path to float32 rounding.
RPX_DEV RayS convert_to_sp_s(RayS ray, F3 normal) {
    F3 E2v = fnorm(fcross(ray.direction, ray.E_vector));
    F3 E1v = fnorm(fcross(E2v, ray.direction));
    F3 S = fcross(ray.direction, normal);
    if (fabsf(S.x) < 1.0e-10f && fabsf(S.y) < 1.0e-10f && fabsf(S.z) < 1.0e-10f) return ray;
    S = fnorm(S);
    F3 P = fnorm(fcross(ray.direction, S));
    float A = fdot(E1v, S), B = fdot(E2v, S);
    Cf Samp = cf(ray.E1.re * A + ray.E2.re * B, ray.E1.im * A + ray.E2.im * B);
    B = fdot(E1v, P);
    A = fdot(E2v, P);
    Cf Pamp = cf(ray.E1.re * B + ray.E2.re * A, ray.E1.im * B + ray.E2.im * A);
    ray.E_vector = S;
    ray.E1 = Samp;
    ray.E2 = Pamp;
    return ray;
}

RPX_DEV float power_frac_s(float nre, Cf a, Cf b, float P_in) {
    return nre * (cfabs2(a) + cfabs2(b)) / P_in;
}

// Float64 refraction of in_dir (unit) at unit normal with index ratio n1/n2
// (cmaterials.pyx:650-660 geometry).
RPX_DEV V3 refract_dir(V3 in_dir, V3 normal, double cosTheta, double ratio, int flip) {
    V3 ctn = mul(normal, cosTheta);
    V3 tg2 = mul(sub(in_dir, ctn), ratio);
    double c2 = sqrt(1 - mag_sq(tg2));
    return sub(tg2, mul(normal, c2 * flip));
}

RPX_DEV V3 reflect_dir(V3 in_dir, V3 normal, double cosTheta) {
    return sub(in_dir, mul(mul(normal, cosTheta), 2));
}

// Common start of every material: S/P basis, optical path of the segment
// just traced (float64: length of this generation times the ray's index),
// child origin and normal.
RPX_DEV RayS begin_child(const RayS &in, double length, F3 normal_f, V3 point, uint32_t idx) {
    RayS sp = convert_to_sp_s(in, normal_f);
    sp.accum += length * (double)sp.n.re;
    sp.origin = f3_of(point);
    sp.normal = normal_f;
    sp.length = __builtin_inff();
    sp.parent_idx = idx;
    return sp;
}

// The grating's children, FP32-split (the float64 grating_children's
// emitter protocol: one child per propagating order, those past kCap
// counted, not built).
template <class Emit>
RPX_DEV void grating_children_split(const rpx_material_t &m, const RayS &in, double length, uint32_t idx, V3 point,
                                    const Orientation &orient, const WlTable &wavelengths, Emit &emit) {
    const V3 normal = norm(orient.normal);
    const F3 normal_f = f3_of(normal);
    const V3 tangent = norm(orient.tangent);
    const V3 tangent2 = cross(normal, tangent);
    const double wavelen = wavelengths[in.wl_idx];
    const double line_spacing = 1000.0 / m.lines_per_mm;
    const V3 dir = norm(v3_of(in.direction));
    const double k_z0 = dot(normal, dir);
    const double k_y = dot(tangent2, dir);
    const double k_x0 = dot(tangent, dir);
    const int sign = (k_z0 < 0.0) ? 1 : -1;
    const int n_orders = grating_orders(m);
    int built = 0;
    for (int o = 0; o < n_orders; ++o) {
        const int order = grating_order(m, o);
        const double k_x = k_x0 - order * wavelen / (line_spacing * (double)in.n.re);
        double k_z = 1 - (k_x * k_x) - (k_y * k_y);
        if (k_z < 0) continue;
        if (built >= Emit::kCap) {
            emit.skip();
            continue;
        }
        ++built;
        k_z = sign * sqrt(k_z);
        const V3 reflected = add(add(mul(tangent, k_x), mul(tangent2, k_y)), mul(normal, k_z));
        RayS sp = begin_child(in, length, normal_f, point, idx);
        sp.length = (float)length;   // the grating child keeps the hit length (cmaterials.pyx:1472-1542)
        const float eff = (float)m.efficiency;
        sp.direction = f3_of(reflected);
        sp.E1 = cfscale(sp.E1, -eff);
        sp.E2 = cfscale(sp.E2, eff);
        sp.ray_type_id |= RPX_REFL_RAY;
        const V3 og = v3(m.origin[0], m.origin[1], m.origin[2]);
        sp.phase += 1000.0 * dot(sub(og, point), tangent) * order * 2 * M_PI / line_spacing;
        emit(sp);
    }
}

// InterfaceMaterial.eval_child_ray_c of every material, FP32-split.
template <uint32_t F, class Emit>
RPX_DEV void eval_children_split(const rpx_material_t &m, const RayS &in, double length, uint32_t idx,
                                 V3 point, const Orientation &orient, const WlTable &wavelengths, Emit &emit) {
    const V3 normal = norm(orient.normal);
    const F3 normal_f = f3_of(normal);
    switch (m.type) {
    case RPX_MAT_TRANSPARENT: {
        if constexpr ((F & M_SIMPLE) == 0) return;
        RayS sp = begin_child(in, length, normal_f, point, idx);
        sp.ray_type_id &= ~RPX_REFL_RAY;
        emit(sp);
    } break;
    case RPX_MAT_PEC: {
        if constexpr ((F & M_SIMPLE) == 0) return;
        RayS sp = begin_child(in, length, normal_f, point, idx);
        V3 d = v3_of(in.direction);
        sp.direction = f3_of(reflect_dir(d, normal, dot(normal, d)));
        sp.E1 = cfneg(sp.E1);
        sp.ray_type_id |= RPX_REFL_RAY;
        emit(sp);
    } break;
    case RPX_MAT_PARTIAL: {
        if constexpr ((F & M_SIMPLE) == 0) return;
        V3 in_dir = norm(v3_of(in.direction));
        RayS sp = begin_child(in, length, normal_f, point, idx);
        RayS sp2 = sp;
        const float Rr = sqrtf((float)m.reflectivity), T = sqrtf((float)(1 - m.reflectivity));
        sp.direction = f3_of(reflect_dir(in_dir, normal, dot(normal, in_dir)));
        sp.E1 = cfscale(sp.E1, Rr);
        sp.E2 = cfscale(sp.E2, Rr);
        sp.ray_type_id |= RPX_REFL_RAY;
        emit(sp);
        sp2.direction = f3_of(in_dir);
        sp2.E1 = cfscale(sp2.E1, T);
        sp2.E2 = cfscale(sp2.E2, T);
        sp2.ray_type_id &= ~RPX_REFL_RAY;
        emit(sp2);
    } break;
    case RPX_MAT_DIELECTRIC: {
        if constexpr ((F & M_SIMPLE) == 0) return;
        V3 in_dir = norm(v3_of(in.direction));
        RayS sp = begin_child(in, length, normal_f, point, idx);
        const double cosTheta = dot(normal, in_dir);
        const double cos1 = fabs(cosTheta);
        double n1, n2;
        int flip;
        if (cosTheta < 0.0) {
            n1 = m.n_outside[0]; n2 = m.n_inside[0];
            sp.n = cf((float)m.n_inside[0], (float)m.n_inside[1]);
            flip = 1;
        } else {
            n1 = m.n_inside[0]; n2 = m.n_outside[0];
            sp.n = cf((float)m.n_outside[0], (float)m.n_outside[1]);
            flip = -1;
        }
        const double q = n2 / n1;
        if ((cosTheta * cosTheta) + (q * q - 1) < 0.0) {   // total internal reflection
            sp.direction = f3_of(reflect_dir(in_dir, normal, cosTheta));
            sp.E1 = cfneg(sp.E1);
            sp.E2 = cfneg(sp.E2);
            sp.ray_type_id |= RPX_REFL_RAY;
        } else {
            V3 t = refract_dir(in_dir, normal, cosTheta, n1 / n2, flip);
            const float cos2 = (float)fabs(dot(t, normal));
            const float c1 = (float)cos1, fn1 = (float)n1, fn2 = (float)n2;
            const float aspect = sqrtf(cos2 / c1) * (2 * fn1 * c1);
            const float T_p = aspect / (fn2 * c1 + fn1 * cos2);
            const float T_s = aspect / (fn2 * cos2 + fn1 * c1);
            sp.direction = f3_of(t);
            sp.E1 = cfscale(sp.E1, T_s);
            sp.E2 = cfscale(sp.E2, T_p);
            sp.ray_type_id &= ~RPX_REFL_RAY;
        }
        emit(sp);
    } break;
    case RPX_MAT_FULL_DIELECTRIC:
    case RPX_MAT_FULL_DISPERSIVE: {
        if constexpr ((F & M_FRESNEL) == 0) return;
        Sides sd = material_indices(m, m.type == RPX_MAT_FULL_DISPERSIVE ? wavelengths[in.wl_idx] : 0.0);
        V3 in_dir = norm(v3_of(in.direction));
        RayS sp = begin_child(in, length, normal_f, point, idx);
        const Cf E1 = sp.E1, E2 = sp.E2;
        const double cosTheta = dot(normal, in_dir);
        const double cos1d = fabs(cosTheta);
        const float cos1 = (float)cos1d;
        const float sin1 = sqrtf((float)fabs(1 - cos1d * cos1d));
        Cx n1d, n2d;
        int flip;
        if (cosTheta < 0.0) { n1d = sd.n_out; n2d = sd.n_in; flip = 1; }
        else { n1d = sd.n_in; n2d = sd.n_out; flip = -1; }
        const Cf n1 = cf_of(n1d), n2 = cf_of(n2d);
        const Cf sin2 = cfdiv(cfscale(n1, sin1), n2);
        const Cf cos2 = cfsqrt(cfsub(cf(1, 0), cfmul(sin2, sin2)));
        const float P_in = n1.re * (cfabs2(E1) + cfabs2(E2));
        if (P_in == 0.0f) return;
        const Cf n2c1 = cfscale(n2, cos1), n1c2 = cfmul(n1, cos2);
        const Cf n2c2 = cfmul(n2, cos2), n1c1 = cfscale(n1, cos1);
        Cf R_p = cfdiv(cfneg(cfsub(n2c1, n1c2)), cfadd(n2c1, n1c2));
        Cf R_s = cfdiv(cfneg(cfsub(n2c2, n1c1)), cfadd(n2c2, n1c1));
        R_s = cfmul(R_s, E1);
        R_p = cfmul(R_p, E2);
        if (power_frac_s(n1.re, R_s, R_p, P_in) > (float)m.reflection_threshold) {
            RayS c = sp;
            c.direction = f3_of(reflect_dir(in_dir, normal, cosTheta));
            c.E1 = R_s;
            c.E2 = cfneg(R_p);
            c.n = n1;
            c.ray_type_id |= RPX_REFL_RAY;
            emit(c);
            sp = c;
        }
        const V3 t = refract_dir(in_dir, normal, cosTheta, n1d.re / n2d.re, flip);
        const float aspect = sqrtf(cos2.re / cos1);
        const Cf num = cfscale(n1, aspect * 2.0f * cos1);
        Cf T_p = cfdiv(num, cfadd(n2c1, n1c2));
        Cf T_s = cfdiv(num, cfadd(n2c2, n1c1));
        T_s = cfmul(T_s, E1);
        T_p = cfmul(T_p, E2);
        if (power_frac_s(n2.re, T_s, T_p, P_in) > (float)m.transmission_threshold) {
            sp.direction = f3_of(t);
            sp.E1 = T_s;
            sp.E2 = T_p;
            sp.n = n2;
            sp.ray_type_id &= ~RPX_REFL_RAY;
            emit(sp);
        }
    } break;
    case RPX_MAT_SINGLE_LAYER:
    case RPX_MAT_COATED_DISPERSIVE: {
        if constexpr ((F & M_COATED) == 0) return;
        const double wavelength = wavelengths[in.wl_idx];
        Sides sd = material_indices(m, wavelength);
        V3 in_dir = norm(v3_of(in.direction));
        RayS sp = begin_child(in, length, normal_f, point, idx);
        const Cf E1 = sp.E1, E2 = sp.E2;
        const double cosTheta = dot(normal, in_dir);
        const double cos1d = fabs(cosTheta);
        const float cos1 = (float)cos1d;
        const float sin1 = sqrtf((float)fabs(1 - cos1d * cos1d));
        Cx n1d, n3d;
        int flip;
        if (cosTheta < 0.0) { n1d = sd.n_out; n3d = sd.n_in; flip = 1; }
        else { n1d = sd.n_in; n3d = sd.n_out; flip = -1; }
        const Cf n1 = cf_of(n1d), n2 = cf_of(sd.n_coat), n3 = cf_of(n3d);
        const Cf n1s1 = cfscale(n1, sin1);
        const Cf sin2 = cfdiv(n1s1, n2);
        const Cf cos2 = cfsqrt(cfsub(cf(1, 0), cfmul(sin2, sin2)));
        const Cf sin3 = cfdiv(n1s1, n3);
        const Cf cos3 = cfsqrt(cfsub(cf(1, 0), cfmul(sin3, sin3)));
        const float P_in = n1.re * (cfabs2(E1) + cfabs2(E2));
        if (P_in == 0.0f) return;
        const Cf n1cos1 = cfscale(n1, cos1), n2cos2 = cfmul(n2, cos2), n3cos3 = cfmul(n3, cos3);
        const float dwc = (float)(2 * M_PI * m.thickness / wavelength);
        const Cf phi = cfdiv(cfmul(cf(0.0f, -dwc), cfsub(n2, cfmul(sin2, sin2))), cos2);
        const Cf ep1 = cfdiv(cfexp(phi), cfmul(cfscale(n2cos2, 4.0f), n3cos3));
        const Cf ep2 = cfexp(cfscale(phi, -2.0f));
        const Cf mep1 = cfneg(ep1);
        // s-polarisation
        Cf a = cfsub(n1cos1, n2cos2), b = cfadd(n1cos1, n2cos2);
        Cf c = cfadd(n2cos2, n3cos3), d = cfsub(n2cos2, n3cos3);
        Cf M00 = cfmul(mep1, cfadd(cfmul(a, c), cfmul(cfmul(b, d), ep2)));
        Cf M01 = cfmul(ep1, cfadd(cfmul(cfmul(a, d), ep2), cfmul(b, c)));
        Cf M10 = cfmul(ep1, cfadd(cfmul(a, d), cfmul(cfmul(b, c), ep2)));
        Cf M11 = cfmul(mep1, cfadd(cfmul(cfmul(a, c), ep2), cfmul(b, d)));
        Cf R_s = cfdiv(cfneg(M00), M01);
        Cf T_s = cfadd(M10, cfmul(M11, R_s));
        // p-polarisation
        const Cf n1cos2 = cfmul(n1, cos2), n2cos1 = cfscale(n2, cos1);
        const Cf n2cos3 = cfmul(n2, cos3), n3cos2 = cfmul(n3, cos2);
        a = cfsub(n1cos2, n2cos1); b = cfadd(n1cos2, n2cos1);
        c = cfadd(n2cos3, n3cos2); d = cfsub(n2cos3, n3cos2);
        M00 = cfmul(mep1, cfadd(cfmul(a, c), cfmul(cfmul(b, d), ep2)));
        M01 = cfmul(ep1, cfadd(cfmul(cfmul(a, d), ep2), cfmul(b, c)));
        M10 = cfmul(ep1, cfadd(cfmul(a, d), cfmul(cfmul(b, c), ep2)));
        M11 = cfmul(mep1, cfadd(cfmul(cfmul(a, c), ep2), cfmul(b, d)));
        Cf R_p = cfdiv(cfneg(M00), M01);
        Cf T_p = cfadd(M10, cfmul(M11, R_p));
        R_s = cfmul(R_s, E1);
        R_p = cfmul(R_p, E2);
        if (power_frac_s(n1.re, R_s, R_p, P_in) > (float)m.reflection_threshold) {
            RayS cr = sp;
            cr.direction = f3_of(reflect_dir(in_dir, normal, cosTheta));
            cr.E1 = R_s;
            cr.E2 = cfneg(R_p);
            cr.n = n1;
            cr.ray_type_id |= RPX_REFL_RAY;
            emit(cr);
            sp = cr;
        }
        const V3 t = refract_dir(in_dir, normal, cosTheta, n1d.re / n3d.re, flip);
        const float aspect = sqrtf(cos3.re / cos1);
        T_s = cfmul(T_s, cfscale(E1, aspect));
        T_p = cfmul(T_p, cfscale(E2, aspect));
        if (power_frac_s(n3.re, T_s, T_p, P_in) > (float)m.transmission_threshold) {
            sp.direction = f3_of(t);
            sp.E1 = T_s;
            sp.E2 = T_p;
            sp.n = n3;
            sp.ray_type_id &= ~RPX_REFL_RAY;
            emit(sp);
        }
    } break;
    case RPX_MAT_GRATING: {
        if constexpr ((F & M_GRATING) == 0) return;
        grating_children_split(m, in, length, idx, point, orient, wavelengths, emit);
    } break;
    default:  // RPX_MAT_NONE, RPX_MAT_OPAQUE: no children
        break;
    }
}

}  // namespace rpx
