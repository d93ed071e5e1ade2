// Interface materials: child-ray emission and parabasal directions (FP64).
//
// Children are produced through an emitter functor in the reference's order
// (reflected before transmitted, cmaterials.pyx:826-872), so one routine serves
// both the counting pass and the writing pass of a generation.
#pragma once
#include "rpx_faces.h"

namespace rpx {

struct Ray {
    V3 origin, direction, normal, E_vector;
    Cx n, E1, E2;
    double length, phase, accum;
    uint32_t wl_idx, parent_idx, end_face_idx, ray_ident, ray_type_id;
};

// Emitters take children in order through operator(); kCap is how many a
// material should build and hand over, skip() counts one more child the
// emitter does not take (a multi-order grating's third order, built again
// after the children's positions are known).
//
// PairEmit collects the (at most two) children of one ray as a two-slot
// shift register: r0 is the latest child, r1 the one before.  Conditional
// or indexed slot writes make the compiler keep the pair in scratch memory;
// the unconditional shift keeps both in registers.
struct PairEmit {
    static constexpr int kCap = 2;
    Ray r0, r1;
    int n = 0;
    RPX_DEV void operator()(const Ray &x) {
        r1 = r0;
        r0 = x;
        ++n;
    }
    RPX_DEV void skip() { ++n; }
    RPX_DEV const Ray &first() const { return n > 1 ? r1 : r0; }
};

// Order k (0-based) of a grating: `order`, then the signed bytes of
// more_orders (multi-order gratings, n_orders 2 or 3).
RPX_DEV int grating_order(const rpx_material_t &m, int k) {
    return k == 0 ? m.order : (int)(int8_t)((uint32_t)m.more_orders >> (8 * (k - 1)));
}
RPX_DEV int grating_orders(const rpx_material_t &m) { return m.n_orders > 1 ? m.n_orders : 1; }

RPX_DEV Ray load_ray(const rpx_rays_t &r, uint64_t i) {
    const double *f = r.f64;
    const uint64_t c = r.capacity;
    Ray o;
    o.origin = v3(f[RPX_OX * c + i], f[RPX_OY * c + i], f[RPX_OZ * c + i]);
    o.direction = v3(f[RPX_DX * c + i], f[RPX_DY * c + i], f[RPX_DZ * c + i]);
    o.normal = v3(f[RPX_NX * c + i], f[RPX_NY * c + i], f[RPX_NZ * c + i]);
    o.E_vector = v3(f[RPX_EX * c + i], f[RPX_EY * c + i], f[RPX_EZ * c + i]);
    o.n = cx(f[RPX_N_RE * c + i], f[RPX_N_IM * c + i]);
    o.E1 = cx(f[RPX_E1_RE * c + i], f[RPX_E1_IM * c + i]);
    o.E2 = cx(f[RPX_E2_RE * c + i], f[RPX_E2_IM * c + i]);
    o.length = f[RPX_LENGTH * c + i];
    o.phase = f[RPX_PHASE * c + i];
    o.accum = f[RPX_ACCUM * c + i];
    const uint32_t *u = r.u32;
    o.wl_idx = u[RPX_WL_IDX * c + i];
    o.parent_idx = u[RPX_PARENT_IDX * c + i];
    o.end_face_idx = u[RPX_END_FACE_IDX * c + i];
    o.ray_ident = u[RPX_RAY_IDENT * c + i];
    o.ray_type_id = u[RPX_RAY_TYPE_ID * c + i];
    return o;
}

RPX_DEV void store_ray(const rpx_rays_t &r, uint64_t i, const Ray &o) {
    double *f = r.f64;
    const uint64_t c = r.capacity;
    f[RPX_OX * c + i] = o.origin.x;  f[RPX_OY * c + i] = o.origin.y;  f[RPX_OZ * c + i] = o.origin.z;
    f[RPX_DX * c + i] = o.direction.x; f[RPX_DY * c + i] = o.direction.y; f[RPX_DZ * c + i] = o.direction.z;
    f[RPX_NX * c + i] = o.normal.x;  f[RPX_NY * c + i] = o.normal.y;  f[RPX_NZ * c + i] = o.normal.z;
    f[RPX_EX * c + i] = o.E_vector.x; f[RPX_EY * c + i] = o.E_vector.y; f[RPX_EZ * c + i] = o.E_vector.z;
    f[RPX_N_RE * c + i] = o.n.re;    f[RPX_N_IM * c + i] = o.n.im;
    f[RPX_E1_RE * c + i] = o.E1.re;  f[RPX_E1_IM * c + i] = o.E1.im;
    f[RPX_E2_RE * c + i] = o.E2.re;  f[RPX_E2_IM * c + i] = o.E2.im;
    f[RPX_LENGTH * c + i] = o.length;
    f[RPX_PHASE * c + i] = o.phase;
    f[RPX_ACCUM * c + i] = o.accum;
    uint32_t *u = r.u32;
    u[RPX_WL_IDX * c + i] = o.wl_idx;
    u[RPX_PARENT_IDX * c + i] = o.parent_idx;
    u[RPX_END_FACE_IDX * c + i] = o.end_face_idx;
    u[RPX_RAY_IDENT * c + i] = o.ray_ident;
    u[RPX_RAY_TYPE_ID * c + i] = o.ray_type_id;
}

// Parabasal ray (ctracer.pxd:54-58)
struct Para {
    V3 origin, direction, normal;
    double length;
};

// ------------------------------------------------------------- dispersion
// BaseDispersionCurve.c_evaluate_n (cmaterials.pyx:206-228) for one
// wavelength; the host rejects out-of-range wavelengths beforehand.
RPX_DEV Cx eval_dispersion(const rpx_dispersion_t &d, double wvl) {
    double nre;
    const double *c = d.coefs;
    switch (d.formula_id) {
    case 0:  // nondispersive_0 (cmaterials.pyx:107)
        nre = c[0];
        break;
    case 1: {  // sellmeier_1 (cmaterials.pyx:110-119)
        double n2 = 1.0, wl2 = wvl * wvl;
        n2 += c[0];
        for (int i = 0; i < (d.n_coefs - 1) / 2; ++i)
            n2 += c[2 * i + 1] * wl2 / (wl2 - c[2 * i + 2] * c[2 * i + 2]);
        nre = sqrt(n2);
    } break;
    case 2: {  // sellmeier_2 (cmaterials.pyx:121-130)
        double n2 = 1.0, wl2 = wvl * wvl;
        n2 += c[0];
        for (int i = 0; i < (d.n_coefs - 1) / 2; ++i)
            n2 += c[2 * i + 1] * wl2 / (wl2 - c[2 * i + 2]);
        nre = sqrt(n2);
    } break;
    case 3: {  // sellmeier_3 polynomial (cmaterials.pyx:133-140)
        double n2 = c[0];
        for (int i = 1; i < d.n_coefs - 1; i += 2) n2 += c[i] * pow(wvl, c[i + 1]);
        nre = sqrt(n2);
    } break;
    default:  // sellmeier_5 returns nothing (cmaterials.pyx:143-149): 0.0
        nre = 0.0;
        break;
    }
    double n_imag = 0.0001 * d.absorption / (4 * M_PI);
    return cx(nre, n_imag * wvl);
}

// ------------------------------------------------------------- S/P basis
// convert_to_sp (cmaterials.pyx:49-91)
RPX_DEV Ray convert_to_sp(Ray ray, V3 normal) {
    Cx E1 = ray.E1, E2 = ray.E2;
    V3 E2v = norm(cross(ray.direction, ray.E_vector));
    V3 E1v = norm(cross(E2v, ray.direction));
    normal = norm(normal);
    V3 S = cross(ray.direction, normal);
    if (fabs(S.x) < 1.0e-10 && fabs(S.y) < 1.0e-10 && fabs(S.z) < 1.0e-10) return ray;
    S = norm(S);
    V3 v = cross(ray.direction, S);
    V3 P = norm(v);
    double A = dot(E1v, S);
    double B = dot(E2v, S);
    Cx Samp = cx(E1.re * A + E2.re * B, E1.im * A + E2.im * B);
    B = dot(E1v, P);
    A = dot(E2v, P);
    Cx Pamp = cx(E1.re * B + E2.re * A, E1.im * B + E2.im * A);
    ray.E_vector = S;
    ray.E1 = Samp;
    ray.E2 = Pamp;
    return ray;
}

RPX_DEV Cx R(double x) { return cx(x, 0.0); }

// Indices on the incident (n1), coating (n2) and far (n3) sides.
struct Sides {
    Cx n_in, n_out, n_coat;
};

RPX_DEV Sides material_indices(const rpx_material_t &m, double wvl) {
    Sides s;
    if (m.type == RPX_MAT_FULL_DISPERSIVE || m.type == RPX_MAT_COATED_DISPERSIVE) {
        s.n_in = eval_dispersion(m.disp_inside, wvl);
        s.n_out = eval_dispersion(m.disp_outside, wvl);
        s.n_coat = (m.type == RPX_MAT_COATED_DISPERSIVE) ? eval_dispersion(m.disp_coating, wvl) : cx(0, 0);
    } else {
        s.n_in = cx(m.n_inside[0], m.n_inside[1]);
        s.n_out = cx(m.n_outside[0], m.n_outside[1]);
        s.n_coat = cx(m.n_coating[0], m.n_coating[1]);
    }
    return s;
}

// Wavelength table lookup; an index outside the table yields NaN (the ray's
// children then fail every threshold test) instead of a fault.
struct WlTable {
    const double *wl;
    uint64_t n;
    RPX_DEV double operator[](uint32_t i) const { return (i < n) ? wl[i] : __builtin_nan(""); }
};

// Whether a material can have children: OpaqueMaterial and the base
// InterfaceMaterial never do, so a ray ending on them needs no material
// evaluation and none of its record beyond the geometry.
RPX_DEV bool material_emits(const rpx_material_t &m) {
    return m.type != RPX_MAT_NONE && m.type != RPX_MAT_OPAQUE;
}

RPX_DEV double power_frac(double nre, Cx a, Cx b, double P_in) {
    double ca = cabs_(a), cb = cabs_(b);
    return nre * (ca * ca + cb * cb) / P_in;
}

// ------------------------------------------------------------- children
// InterfaceMaterial.eval_child_ray_c of every supported material.
// DiffractionGratingMaterial.eval_child_ray_c (cmaterials.pyx:1472-1542):
// one child per propagating order, orders in list order (one order in the
// reference); children past the emitter's kCap are counted, not built.
template <class Emit>
RPX_DEV void grating_children(const rpx_material_t &m, const Ray &in, uint32_t idx, V3 point,
                              const Orientation &orient, const WlTable &wavelengths, Emit &emit) {
    V3 normal = norm(orient.normal);
    V3 tangent = norm(orient.tangent);
    V3 tangent2 = cross(normal, tangent);
    double wavelen = wavelengths[in.wl_idx];
    double line_spacing = 1000.0 / m.lines_per_mm;
    const V3 dir = norm(in.direction);
    const double k_z0 = dot(normal, dir);
    const double k_y = dot(tangent2, dir);
    const double k_x0 = dot(tangent, dir);
    int sign = (k_z0 < 0.0) ? 1 : -1;
    double n_re = in.n.re;
    const int n_orders = grating_orders(m);
    int built = 0;
    for (int o = 0; o < n_orders; ++o) {
        const int order = grating_order(m, o);
        double k_x = k_x0 - order * wavelen / (line_spacing * n_re);
        double k_z = 1 - (k_x * k_x) - (k_y * k_y);
        if (k_z < 0) continue;
        if (built >= Emit::kCap) {
            emit.skip();
            continue;
        }
        ++built;
        k_z = sign * sqrt(k_z);
        V3 reflected = mul(tangent, k_x);
        reflected = add(reflected, mul(tangent2, k_y));
        reflected = add(reflected, mul(normal, k_z));
        Ray sp = convert_to_sp(in, normal);
        sp.accum += sp.length * sp.n.re;
        sp.origin = point;
        sp.normal = normal;
        sp.direction = reflected;
        sp.E1.re = -sp.E1.re * m.efficiency;
        sp.E1.im = -sp.E1.im * m.efficiency;
        sp.E2.re = sp.E2.re * m.efficiency;
        sp.E2.im = sp.E2.im * m.efficiency;
        sp.parent_idx = idx;
        sp.ray_type_id |= RPX_REFL_RAY;
        V3 og = v3(m.origin[0], m.origin[1], m.origin[2]);
        sp.phase += 1000.0 * dot(sub(og, point), tangent) * order * 2 * M_PI / line_spacing;
        emit(sp);
    }
}

template <uint32_t F, class Emit>
RPX_DEV void eval_children(const rpx_material_t &m, const Ray &in, uint32_t idx, V3 point,
                           const Orientation &orient, const WlTable &wavelengths, Emit &emit) {
    switch (m.type) {
    case RPX_MAT_TRANSPARENT: {
        if constexpr ((F & M_SIMPLE) == 0) return;  // cmaterials.pyx:260-278
        V3 normal = norm(orient.normal);
        Ray sp = convert_to_sp(in, normal);
        sp.accum += sp.length * sp.n.re;
        sp.origin = point;
        sp.normal = normal;
        sp.length = RPX_INF;
        sp.parent_idx = idx;
        sp.ray_type_id &= ~RPX_REFL_RAY;
        emit(sp);
    } break;
    case RPX_MAT_PEC: {
        if constexpr ((F & M_SIMPLE) == 0) return;  // cmaterials.pyx:285-319
        V3 normal = norm(orient.normal);
        Ray sp = convert_to_sp(in, normal);
        double cosTheta = dot(normal, in.direction);
        V3 ctn = mul(normal, cosTheta);
        V3 reflected = sub(in.direction, mul(ctn, 2));
        sp.accum += sp.length * sp.n.re;
        sp.origin = point;
        sp.normal = normal;
        sp.direction = reflected;
        sp.length = RPX_INF;
        sp.E1.re = -sp.E1.re;
        sp.E1.im = -sp.E1.im;
        sp.parent_idx = idx;
        sp.ray_type_id |= RPX_REFL_RAY;
        emit(sp);
    } break;
    case RPX_MAT_PARTIAL: {
        if constexpr ((F & M_SIMPLE) == 0) return;  // cmaterials.pyx:345-397
        V3 normal = norm(orient.normal);
        V3 in_dir = norm(in.direction);
        Ray sp = convert_to_sp(in, normal);
        double cosTheta = dot(normal, in_dir);
        V3 ctn = mul(normal, cosTheta);
        sp.accum += sp.length * sp.n.re;
        Ray sp2 = sp;
        double Rr = sqrt(m.reflectivity);
        double T = sqrt(1 - m.reflectivity);
        V3 reflected = sub(in_dir, mul(ctn, 2));
        sp.origin = point;
        sp.normal = normal;
        sp.direction = reflected;
        sp.length = RPX_INF;
        sp.E1 = cmul(sp.E1, R(Rr));
        sp.E2 = cmul(sp.E2, R(Rr));
        sp.parent_idx = idx;
        sp.ray_type_id |= RPX_REFL_RAY;
        emit(sp);
        sp2.origin = point;
        sp2.normal = normal;
        sp2.direction = in_dir;
        sp2.length = RPX_INF;
        sp2.E1 = cmul(sp2.E1, R(T));
        sp2.E2 = cmul(sp2.E2, R(T));
        sp2.parent_idx = idx;
        sp2.ray_type_id &= ~RPX_REFL_RAY;
        emit(sp2);
    } break;
    case RPX_MAT_DIELECTRIC: {
        if constexpr ((F & M_SIMPLE) == 0) return;  // cmaterials.pyx:587-680
        V3 normal = norm(orient.normal);
        V3 in_dir = norm(in.direction);
        Ray sp = convert_to_sp(in, normal);
        sp.accum += sp.length * sp.n.re;
        double cosTheta = dot(normal, in_dir);
        double cos1 = fabs(cosTheta);
        double n1, n2;
        int flip;
        if (cosTheta < 0.0) {
            n1 = m.n_outside[0];
            n2 = m.n_inside[0];
            sp.n = cx(m.n_inside[0], m.n_inside[1]);
            flip = 1;
        } else {
            n1 = m.n_inside[0];
            n2 = m.n_outside[0];
            sp.n = cx(m.n_outside[0], m.n_outside[1]);
            flip = -1;
        }
        double q = n2 / n1;
        double N2 = q * q;
        double N2_sin2 = (cosTheta * cosTheta) + (N2 - 1);
        V3 ctn = mul(normal, cosTheta);
        if (N2_sin2 < 0.0) {
            V3 reflected = sub(in_dir, mul(ctn, 2));
            sp.origin = point;
            sp.normal = normal;
            sp.direction = reflected;
            sp.length = RPX_INF;
            sp.E1.re *= -1; sp.E1.im *= -1;
            sp.E2.re *= -1; sp.E2.im *= -1;
            sp.parent_idx = idx;
            sp.ray_type_id |= RPX_REFL_RAY;
        } else {
            V3 tangent = sub(in_dir, ctn);
            V3 tg2 = mul(tangent, n1 / n2);
            double tan_mag_sq = mag_sq(tg2);
            double c2 = sqrt(1 - tan_mag_sq);
            V3 transmitted = sub(tg2, mul(normal, c2 * flip));
            double cos2 = fabs(dot(transmitted, normal));
            double Two_n1_cos1 = (2 * n1) * cos1;
            double aspect = sqrt(cos2 / cos1) * Two_n1_cos1;
            double T_p = aspect / (n2 * cos1 + n1 * cos2);
            double T_s = aspect / (n2 * cos2 + n1 * cos1);
            sp.origin = point;
            sp.normal = normal;
            sp.direction = transmitted;
            sp.length = RPX_INF;
            sp.E1.re *= T_s; sp.E1.im *= T_s;
            sp.E2.re *= T_p; sp.E2.im *= T_p;
            sp.parent_idx = idx;
            sp.ray_type_id &= ~RPX_REFL_RAY;
        }
        emit(sp);
    } break;
    case RPX_MAT_FULL_DIELECTRIC:
    case RPX_MAT_FULL_DISPERSIVE: {
        if constexpr ((F & M_FRESNEL) == 0) return;  // cmaterials.pyx:755-872, 896-1013
        Sides sd = material_indices(m, m.type == RPX_MAT_FULL_DISPERSIVE ? wavelengths[in.wl_idx] : 0.0);
        V3 normal = norm(orient.normal);
        V3 in_dir = norm(in.direction);
        Ray sp = convert_to_sp(in, normal);
        sp.accum += sp.length * sp.n.re;
        Cx E1 = sp.E1, E2 = sp.E2;
        double cosTheta = dot(normal, in_dir);
        double cos1 = fabs(cosTheta);
        double sin1 = sqrt(fabs(1 - cos1 * cos1));
        Cx n1, n2;
        int flip;
        if (cosTheta < 0.0) { n1 = sd.n_out; n2 = sd.n_in; flip = 1; }
        else { n1 = sd.n_in; n2 = sd.n_out; flip = -1; }
        Cx sin2 = cdiv(cmul(n1, R(sin1)), n2);
        Cx cos2 = csqrt_(csub(R(1), cmul(sin2, sin2)));
        V3 ctn = mul(normal, cosTheta);
        double P_in = n1.re * (E1.re * E1.re + E1.im * E1.im + E2.re * E2.re + E2.im * E2.im);
        if (P_in == 0.0) return;
        Cx n2c1 = cmul(n2, R(cos1)), n1c2 = cmul(n1, cos2);
        Cx n2c2 = cmul(n2, cos2), n1c1 = cmul(n1, R(cos1));
        Cx R_p = cdiv(cneg(csub(n2c1, n1c2)), cadd(n2c1, n1c2));
        Cx R_s = cdiv(cneg(csub(n2c2, n1c1)), cadd(n2c2, n1c1));
        R_s = cmul(R_s, E1);
        R_p = cmul(R_p, E2);
        if (power_frac(n1.re, R_s, R_p, P_in) > m.reflection_threshold) {
            Ray c = sp;
            c.origin = point;
            c.normal = normal;
            c.direction = sub(in_dir, mul(ctn, 2));
            c.length = RPX_INF;
            c.E1 = R_s;
            c.E2 = cx(-R_p.re, -R_p.im);
            c.parent_idx = idx;
            c.n = n1;
            c.ray_type_id |= RPX_REFL_RAY;
            emit(c);
            sp = c;
        }
        V3 tangent = sub(in_dir, ctn);
        V3 tg2 = mul(tangent, n1.re / n2.re);
        double tan_mag_sq = mag_sq(tg2);
        double c2 = sqrt(1 - tan_mag_sq);
        V3 transmitted = sub(tg2, mul(normal, c2 * flip));
        double aspect = sqrt(cos2.re / cos1);
        Cx num = cmul(R(aspect), cmul(R(2.0 * cos1), n1));
        Cx T_p = cdiv(num, cadd(n2c1, n1c2));
        Cx T_s = cdiv(num, cadd(n2c2, n1c1));
        T_s = cmul(T_s, E1);
        T_p = cmul(T_p, E2);
        if (power_frac(n2.re, T_s, T_p, P_in) > m.transmission_threshold) {
            sp.origin = point;
            sp.normal = normal;
            sp.direction = transmitted;
            sp.length = RPX_INF;
            sp.E1 = T_s;
            sp.E2 = T_p;
            sp.parent_idx = idx;
            sp.n = n2;
            sp.ray_type_id &= ~RPX_REFL_RAY;
            emit(sp);
        }
    } break;
    case RPX_MAT_SINGLE_LAYER:
    case RPX_MAT_COATED_DISPERSIVE: {
        if constexpr ((F & M_COATED) == 0) return;  // cmaterials.pyx:1026-1182, 1228-1391
        const double wavelength = wavelengths[in.wl_idx];  // both read _wavelengths
        Sides sd = material_indices(m, wavelength);
        V3 normal = norm(orient.normal);
        V3 in_dir = norm(in.direction);
        Ray sp = convert_to_sp(in, normal);
        sp.accum += sp.length * sp.n.re;
        Cx E1 = sp.E1, E2 = sp.E2;
        double cosTheta = dot(normal, in_dir);
        double cos1 = fabs(cosTheta);
        double sin1 = sqrt(fabs(1 - cos1 * cos1));
        Cx n2 = sd.n_coat, n1, n3;
        int flip;
        if (cosTheta < 0.0) { n1 = sd.n_out; n3 = sd.n_in; flip = 1; }
        else { n1 = sd.n_in; n3 = sd.n_out; flip = -1; }
        Cx n1s1 = cmul(n1, R(sin1));
        Cx sin2 = cdiv(n1s1, n2);
        Cx cos2 = csqrt_(csub(R(1), cmul(sin2, sin2)));
        Cx sin3 = cdiv(n1s1, n3);
        Cx cos3 = csqrt_(csub(R(1), cmul(sin3, sin3)));
        V3 ctn = mul(normal, cosTheta);
        double P_in = n1.re * (E1.re * E1.re + E1.im * E1.im + E2.re * E2.re + E2.im * E2.im);
        if (P_in == 0.0) return;
        Cx n1cos1 = cmul(n1, R(cos1));
        Cx n2cos2 = cmul(n2, cos2);
        Cx n3cos3 = cmul(n3, cos3);
        double dwc = 2 * M_PI * m.thickness / wavelength;
        Cx phi = cdiv(cmul(cmul(cx(-0.0, -1.0), R(dwc)), csub(n2, cmul(sin2, sin2))), cos2);
        Cx ep1 = cdiv(cexp_(phi), cmul(cmul(R(4), n2cos2), n3cos3));
        Cx ep2 = cexp_(cmul(R(-2), phi));
        Cx mep1 = cneg(ep1);
        // s-polarisation
        Cx a = csub(n1cos1, n2cos2), b = cadd(n1cos1, n2cos2);
        Cx c = cadd(n2cos2, n3cos3), d = csub(n2cos2, n3cos3);
        Cx M00 = cmul(mep1, cadd(cmul(a, c), cmul(cmul(b, d), ep2)));
        Cx M01 = cmul(ep1, cadd(cmul(cmul(a, d), ep2), cmul(b, c)));
        Cx M10 = cmul(ep1, cadd(cmul(a, d), cmul(cmul(b, c), ep2)));
        Cx M11 = cmul(mep1, cadd(cmul(cmul(a, c), ep2), cmul(b, d)));
        Cx R_s = cdiv(cneg(M00), M01);
        Cx T_s = cadd(M10, cmul(M11, R_s));
        // p-polarisation
        Cx n1cos2 = cmul(n1, cos2), n2cos1 = cmul(n2, R(cos1));
        Cx n2cos3 = cmul(n2, cos3), n3cos2 = cmul(n3, cos2);
        a = csub(n1cos2, n2cos1); b = cadd(n1cos2, n2cos1);
        c = cadd(n2cos3, n3cos2); d = csub(n2cos3, n3cos2);
        M00 = cmul(mep1, cadd(cmul(a, c), cmul(cmul(b, d), ep2)));
        M01 = cmul(ep1, cadd(cmul(cmul(a, d), ep2), cmul(b, c)));
        M10 = cmul(ep1, cadd(cmul(a, d), cmul(cmul(b, c), ep2)));
        M11 = cmul(mep1, cadd(cmul(cmul(a, c), ep2), cmul(b, d)));
        Cx R_p = cdiv(cneg(M00), M01);
        Cx T_p = cadd(M10, cmul(M11, R_p));
        R_s = cmul(R_s, E1);
        R_p = cmul(R_p, E2);
        if (power_frac(n1.re, R_s, R_p, P_in) > m.reflection_threshold) {
            Ray cr = sp;
            cr.origin = point;
            cr.normal = normal;
            cr.direction = sub(in_dir, mul(ctn, 2));
            cr.length = RPX_INF;
            cr.E1 = R_s;
            cr.E2 = cx(-R_p.re, -R_p.im);
            cr.parent_idx = idx;
            cr.n = n1;
            cr.ray_type_id |= RPX_REFL_RAY;
            emit(cr);
            sp = cr;
        }
        V3 tangent = sub(in_dir, ctn);
        V3 tg2 = mul(tangent, n1.re / n3.re);
        double tan_mag_sq = mag_sq(tg2);
        double c2 = sqrt(1 - tan_mag_sq);
        V3 transmitted = sub(tg2, mul(normal, c2 * flip));
        double aspect = sqrt(cos3.re / cos1);
        T_s = cmul(T_s, cmul(E1, R(aspect)));
        T_p = cmul(T_p, cmul(E2, R(aspect)));
        if (power_frac(n3.re, T_s, T_p, P_in) > m.transmission_threshold) {
            sp.origin = point;
            sp.normal = normal;
            sp.direction = transmitted;
            sp.length = RPX_INF;
            sp.E1 = T_s;
            sp.E2 = T_p;
            sp.parent_idx = idx;
            sp.n = n3;
            sp.ray_type_id &= ~RPX_REFL_RAY;
            emit(sp);
        }
    } break;
    case RPX_MAT_GRATING: {
        if constexpr ((F & M_GRATING) == 0) return;
        grating_children(m, in, idx, point, orient, wavelengths, emit);
    } break;
    default:  // RPX_MAT_NONE, RPX_MAT_OPAQUE: no children
        break;
    }
}

// eval_parabasal_ray_c: InterfaceMaterial default (ctracer.pyx:1588-1610),
// DielectricMaterial (cmaterials.pyx:683-724, inherited by FullDielectric and
// SingleLayerCoated), CoatedDispersiveMaterial (:1393-1434), grating (:1544-1599).
template <uint32_t F>
RPX_DEV Para eval_parabasal(const rpx_material_t &m, const Ray &base, V3 direction, V3 point,
                            const Orientation &orient, uint32_t ray_type_id, const WlTable &wavelengths) {
    Para out;
    V3 normal = norm(orient.normal);
    switch (m.type) {
    case RPX_MAT_DIELECTRIC:
    case RPX_MAT_FULL_DIELECTRIC:
    case RPX_MAT_SINGLE_LAYER:
    case RPX_MAT_COATED_DISPERSIVE: {
        if constexpr ((F & (M_SIMPLE | M_FRESNEL | M_COATED)) == 0) break;
        direction = norm(direction);
        double cosTheta = dot(normal, direction);
        V3 ctn = mul(normal, cosTheta);
        double n1, n2;
        int flip;
        double nin, nout;
        if (m.type == RPX_MAT_COATED_DISPERSIVE) {
            double wvl = wavelengths[base.wl_idx];
            nin = eval_dispersion(m.disp_inside, wvl).re;
            nout = eval_dispersion(m.disp_outside, wvl).re;
        } else {
            nin = m.n_inside[0];
            nout = m.n_outside[0];
        }
        if (cosTheta < 0.0) { n1 = nout; n2 = nin; flip = 1; }
        else { n1 = nin; n2 = nout; flip = -1; }
        if (ray_type_id & RPX_REFL_RAY) {
            out.direction = sub(direction, mul(ctn, 2));
        } else {
            V3 tangent = sub(direction, ctn);
            V3 tg2 = mul(tangent, n1 / n2);
            double tan_mag_sq = mag_sq(tg2);
            double c2 = sqrt(1 - tan_mag_sq);
            out.direction = sub(tg2, mul(normal, c2 * flip));
        }
    } break;
    case RPX_MAT_GRATING: {
        if constexpr ((F & M_GRATING) == 0) break;
        V3 tangent = norm(orient.tangent);
        V3 tangent2 = cross(normal, tangent);
        double wavelen = wavelengths[base.wl_idx];
        double line_spacing = 1000.0 / m.lines_per_mm;
        V3 reflected = norm(direction);
        double k_z = dot(normal, reflected);
        double k_y = dot(tangent2, reflected);
        double k_x = dot(tangent, reflected);
        int sign = (k_z < 0.0) ? 1 : -1;
        k_x = k_x - m.order * wavelen / (line_spacing * base.n.re);
        k_z = 1 - (k_x * k_x) - (k_y * k_y);
        k_z = sign * sqrt(k_z);
        reflected = mul(tangent, k_x);
        reflected = add(reflected, mul(tangent2, k_y));
        reflected = add(reflected, mul(normal, k_z));
        out.direction = reflected;
    } break;
    default:
        if (ray_type_id & RPX_REFL_RAY) {
            double cosTheta = dot(normal, direction);
            V3 ctn = mul(normal, cosTheta);
            out.direction = sub(direction, mul(ctn, 2));
        } else {
            out.direction = direction;
        }
        break;
    }
    out.origin = point;
    out.normal = normal;
    out.length = RPX_INF;
    return out;
}

}  // namespace rpx
