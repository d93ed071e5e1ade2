// Face intersection, normals, aperture shapes and Zernike distortions (FP64).
//
// Each routine follows the arithmetic of the reference face it stands in for;
// the reference line is cited beside it.  `is_base` is the reference's
// is_base_ray flag: parabasal rays skip most aperture tests.
#pragma once
#include "rpx_internal.h"

namespace rpx {

// --------------------------------------------------------------- shapes
// PolygonShape.point_inside_c (cshapes.pyx:150-168)
RPX_DEV bool polygon_shape_inside(const double *pts, int size, double X, double Y) {
    bool ct = false;
    double y1 = pts[2 * (size - 1) + 1], x1 = pts[2 * (size - 1)];
    for (int i = 0; i < size; ++i) {
        double y2 = pts[2 * i + 1], x2 = pts[2 * i];
        if ((y1 <= Y && Y < y2) || (y2 <= Y && Y < y1)) {
            if ((x1 + ((Y - y1) * (x2 - x1)) / (y2 - y1)) > X) ct = !ct;
        }
        y1 = y2;
        x1 = x2;
    }
    return ct;
}

// point_in_polygon_c used by PolygonFace (cfaces.pyx:1050-1068)
RPX_DEV bool polygon_face_inside(const double *pts, int size, double X, double Y) {
    bool ct = false;
    double y1 = pts[2 * (size - 1) + 1], x1 = pts[2 * (size - 1)];
    for (int i = 0; i < size; ++i) {
        double y2 = pts[2 * i + 1], x2 = pts[2 * i];
        double h = (Y - y1) / (y2 - y1);
        if (0.0 < h && h <= 1.0) {
            double x = x1 + h * (x2 - x1);
            if (x > X) ct = !ct;
        }
        y1 = y2;
        x1 = x2;
    }
    return ct;
}

// Shape tree (cshapes.pyx:34-168) evaluated as a postfix program.
RPX_DEV bool shape_inside(const SceneView &S, int root, double x, double y) {
    if (root < 0) return true;
    uint32_t stack = 0;
    int sp = 0;
    int begin = S.shape_ops[root].begin;
    for (int i = begin; i <= root; ++i) {
        const rpx_shape_op_t &op = S.shape_ops[i];
        uint32_t v = 0;
        switch (op.op) {
        case RPX_SHAPE_TRUE:
            v = 1;
            break;
        case RPX_SHAPE_CIRCLE: {
            double dx = x - op.p[0], dy = y - op.p[1];
            v = ((dx * dx) + (dy * dy) < (op.p[2] * op.p[2])) ? 1u : 0u;
        } break;
        case RPX_SHAPE_RECT: {
            double dx = x - op.p[0], dy = y - op.p[1];
            v = ((2 * fabs(dx) < op.p[2]) && (2 * fabs(dy)) < op.p[3]) ? 1u : 0u;
        } break;
        case RPX_SHAPE_POLYGON:
            v = polygon_shape_inside(S.vertices + 2 * op.a, op.b, x, y) ? 1u : 0u;
            break;
        case RPX_SHAPE_NOT:
            --sp;
            v = ((stack >> sp) & 1u) ^ 1u;
            stack &= ~(1u << sp);
            break;
        default: {  // AND / OR / XOR
            --sp;
            uint32_t b = (stack >> sp) & 1u;
            stack &= ~(1u << sp);
            --sp;
            uint32_t a = (stack >> sp) & 1u;
            stack &= ~(1u << sp);
            v = op.op == RPX_SHAPE_AND ? (a & b) : op.op == RPX_SHAPE_OR ? (a | b) : (a ^ b);
        } break;
        }
        stack |= (v << sp);
        ++sp;
    }
    return ((stack >> (sp - 1)) & 1u) != 0;
}

// ------------------------------------------------------------- Zernike
// The memoised radial recursion of cdistortions.pyx:149-316 is expanded at
// scene build into integer polynomials in r (rpx_zernike_coef_t.poly, the
// exact values of the compiled recursion, slot sharing included); a
// coefficient's R, R', R/r are evaluated here by Horner's rule over the
// zero-padded RPX_ZPOLY coefficients (the loads are independent, so they
// issue together).  When every active lane evaluates the same distortion the
// tables are read through the constant address space, i.e. with scalar loads
// into SGPRs; otherwise (per-lane faces: normals, parabasal rays) with vector
// loads.  The scene tables are not written while kernels run.
typedef __attribute__((address_space(4))) const rpx_zernike_t ZernikeK;
typedef __attribute__((address_space(4))) const rpx_zernike_coef_t ZCoefK;

template <class CT>
RPX_DEV double zpoly(CT &c, int k, double r) {
    double v = c.poly[k][RPX_ZPOLY - 1];
#pragma unroll
    for (int i = RPX_ZPOLY - 2; i >= 0; --i) v = fma(v, r, c.poly[k][i]);
    return v;
}

// cos(k theta) and sin(k theta) for theta = atan2(y, x), from cos = x/r and
// sin = y/r by the angle-addition recurrence (equal to the reference's
// cos(m * theta) to a few ulp for the orders a distortion uses); at r = 0
// the reference's atan2 branch value is kept.
struct Harmonic {
    double c1, s1;      // cos theta, sin theta
    double theta;       // used only at r = 0
    bool origin;
};

RPX_DEV Harmonic harmonic(double x, double y, double r) {
    Harmonic h;
    h.origin = r == 0.0;
    h.theta = h.origin ? atan2(y, x) : 0.0;
    h.c1 = h.origin ? 1.0 : x / r;
    h.s1 = h.origin ? 0.0 : y / r;
    return h;
}

RPX_DEV void harmonic_k(const Harmonic &h, int k, double *ck, double *sk) {
    if (h.origin) {
        *ck = cos(k * h.theta);
        *sk = sin(k * h.theta);
        return;
    }
    double c = 1.0, s = 0.0;
    for (int i = 0; i < k; ++i) {
        const double cn = c * h.c1 - s * h.s1;
        const double sn = s * h.c1 + c * h.s1;
        c = cn;
        s = sn;
    }
    *ck = c;
    *sk = s;
}

// ZernikeDistortion.z_offset_c (cdistortions.pyx:416-450)
template <class ZT, class CT>
RPX_DEV double zernike_sag_t(ZT &Z, CT *coefs, double x, double y) {
    x /= Z.unit_radius;
    y /= Z.unit_radius;
    double r = sqrt(x * x + y * y);
    const Harmonic hm = harmonic(x, y, r);
    double z = 0.0;
    for (int i = 0; i < Z.n_coefs; ++i) {
        CT &c = coefs[Z.coef_begin + i];
        double ck, sk;
        harmonic_k(hm, c.m >= 0 ? c.m : -c.m, &ck, &sk);
        // cos(m theta), or -sin(m theta) = sin(|m| theta) for m < 0
        const double PH = (c.m >= 0) ? ck : sk;
        z += c.nval * zpoly(c, RPX_ZP_SAG, r) * PH;
    }
    return z;
}

RPX_DEV double zernike_sag(const SceneView &S, int zi, double x, double y) {
    const int z0 = __builtin_amdgcn_readfirstlane(zi);
    if (__ballot(zi != z0) == 0)
        return zernike_sag_t(((ZernikeK *)S.zernikes)[z0], (ZCoefK *)S.zcoefs, x, y);
    return zernike_sag_t(S.zernikes[zi], S.zcoefs, x, y);
}

// ZernikeDistortion.z_offset_and_gradient_c (cdistortions.pyx:456-514):
// returns (dz/dx, dz/dy, z).
template <class ZT, class CT>
RPX_DEV V3 zernike_grad_t(ZT &Z, CT *coefs, double x, double y) {
    x /= Z.unit_radius;
    y /= Z.unit_radius;
    double r = sqrt(x * x + y * y);
    const Harmonic hm = harmonic(x, y, r);
    V3 out = v3(0.0, 0.0, 0.0);
    const double st = hm.origin ? sin(hm.theta) : hm.s1, ct = hm.origin ? cos(hm.theta) : hm.c1;
    for (int i = 0; i < Z.n_coefs; ++i) {
        CT &c = coefs[Z.coef_begin + i];
        const int am = c.m >= 0 ? c.m : -c.m;
        double ck, sk;
        harmonic_k(hm, am, &ck, &sk);
        double PH, PHp;
        if (c.m >= 0) {
            PH = ck;                       // cos(m theta)
            PHp = (double)(-c.m) * sk;     // -m sin(m theta)
        } else {
            PH = sk;                       // -sin(m theta) = sin(|m| theta)
            PHp = (double)(-c.m) * ck;     // -m cos(m theta) = |m| cos(|m| theta)
        }
        const double R = zpoly(c, RPX_ZP_R, r), Rp = zpoly(c, RPX_ZP_RP, r), Ror = zpoly(c, RPX_ZP_ROR, r);
        const double N = c.nval;
        out.z += N * R * PH;
        out.x += N * (Rp * ct * PH + Ror * (-st) * PHp);
        out.y += N * (Rp * st * PH + Ror * (ct)*PHp);
    }
    out.x /= Z.unit_radius;
    out.y /= Z.unit_radius;
    return out;
}

RPX_DEV V3 zernike_grad(const SceneView &S, int zi, double x, double y) {
    const int z0 = __builtin_amdgcn_readfirstlane(zi);
    if (__ballot(zi != z0) == 0)
        return zernike_grad_t(((ZernikeK *)S.zernikes)[z0], (ZCoefK *)S.zcoefs, x, y);
    return zernike_grad_t(S.zernikes[zi], S.zcoefs, x, y);
}

// ----------------------------------------------------------- conic / asphere
// intersect_conic (cfaces.pyx:1695-1747)
RPX_DEV double intersect_conic(V3 a, V3 d, double curvature, double conic_const) {
    double beta = 1 + conic_const;
    double R = -curvature;
    double b2 = beta * beta;
    double A = b2 * (d.z * d.z) + beta * (d.x * d.x) + beta * (d.y * d.y);
    double B = -2 * R * beta * d.z + 2 * a.x * beta * d.x + 2 * a.y * beta * d.y + 2 * a.z * b2 * d.z;
    double C = -2 * R * a.z * beta + (a.x * a.x) * beta + (a.y * a.y) * beta + (a.z * a.z) * b2;
    double D = B * B - 4 * A * C;
    if (D < 0) return -1;
    D = sqrt(D);
    if (R * beta * d.z <= 0) return (-B + D) / (2 * A);
    return (-B - D) / (2 * A);
}

// x^2 .. x^nmax into pw[2..nmax].  x^2 is the rounded product (the
// reference's x**2); higher powers come from a double-double product chain
// rounded once, i.e. correctly rounded like the C library's pow for these
// integer exponents, at a fraction of pow's cost.
RPX_DEV void pow_chain(double x, double *pw, int nmax) {
    double hi = x * x;
    double lo = fma(x, x, -hi);
    pw[2] = hi;
    for (int k = 3; k <= nmax; ++k) {
        double p = hi * x;
        double e = fma(hi, x, -p);
        lo = fma(lo, x, e);
        double s = p + lo;
        lo = lo - (s - p);
        hi = s;
        pw[k] = hi + lo;
    }
}

struct Asph {
    double R, beta, A4, A6, A8, A10, A12, A14, A16, R3;
    V3 a, d;
};

// eval_aspheric_impf (cfaces.pyx:1854-1863)
RPX_DEV double asph_impf(const Asph &A, double alpha) {
    double tx = A.a.x + alpha * A.d.x, ty = A.a.y + alpha * A.d.y;
    double r2 = (tx * tx + ty * ty);
    double pw[9];
    pow_chain(r2, pw, 8);
    double out = r2;
    out /= A.R * (1 + sqrt(1 - A.beta * r2 / (A.R * A.R)));
    out -= A.a.z + alpha * A.d.z;
    out += A.A4 * pw[2] + A.A6 * pw[3] + A.A8 * pw[4] + A.A10 * pw[5] + A.A12 * pw[6] + A.A14 * pw[7] +
           A.A16 * pw[8];
    return out;
}

// eval_aspheric_grad (cfaces.pyx:1866-1882)
RPX_DEV double asph_grad(const Asph &A, double alpha) {
    double tx = A.a.x + alpha * A.d.x, ty = A.a.y + alpha * A.d.y;
    double r2 = (tx * tx + ty * ty);
    double dx = A.d.x * tx;
    double dy = A.d.y * ty;
    double pw[8];
    pow_chain(r2, pw, 7);
    double out = A.A10 * (10 * dx + 10 * dy) * pw[4];
    out += A.A12 * (12 * dx + 12 * dy) * pw[5];
    out += A.A14 * (14 * dx + 14 * dy) * pw[6];
    out += A.A16 * (16 * dx + 16 * dy) * pw[7];
    out += A.A4 * (4 * dx + 4 * dy) * (r2);
    out += A.A6 * (6 * dx + 6 * dy) * pw[2];
    out += A.A8 * (8 * dx + 8 * dy) * pw[3] - A.d.z;
    out += (2 * dx + 2 * dy) / (A.R * (sqrt(1 - A.beta * (r2) / (A.R * A.R)) + 1));
    double s = sqrt(1 - A.beta * r2 / (A.R * A.R));
    double s1 = s + 1;
    out += A.beta * (2 * dx + 2 * dy) * (r2) / (2 * A.R3 * s * (s1 * s1));
    return out;
}

// A base-ray hit on an aspheric face is accepted only inside the aperture,
// where the surface lies within z_height +/- p[12]; the converged root
// satisfies the sag equation to within the Newton tolerance.  A segment whose
// z-range misses that band cannot be accepted.
RPX_DEV bool asph_outside_band(const rpx_face_t &f, V3 p1, V3 p2, int is_base) {
    if (!is_base || f.p[12] < 0.0) return false;
    double len = sep(p1, p2);
    double m = f.p[12] + 1e-9 * (1.0 + fabs(f.p[1]) + f.p[12]) + 10.0 * f.p[10] * len * (1.0 + f.p[13]);
    double zlo = fmin(p1.z, p2.z), zhi = fmax(p1.z, p2.z);
    return zhi < f.p[1] - m || zlo > f.p[1] + m;
}

// A base ray's nearest-face search only accepts a face's hit closer than
// the nearest one so far.  p[14]..p[15] of a face record hold a z band (in
// its face list's frame, widened for rounding on the host; NaN when not
// bounded) containing every point at which the face can accept a base-ray
// hit; an aspheric face's Newton tolerance widens it by a multiple of the
// segment length, as in asph_outside_band.  If the part of the segment
// closer than the current nearest hit -- z from z1 over zspan -- misses the
// band, the face cannot give a nearer hit and its exact test is skipped.
// Skipping removes only candidates the exact test would reject, so the
// nearest face, its distance and the tie rule are unchanged.
RPX_DEV bool face_out_of_reach(const rpx_face_t &f, double z1, double zspan, double slen) {
    const double m = f.type == RPX_FACE_ASPHERIC ? 10.0 * f.p[10] * slen * (1.0 + f.p[13]) : 0.0;
    const double z2 = z1 + zspan;
    return fmax(z1, z2) < f.p[14] - m || fmin(z1, z2) > f.p[15] + m;
}

// ------------------------------------------------------------- intersect
// Face record k of a table in global memory, for an index that is the same
// in every lane: read through the constant address space, i.e. with scalar
// loads into SGPRs (word by word, so that the host pass of the compiler,
// which cannot copy-construct from that address space, accepts it).  Scene
// tables are not written while kernels run.
RPX_DEV rpx_face_t face_uniform(const rpx_face_t *faces, int k) {
    typedef __attribute__((address_space(4))) const uint64_t WordK;
    static_assert(sizeof(rpx_face_t) % 8 == 0, "face records are copied as 64-bit words");
    const WordK *w = (const WordK *)(faces + k);
    rpx_face_t f;
    uint64_t *d = (uint64_t *)&f;
#pragma unroll
    for (int i = 0; i < (int)(sizeof(rpx_face_t) / 8); ++i) d[i] = w[i];
    return f;
}

// Base-face intersection used by DistortionFace (no recursion into a second
// distortion level: the host refuses nested distortions).
template <uint32_t F>
RPX_DEV double plain_intersect(const SceneView &S, const rpx_face_t &f, V3 p1, V3 p2, int is_base) {
    const double tol = f.tolerance;
    switch (f.type) {
    case RPX_FACE_CIRCULAR: {
        if constexpr ((F & F_PLANE) == 0) return -1;  // cfaces.pyx:151-178
        double max_length = sep(p1, p2);
        double h = (f.p[2] - p1.z) / (p2.z - p1.z);
        double d = f.p[0];
        if ((h < tol) || (h > 1.0)) return -1;
        double X = p1.x + h * (p2.x - p1.x) - f.p[1];
        double Y = p1.y + h * (p2.y - p1.y);
        if (is_base && (X * X + Y * Y) > (d * d / 4)) return -1;
        return h * max_length;
    }
    case RPX_FACE_SHAPED_PLANAR: {
        if constexpr ((F & F_PLANE) == 0) return -1;  // cfaces.pyx:201-226
        double max_length = sep(p1, p2);
        double h = (f.p[0] - p1.z) / (p2.z - p1.z);
        if ((h < tol) || (h > 1.0)) return -1;
        double X = p1.x + h * (p2.x - p1.x);
        double Y = p1.y + h * (p2.y - p1.y);
        if (is_base && !shape_inside(S, f.shape, X, Y)) return -1;
        return h * max_length;
    }
    case RPX_FACE_RECTANGULAR: {
        if constexpr ((F & F_PLANE) == 0) return -1;  // cfaces.pyx:365-397
        double max_length = sep(p1, p2);
        double h = (f.p[3] - p1.z) / (p2.z - p1.z);
        double lngth = f.p[0], wdth = f.p[1];
        if ((h < tol) || (h > 1.0)) return -1;
        if (is_base) {
            double X = p1.x + h * (p2.x - p1.x) - f.p[2];
            double Y = p1.y + h * (p2.y - p1.y);
            if (X * X > lngth * lngth / 4) return -1;
            if (Y * Y > wdth * wdth / 4) return -1;
        }
        return h * max_length;
    }
    case RPX_FACE_POLYGON: {
        if constexpr ((F & F_PLANE) == 0) return -1;  // cfaces.pyx:1093-1108
        double max_length = sep(p1, p2);
        double h = (f.p[0] - p1.z) / (p2.z - p1.z);
        if ((h < tol) || (h > 1.0)) return -1;
        double X = p1.x + h * (p2.x - p1.x);
        double Y = p1.y + h * (p2.y - p1.y);
        if (is_base && polygon_face_inside(S.vertices + 2 * f.aux0, f.aux1, X, Y)) return h * max_length;
        return -1;
    }
    case RPX_FACE_SPHERICAL:
    case RPX_FACE_SHAPED_SPHERICAL: {
        if constexpr ((F & F_SPHERE) == 0) return -1;  // cfaces.pyx:421-486 / 513-576
        const bool shaped = f.type == RPX_FACE_SHAPED_SPHERICAL;
        const double curvature = shaped ? f.p[0] : f.p[1];
        const double z_height = shaped ? f.p[1] : f.p[2];
        V3 s = sub(p2, p1);
        double cz = z_height - curvature;
        V3 d = p1;
        d.z -= cz;
        double A = mag_sq(s);
        double B = 2 * dot(s, d);
        double C = mag_sq(d) - curvature * curvature;
        double D = B * B - 4 * A * C;
        if (D < 0) return -1;
        D = sqrt(D);
        double a1 = (-B + D) / (2 * A);
        V3 pt1 = add(p1, mul(s, a1));
        double a2 = (-B - D) / (2 * A);
        V3 pt2 = add(p1, mul(s, a2));
        if (curvature >= 0) {
            if (pt1.z < cz) a1 = RPX_INF;
            if (pt2.z < cz) a2 = RPX_INF;
        } else {
            if (pt1.z > cz) a1 = RPX_INF;
            if (pt2.z > cz) a2 = RPX_INF;
        }
        if (is_base) {
            if (shaped) {
                if (!shape_inside(S, f.shape, pt1.x, pt1.y)) a1 = RPX_INF;
                if (!shape_inside(S, f.shape, pt2.x, pt2.y)) a2 = RPX_INF;
            } else {
                double DD = f.p[0] * f.p[0] / 4.;
                if ((pt1.x * pt1.x + pt1.y * pt1.y) > DD) a1 = RPX_INF;
                if ((pt2.x * pt2.x + pt2.y * pt2.y) > DD) a2 = RPX_INF;
            }
        }
        if (a2 < a1) a1 = a2;
        if (a1 > 1.0 || a1 < tol) return -1;
        return a1 * sep(p1, p2);
    }
    case RPX_FACE_EXTRUDED_PLANAR: {
        if constexpr ((F & F_PLANE) == 0) return -1;  // cfaces.pyx:665-699
        double ux = f.p[0], uy = f.p[1];
        double vx = f.p[2] - ux, vy = f.p[3] - uy;
        V3 s = sub(p2, p1);
        double a;
        if (is_base) {
            a = (s.y * (ux - p1.x) - s.x * (uy - p1.y)) / (s.x * vy - s.y * vx);
            if (a < 0) return -1;
            if (a > 1) return -1;
        }
        a = (vx * (p1.y - uy) - vy * (p1.x - ux)) / (s.x * vy - s.y * vx);
        if (is_base) {
            double dz = a * (p2.z - p1.z);
            double zz = p1.z + dz;
            if (f.p[4] < zz && zz < f.p[5]) return a * mag(s);
            return -1;
        }
        return a * mag(s);
    }
    case RPX_FACE_CONIC: {
        if constexpr ((F & F_CONIC) == 0) return -1;  // cfaces.pyx:1767-1798
        V3 d = sub(p2, p1);
        V3 a = p1;
        a.z -= f.p[1];
        double a1 = intersect_conic(a, d, f.p[0], f.p[2]);
        V3 pt1 = add(a, mul(d, a1));
        if (is_base && !shape_inside(S, f.shape, pt1.x, pt1.y)) return -1;
        if (a1 > 1.0 || a1 < tol) return -1;
        return a1 * sep(p1, p2);
    }
    case RPX_FACE_ASPHERIC: {
        if constexpr ((F & F_ASPHERE) == 0) return -1;  // cfaces.pyx:1909-1976
        double atol = f.p[10];
        double tl = atol * atol;
        if (asph_outside_band(f, p1, p2, is_base)) return -1;
        V3 d = sub(p2, p1);
        V3 a = p1;
        a.z -= f.p[1];
        double a1 = intersect_conic(a, d, f.p[0], f.p[2]);
        Asph A;
        A.R = -f.p[0];
        A.beta = 1 + f.p[2];
        A.A4 = f.p[3]; A.A6 = f.p[4]; A.A8 = f.p[5]; A.A10 = f.p[6];
        A.A12 = f.p[7]; A.A14 = f.p[8]; A.A16 = f.p[9];
        A.R3 = f.p[11];
        A.a = a;
        A.d = d;
        double fv = asph_impf(A, a1), f_last = fv;
        double dz = -fv / asph_grad(A, a1);
        bool converged = false;
        for (int i = 0; i < 100; ++i) {
            a1 += dz;
            if (dz * dz < tl) { converged = true; break; }
            fv = asph_impf(A, a1);
            if (fabs(fv) > fabs(f_last)) return -1;
            f_last = fv;
            dz = -fv / asph_grad(A, a1);
        }
        if (!converged) return -1;
        V3 pt1 = add(a, mul(d, a1));
        if (is_base && !shape_inside(S, f.shape, pt1.x, pt1.y)) return -1;
        if (a1 > 1.0 || a1 < tol) return -1;
        return a1 * sep(p1, p2);
    }
    default:
        return -1;
    }
}

// Local-frame normal of a face (compute_normal_c of each class).
template <uint32_t F>
RPX_DEV V3 plain_normal(const SceneView &S, const rpx_face_t &f, V3 p) {
    switch (f.type) {
    case RPX_FACE_CIRCULAR:  // cfaces.pyx:180-191
        return v3(0, 0, f.invert_normals ? 1 : -1);
    case RPX_FACE_SHAPED_PLANAR:  // cfaces.pyx:228-236
        return v3(0, 0, 1);
    case RPX_FACE_RECTANGULAR:
    case RPX_FACE_POLYGON:  // cfaces.pyx:399-407, 1110-1118
        return v3(0, 0, -1);
    case RPX_FACE_SPHERICAL:
    case RPX_FACE_SHAPED_SPHERICAL: {
        if constexpr ((F & F_SPHERE) == 0) return p;  // cfaces.pyx:488-498
        const bool shaped = f.type == RPX_FACE_SHAPED_SPHERICAL;
        const double curvature = shaped ? f.p[0] : f.p[1];
        const double z_height = shaped ? f.p[1] : f.p[2];
        p.z -= (z_height - curvature);
        if (curvature < 0) {
            p.z = -p.z;
            p.y = -p.y;
            p.x = -p.x;
        }
        return norm(p);
    }
    case RPX_FACE_EXTRUDED_PLANAR:  // cfaces.pyx:701-702
        return v3(f.p[6], f.p[7], f.p[8]);
    case RPX_FACE_CONIC: {
        if constexpr ((F & F_CONIC) == 0) return p;  // cfaces.pyx:1800-1823
        double R = -f.p[0];
        double beta = 1 + f.p[2];
        int sign = f.invert_normals ? -1 : 1;
        p.z -= f.p[1];
        V3 g;
        g.z = 2 * beta * (R - beta * p.z);
        g.x = -p.x * 2 * beta;
        g.y = -p.y * 2 * beta;
        if ((R * beta) < 0) sign *= -1;
        g.z *= sign;
        g.y *= sign;
        g.x *= sign;
        return norm(g);
    }
    case RPX_FACE_ASPHERIC: {
        if constexpr ((F & F_ASPHERE) == 0) return p;  // cfaces.pyx:1978-2007
        double R = -f.p[0];
        double beta = 1 + f.p[2];
        int sign = f.invert_normals ? -1 : 1;
        p.z -= f.p[1];
        double r2 = p.x * p.x + p.y * p.y;
        double root = sqrt(1 - (beta * (r2) / (R * R)));
        double A4 = f.p[3], A6 = f.p[4], A8 = f.p[5], A10 = f.p[6], A12 = f.p[7], A14 = f.p[8], A16 = f.p[9];
        double pw[8];
        pow_chain(r2, pw, 7);
        double df = 10 * A10 * pw[4] + 8 * A8 * pw[3] + 6 * A6 * pw[2] + 4 * A4 * r2;
        df += 16 * A16 * pw[7] + 14 * A14 * pw[6] + 12 * A12 * pw[5];
        df += 2 / (R * (1 + root));
        double r1 = 1 + root;
        df += beta * (r2) / (f.p[11] * root * (r1 * r1));
        V3 g;
        g.z = 1.0;
        g.x = -df * p.x;
        g.y = -df * p.y;
        g.z *= sign;
        g.y *= sign;
        g.x *= sign;
        return norm(g);
    }
    default:
        return p;
    }
}

// DistortionFace.intersect_c (cfaces.pyx:2339-2416)
template <uint32_t F>
RPX_DEV double distortion_intersect(const SceneView &S, const rpx_face_t &f, V3 p1, V3 p2) {
    const rpx_face_t &face = S.faces[f.aux0];
    const int zi = f.aux1;
    double a1, a2, h = sep(p2, p1);
    double tolerance = f.p[0];
    a2 = plain_intersect<F>(S, face, p1, p2, 0);
    if (a2 > h || a2 < f.tolerance) return -1;
    V3 d = sub(p2, p1);
    V3 pt1 = add(p1, mul(d, a2 / h));
    V3 dxdyz = zernike_grad(S, zi, pt1.x, pt1.y);
    V3 n = plain_normal<F>(S, face, pt1);
    pt1.z += dxdyz.z;
    n.x /= n.z;
    n.y /= n.z;
    n.x -= dxdyz.x;
    n.y -= dxdyz.y;
    V3 o = sub(p1, pt1);
    a1 = -h * dot(o, n) / dot(d, n);
    if (a1 < fabs(dxdyz.z)) return -1;
    for (int i = 0; i < 20; ++i) {
        pt1 = add(p1, mul(d, a1 / h));
        if (fabs(a1 - a2) < tolerance) break;
        double z_shift = zernike_sag(S, zi, pt1.x, pt1.y);
        V3 q1 = p1, q2 = p2;
        q1.z -= z_shift;
        q2.z -= z_shift;
        a2 = plain_intersect<F>(S, face, q1, q2, 0);
        pt1 = add(q1, mul(d, a2 / h));
        n = plain_normal<F>(S, face, pt1);
        dxdyz = zernike_grad(S, zi, pt1.x, pt1.y);
        pt1.z += dxdyz.z;
        n.x /= n.z;
        n.y /= n.z;
        n.x -= dxdyz.x;
        n.y -= dxdyz.y;
        o = sub(p1, pt1);
        a2 = a1;
        a1 = -h * dot(o, n) / dot(d, n);
    }
    if (!shape_inside(S, f.shape, pt1.x, pt1.y)) return -1;
    return a1;
}

template <uint32_t F>
RPX_DEV double face_intersect(const SceneView &S, const rpx_face_t &f, V3 p1, V3 p2, int is_base) {
    if constexpr ((F & F_DISTORT) != 0) {
        if (f.type == RPX_FACE_DISTORTION) return distortion_intersect<F>(S, f, p1, p2);
    }
    return plain_intersect<F>(S, f, p1, p2, is_base);
}

// DistortionFace.compute_normal_c (cfaces.pyx:2418-2431)
template <uint32_t F>
RPX_DEV V3 face_normal(const SceneView &S, const rpx_face_t &f, V3 p) {
    if constexpr ((F & F_DISTORT) == 0) {
        return plain_normal<F>(S, f, p);
    }
    if (f.type != RPX_FACE_DISTORTION) return plain_normal<F>(S, f, p);
    V3 dxdyz = zernike_grad(S, f.aux1, p.x, p.y);
    V3 p1 = p;
    p1.z -= dxdyz.z;
    V3 n = plain_normal<F>(S, S.faces[f.aux0], p1);
    n.x /= n.z;
    n.y /= n.z;
    n.z = 1.0;
    n.x -= dxdyz.x;
    n.y -= dxdyz.y;
    return norm(n);
}

// compute_tangent_c: (1,0,0) by default, (0,0,1) for ExtrudedPlanarFace
// (ctracer.pyx:1786-1791, cfaces.pyx:704-709).
RPX_DEV V3 face_tangent(const rpx_face_t &f) {
    if (f.type == RPX_FACE_EXTRUDED_PLANAR) return v3(0.0, 0.0, 1.0);
    return v3(1.0, 0.0, 0.0);
}

struct Orientation {
    V3 normal, tangent;
};

// FaceList.compute_orientation_c (ctracer.pyx:1939-1953)
template <uint32_t F>
RPX_DEV Orientation compute_orientation(const SceneView &S, const rpx_face_list_t &fl,
                                        const rpx_face_t &f, V3 point) {
    point = transform(fl.inv_trans, point);
    Orientation o;
    o.normal = face_normal<F>(S, f, point);
    o.tangent = face_tangent(f);
    if (f.invert_normal) {
        o.normal = neg(o.normal);
        o.tangent = neg(o.tangent);
    }
    o.normal = rotate(fl.trans, o.normal);
    o.tangent = rotate(fl.trans, o.tangent);
    return o;
}

}  // namespace rpx
