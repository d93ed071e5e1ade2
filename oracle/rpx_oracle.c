/*
 * rpx_oracle.c -- CPU restatement of raypier's tracing loop and gausslet field
 * summation, used ONLY as test infrastructure: the parity tests, the smoke check
 * and the CPU-baseline leg of the benchmark call it as the checker.  Nothing in
 * the product library links or calls this file.
 *
 * It follows the reference Cython (raypier/core *.pyx modules) statement by statement,
 * with the same floating-point operation order and the same libm / C99-complex
 * calls, on the packed 188-byte ray_t and 668-byte gausslet_t records.  Mixed
 * real/complex expressions promote the real operand to (x, 0) as the
 * Cython-generated C does.  Build: oracle/Makefile (gcc -O2 -ffp-contract=off).
 *
 * Scene tables use the layouts of include/rpx.h (plain data).  The Zernike
 * distortion here runs the reference's memoised recursion literally, over its
 * NaN-initialised workspace; the compiled instruction lists and polynomial
 * tables of the scene table are ignored.
 *
 * Parity unpinned (no known answer, closed form or invariant checks these;
 * they rest on this restatement alone): the libm boundary (pow, sin, cos,
 * atan2, exp, hypot, csqrt, cexp; the device is compared to 1e-9); the
 * DistortionFace result for rays whose fixed-point iteration has not
 * converged in its 20 steps; the Zernike derivative values for n >= 5 (the
 * reference's slot-sharing quirk).
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/rpx.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef double complex cplx;
#define RE(x) CMPLX((x), 0.0)

/* ------------------------------------------------------------ records */
typedef struct { double x, y, z; } vec_t;

#pragma pack(push, 1)
typedef struct {  /* ctracer.pxd:40-48 */
    vec_t origin, direction, normal, E_vector;
    double n_re, n_im, e1_re, e1_im, e2_re, e2_im;
    double length, phase, accumulated_path;
    uint32_t wavelength_idx, parent_idx, end_face_idx, ray_ident, ray_type_id;
} oray_t;
typedef struct {  /* ctracer.pxd:54-58 */
    vec_t origin, direction, normal;
    double length;
} opara_t;
typedef struct {  /* ctracer.pxd:62-64 */
    oray_t base_ray;
    opara_t para[6];
} ogausslet_t;
#pragma pack(pop)

typedef struct { vec_t normal, tangent; } orient_t;

static const double O_INF = HUGE_VAL;

/* --------------------------------------------------- vector maths (ctracer.pyx:83-262) */
static vec_t V(double x, double y, double z) { vec_t v = {x, y, z}; return v; }
static vec_t transform_c(const rpx_transform_t *t, vec_t p) {
    vec_t o;
    o.x = p.x * t->m[0] + p.y * t->m[1] + p.z * t->m[2] + t->t[0];
    o.y = p.x * t->m[3] + p.y * t->m[4] + p.z * t->m[5] + t->t[1];
    o.z = p.x * t->m[6] + p.y * t->m[7] + p.z * t->m[8] + t->t[2];
    return o;
}
static vec_t rotate_c(const rpx_transform_t *t, vec_t p) {
    vec_t o;
    o.x = p.x * t->m[0] + p.y * t->m[1] + p.z * t->m[2];
    o.y = p.x * t->m[3] + p.y * t->m[4] + p.z * t->m[5];
    o.z = p.x * t->m[6] + p.y * t->m[7] + p.z * t->m[8];
    return o;
}
static double sep_(vec_t p1, vec_t p2) {
    double a = (p2.x - p1.x), b = (p2.y - p1.y), c = (p2.z - p1.z);
    return sqrt((a * a) + (b * b) + (c * c));
}
static vec_t invert_(vec_t v) { return V(-v.x, -v.y, -v.z); }
static vec_t multvs_(vec_t a, double b) { return V(a.x * b, a.y * b, a.z * b); }
static vec_t addvv_(vec_t a, vec_t b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static vec_t subvv_(vec_t a, vec_t b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static double mag_(vec_t a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
static double mag_sq_(vec_t a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
static double dotprod_(vec_t a, vec_t b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static vec_t cross_(vec_t a, vec_t b) {
    vec_t c;
    c.x = a.y * b.z - a.z * b.y;
    c.y = a.z * b.x - a.x * b.z;
    c.z = a.x * b.y - a.y * b.x;
    return c;
}
static vec_t norm_(vec_t a) {
    double mag = sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    a.x /= mag;
    a.y /= mag;
    a.z /= mag;
    return a;
}

/* ------------------------------------------------------------- shapes (cshapes.pyx) */
static int polygon_shape(const double *pts, int size, double X, double Y) {  /* :150-168 */
    int ct = 0;
    double y1 = pts[2 * (size - 1) + 1], x1 = pts[2 * (size - 1)];
    for (int i = 0; i < size; i++) {
        double y2 = pts[2 * i + 1], x2 = pts[2 * i];
        if ((y1 <= Y && Y < y2) || (y2 <= Y && Y < y1)) {
            if ((x1 + (Y - y1) * (x2 - x1) / (y2 - y1)) > X) ct = !ct;
        }
        y1 = y2;
        x1 = x2;
    }
    return ct;
}

/* Recursive evaluation of the postfix program, mirroring the object tree. */
static int shape_eval(const rpx_scene_desc_t *S, int idx, double x, double y, int *consumed_begin) {
    const rpx_shape_op_t *op = &S->shape_ops[idx];
    int b;
    switch (op->op) {
    case RPX_SHAPE_TRUE: *consumed_begin = idx; return 1;            /* ctracer.pyx:1660 */
    case RPX_SHAPE_CIRCLE: {                                          /* cshapes.pyx:109 */
        double dx = x - op->p[0], dy = y - op->p[1];
        *consumed_begin = idx;
        return ((dx * dx) + (dy * dy) < (op->p[2] * op->p[2])) ? 1 : 0;
    }
    case RPX_SHAPE_RECT: {                                            /* cshapes.pyx:128 */
        double dx = x - op->p[0], dy = y - op->p[1];
        *consumed_begin = idx;
        return ((2 * fabs(dx) < op->p[2]) && (2 * fabs(dy)) < op->p[3]) ? 1 : 0;
    }
    case RPX_SHAPE_POLYGON:
        *consumed_begin = idx;
        return polygon_shape(S->vertices + 2 * op->a, op->b, x, y);
    case RPX_SHAPE_NOT: {                                             /* cshapes.pyx:41 */
        int v = shape_eval(S, idx - 1, x, y, &b);
        *consumed_begin = b;
        return 1 & (~v);
    }
    default: {                                                        /* cshapes.pyx:56-67 */
        int rb, lb;
        int r = shape_eval(S, idx - 1, x, y, &rb);
        int l = shape_eval(S, rb - 1, x, y, &lb);
        *consumed_begin = lb;
        if (op->op == RPX_SHAPE_AND) return l & r;
        if (op->op == RPX_SHAPE_OR) return l | r;
        return l ^ r;
    }
    }
}

static int point_inside(const rpx_scene_desc_t *S, int root, double x, double y) {
    int b;
    if (root < 0) return 1;
    return shape_eval(S, root, x, y, &b);
}

/* ---------------------------------------------------------- Zernike (cdistortions.pyx) */
typedef struct {
    int kmax;
    double *ws;  /* 3 x kmax, flat: rows share one contiguous block as in the reference */
} zws_t;

#define WS(w, row, k) ((w)->ws[(row) * (w)->kmax + (k)])

static double zernike_R_c(double r, int k, int n, int m, zws_t *w) {  /* :149-178 */
    int kA, kB, kC, nA, nB, nC, mA, mB, mC, half_n;
    double val;
    if (n < m) return 0.0;
    if (n == 0) return 1.0;
    val = WS(w, 0, k);
    if (isnan(val)) {
        nA = n - 1;
        mA = abs(m - 1);
        half_n = nA / 2;
        kA = half_n * (half_n + 1) + abs(mA);
        nB = nA;
        mB = m + 1;
        kB = half_n * (half_n + 1) + abs(mB);
        nC = n - 2;
        mC = m;
        half_n = nC / 2;
        kC = half_n * (half_n + 1) + abs(mC);
        double ra = zernike_R_c(r, kA, nA, mA, w);
        double rb = zernike_R_c(r, kB, nB, mB, w);
        val = r * (ra + rb);
        val -= zernike_R_c(r, kC, nC, mC, w);
        WS(w, 0, k) = val;
        return val;
    }
    return val;
}

static double zernike_Rprime_c(double r, int k, int n, int m, zws_t *w) {  /* :189-219 */
    int kA, kB, kC, nA, nB, nC, mA, mB, mC, half_n;
    double val;
    if (n < m) return 0.0;
    if (n == 0) return 0.0;
    val = WS(w, 1, k);
    if (isnan(val)) {
        nA = n - 1;
        mA = abs(m - 1);
        half_n = nA / 2;
        kA = half_n * (half_n + 1) + abs(mA);
        nB = nA;
        mB = m + 1;
        kB = half_n * (half_n + 1) + abs(mB);
        nC = n - 2;
        mC = m;
        kC = (half_n - 1) * half_n + abs(mC);   /* as written in the reference */
        double ra = zernike_R_c(r, kA, nA, mA, w);
        double rb = zernike_R_c(r, kB, nB, mB, w);
        val = ra + rb;
        double pa = zernike_Rprime_c(r, kA, nA, mA, w);
        double pb = zernike_Rprime_c(r, kB, nB, mB, w);
        val += r * (pa + pb);
        val -= zernike_Rprime_c(r, kC, nC, mC, w);
        WS(w, 1, k) = val;
        return val;
    }
    return val;
}

static double zernike_R_over_r_c(double r, int k, int n, int m, zws_t *w) {  /* :288-316 */
    int kA, kB, kC, nA, nB, nC, mA, mB, mC, half_n;
    double val;
    if (n < m) return 0.0;
    val = WS(w, 2, k);
    if (isnan(val)) {
        nA = n - 1;
        mA = abs(m - 1);
        half_n = nA / 2;
        kA = half_n * (half_n + 1) + abs(mA);
        nB = nA;
        mB = m + 1;
        kB = half_n * (half_n + 1) + abs(mB);
        nC = n - 2;
        mC = m;
        half_n = nC / 2;
        kC = half_n * (half_n + 1) + abs(mC);
        double ra = zernike_R_c(r, kA, nA, mA, w);
        double rb = zernike_R_c(r, kB, nB, mB, w);
        val = ra + rb;
        val -= zernike_R_over_r_c(r, kC, nC, mC, w);
        WS(w, 2, k) = val;
    }
    return val;
}

static int zk_of(int n, int m) { int h = (int)floor(n / 2.); return h * (h + 1) + abs(m); }

static int z_kmax(const rpx_scene_desc_t *S, const rpx_zernike_t *Z) {
    int n_max = 0;
    for (int i = 0; i < Z->n_coefs; i++) {
        const rpx_zernike_coef_t *c = &S->zernike_coefs[Z->coef_begin + i];
        if (c->n > n_max) n_max = c->n;
    }
    return (n_max / 2) * (n_max / 2 + 1) + n_max + 1;   /* cdistortions.pyx:391 */
}

/* Sized with slack: the reference's Rprime index may step past k_max into the
 * following workspace rows, which belong to the same contiguous block. */
static double z_offset_c(const rpx_scene_desc_t *S, int zi, double x, double y) {  /* :416-450 */
    const rpx_zernike_t *Z = &S->zernikes[zi];
    double buf[3 * 64];
    zws_t w;
    double r, theta, Zs = 0.0, N, PH, R;
    w.kmax = z_kmax(S, Z);
    w.ws = buf;
    x /= Z->unit_radius;
    y /= Z->unit_radius;
    r = sqrt(x * x + y * y);
    theta = atan2(y, x);
    for (int i = 0; i < w.kmax; i++) WS(&w, 0, i) = NAN;
    WS(&w, 0, 0) = 1.0;
    for (int i = 0; i < Z->n_coefs; i++) {
        const rpx_zernike_coef_t *c = &S->zernike_coefs[Z->coef_begin + i];
        if (c->m == 0) N = sqrt(c->n + 1);
        else N = sqrt(2 * (c->n + 1));
        N *= c->value;
        if (c->m >= 0) PH = cos(c->m * theta);
        else PH = -sin(c->m * theta);
        R = zernike_R_c(r, zk_of(c->n, c->m), c->n, abs(c->m), &w);
        Zs += N * R * PH;
    }
    return Zs;
}

static vec_t z_offset_and_gradient_c(const rpx_scene_desc_t *S, int zi, double x, double y) {  /* :456-514 */
    const rpx_zernike_t *Z = &S->zernikes[zi];
    double buf[3 * 64];
    zws_t w;
    double r, theta, N, PH, PHprime, R, Rprime, R_over_r;
    vec_t out;
    w.kmax = z_kmax(S, Z);
    w.ws = buf;
    x /= Z->unit_radius;
    y /= Z->unit_radius;
    r = sqrt(x * x + y * y);
    theta = atan2(y, x);
    for (int i = 0; i < w.kmax; i++) {
        WS(&w, 0, i) = NAN;
        WS(&w, 1, i) = NAN;
        WS(&w, 2, i) = NAN;
    }
    out.x = 0.0;
    out.y = 0.0;
    out.z = 0.0;
    for (int i = 0; i < Z->n_coefs; i++) {
        const rpx_zernike_coef_t *c = &S->zernike_coefs[Z->coef_begin + i];
        if (c->m >= 0) {
            PH = cos(c->m * theta);
            PHprime = -c->m * sin(c->m * theta);
        } else {
            PH = -sin(c->m * theta);
            PHprime = -c->m * cos(c->m * theta);
        }
        int k = zk_of(c->n, c->m);
        R = zernike_R_c(r, k, c->n, abs(c->m), &w);
        Rprime = zernike_Rprime_c(r, k, c->n, abs(c->m), &w);
        R_over_r = zernike_R_over_r_c(r, k, c->n, abs(c->m), &w);
        if (c->m == 0) N = sqrt(c->n + 1);
        else N = sqrt(2 * (c->n + 1));
        N *= c->value;
        out.z += N * R * PH;
        out.x += N * (Rprime * cos(theta) * PH + R_over_r * (-sin(theta)) * PHprime);
        out.y += N * (Rprime * sin(theta) * PH + R_over_r * (cos(theta)) * PHprime);
    }
    out.x /= Z->unit_radius;
    out.y /= Z->unit_radius;
    return out;
}

/* ------------------------------------------------------------------- faces (cfaces.pyx) */
static double intersect_conic(vec_t a, vec_t d, double curvature, double conic_const) {  /* :1695-1747 */
    double beta = 1 + conic_const;
    double R = -curvature;
    double A, B, C, D;
    A = (beta * beta) * (d.z * d.z) + beta * (d.x * d.x) + beta * (d.y * d.y);
    B = -2 * R * beta * d.z + 2 * a.x * beta * d.x + 2 * a.y * beta * d.y + 2 * a.z * (beta * beta) * d.z;
    C = -2 * R * a.z * beta + (a.x * a.x) * beta + (a.y * a.y) * beta + (a.z * a.z) * (beta * beta);
    D = B * B - 4 * A * C;
    if (D < 0) return -1;
    D = sqrt(D);
    if (R * beta * d.z <= 0) return (-B + D) / (2 * A);
    return (-B - D) / (2 * A);
}

typedef struct {
    double R, beta, A4, A6, A8, A10, A12, A14, A16;
    vec_t a, d;
} aspheric_t;

static double eval_aspheric_impf(const aspheric_t *A, double alpha) {  /* :1854-1863 */
    double tx = A->a.x + alpha * A->d.x, ty = A->a.y + alpha * A->d.y;
    double r2 = (tx * tx + ty * ty);
    double out = r2;
    out /= A->R * (1 + sqrt(1 - A->beta * r2 / (A->R * A->R)));
    out -= A->a.z + alpha * A->d.z;
    out += A->A4 * (r2 * r2) + A->A6 * pow(r2, 3) + A->A8 * pow(r2, 4) + A->A10 * pow(r2, 5) +
           A->A12 * pow(r2, 6) + A->A14 * pow(r2, 7) + A->A16 * pow(r2, 8);
    return out;
}

static double eval_aspheric_grad(const aspheric_t *A, double alpha) {  /* :1866-1882 */
    double tx = A->a.x + alpha * A->d.x, ty = A->a.y + alpha * A->d.y;
    double r2 = (tx * tx + ty * ty);
    double dx = A->d.x * tx;
    double dy = A->d.y * ty;
    double out = A->A10 * (10 * dx + 10 * dy) * pow(r2, 4);
    out += A->A12 * (12 * dx + 12 * dy) * pow(r2, 5);
    out += A->A14 * (14 * dx + 14 * dy) * pow(r2, 6);
    out += A->A16 * (16 * dx + 16 * dy) * pow(r2, 7);
    out += A->A4 * (4 * dx + 4 * dy) * (r2);
    out += A->A6 * (6 * dx + 6 * dy) * (r2 * r2);
    out += A->A8 * (8 * dx + 8 * dy) * pow(r2, 3) - A->d.z;
    out += (2 * dx + 2 * dy) / (A->R * (sqrt(1 - A->beta * (r2) / (A->R * A->R)) + 1));
    double q = sqrt(1 - A->beta * r2 / (A->R * A->R)) + 1;
    out += A->beta * (2 * dx + 2 * dy) * (r2) /
           (2 * pow(A->R, 3) * sqrt(1 - A->beta * r2 / (A->R * A->R)) * (q * q));
    return out;
}

static double face_intersect(const rpx_scene_desc_t *S, const rpx_face_t *f, vec_t p1, vec_t p2, int is_base_ray);

static double base_intersect(const rpx_scene_desc_t *S, const rpx_face_t *f, vec_t p1, vec_t p2, int is_base_ray) {
    double tol = f->tolerance;
    switch (f->type) {
    case RPX_FACE_CIRCULAR: {  /* :151-178 */
        double max_length = sep_(p1, p2);
        double h = (f->p[2] - p1.z) / (p2.z - p1.z);
        double X, Y, d = f->p[0];
        if ((h < tol) || (h > 1.0)) return -1;
        X = p1.x + h * (p2.x - p1.x) - f->p[1];
        Y = p1.y + h * (p2.y - p1.y);
        if (is_base_ray && (X * X + Y * Y) > (d * d / 4)) return -1;
        return h * max_length;
    }
    case RPX_FACE_SHAPED_PLANAR: {  /* :201-226 */
        double max_length = sep_(p1, p2);
        double h = (f->p[0] - p1.z) / (p2.z - p1.z);
        double X, Y;
        if ((h < tol) || (h > 1.0)) return -1;
        X = p1.x + h * (p2.x - p1.x);
        Y = p1.y + h * (p2.y - p1.y);
        if (is_base_ray && !point_inside(S, f->shape, X, Y)) return -1;
        return h * max_length;
    }
    case RPX_FACE_RECTANGULAR: {  /* :365-397 */
        double max_length = sep_(p1, p2);
        double h = (f->p[3] - p1.z) / (p2.z - p1.z);
        double X, Y, lngth = f->p[0], wdth = f->p[1];
        if ((h < tol) || (h > 1.0)) return -1;
        if (is_base_ray) {
            X = p1.x + h * (p2.x - p1.x) - f->p[2];
            Y = p1.y + h * (p2.y - p1.y);
            if (X * X > lngth * lngth / 4) return -1;
            if (Y * Y > wdth * wdth / 4) return -1;
        }
        return h * max_length;
    }
    case RPX_FACE_POLYGON: {  /* :1093-1108 with point_in_polygon_c :1050-1068 */
        double max_length = sep_(p1, p2);
        double h = (f->p[0] - p1.z) / (p2.z - p1.z);
        double X, Y;
        if ((h < tol) || (h > 1.0)) return -1;
        X = p1.x + h * (p2.x - p1.x);
        Y = p1.y + h * (p2.y - p1.y);
        if (is_base_ray) {
            const double *pts = S->vertices + 2 * f->aux0;
            int size = f->aux1, ct = 0;
            double y1 = pts[2 * (size - 1) + 1], x1 = pts[2 * (size - 1)];
            for (int i = 0; i < size; i++) {
                double y2 = pts[2 * i + 1], x2 = pts[2 * i];
                double hh = (Y - y1) / (y2 - y1);
                if (0 < hh && hh <= 1.0) {
                    double x = x1 + hh * (x2 - x1);
                    if (x > X) ct = !ct;
                }
                y1 = y2;
                x1 = x2;
            }
            if (ct == 1) return h * max_length;
        }
        return -1;
    }
    case RPX_FACE_SPHERICAL:
    case RPX_FACE_SHAPED_SPHERICAL: {  /* :421-486, :513-576 */
        int shaped = f->type == RPX_FACE_SHAPED_SPHERICAL;
        double curvature = shaped ? f->p[0] : f->p[1];
        double z_height = shaped ? f->p[1] : f->p[2];
        double A, B, C, D, cz, a1, a2;
        vec_t s, d, pt1, pt2;
        s = subvv_(p2, p1);
        cz = z_height - curvature;
        d = p1;
        d.z -= cz;
        A = mag_sq_(s);
        B = 2 * dotprod_(s, d);
        C = mag_sq_(d) - curvature * curvature;
        D = B * B - 4 * A * C;
        if (D < 0) return -1;
        D = sqrt(D);
        a1 = (-B + D) / (2 * A);
        pt1 = addvv_(p1, multvs_(s, a1));
        a2 = (-B - D) / (2 * A);
        pt2 = addvv_(p1, multvs_(s, a2));
        if (curvature >= 0) {
            if (pt1.z < cz) a1 = O_INF;
            if (pt2.z < cz) a2 = O_INF;
        } else {
            if (pt1.z > cz) a1 = O_INF;
            if (pt2.z > cz) a2 = O_INF;
        }
        if (is_base_ray) {
            if (shaped) {
                if (!point_inside(S, f->shape, pt1.x, pt1.y)) a1 = O_INF;
                if (!point_inside(S, f->shape, pt2.x, pt2.y)) a2 = O_INF;
            } else {
                D = f->p[0] * f->p[0] / 4.;
                if ((pt1.x * pt1.x + pt1.y * pt1.y) > D) a1 = O_INF;
                if ((pt2.x * pt2.x + pt2.y * pt2.y) > D) a2 = O_INF;
            }
        }
        if (a2 < a1) a1 = a2;
        if (a1 > 1.0 || a1 < tol) return -1;
        return a1 * sep_(p1, p2);
    }
    case RPX_FACE_EXTRUDED_PLANAR: {  /* :665-699 */
        vec_t s, u, v;
        double a, dz;
        u.x = f->p[0];
        u.y = f->p[1];
        v.x = f->p[2] - u.x;
        v.y = f->p[3] - u.y;
        s = subvv_(p2, p1);
        if (is_base_ray) {
            a = (s.y * (u.x - p1.x) - s.x * (u.y - p1.y)) / (s.x * v.y - s.y * v.x);
            if (a < 0) return -1;
            if (a > 1) return -1;
        }
        a = (v.x * (p1.y - u.y) - v.y * (p1.x - u.x)) / (s.x * v.y - s.y * v.x);
        if (is_base_ray) {
            dz = a * (p2.z - p1.z);
            if (f->p[4] < (p1.z + dz) && (p1.z + dz) < f->p[5]) return a * mag_(s);
        } else {
            return a * mag_(s);
        }
        return -1;
    }
    case RPX_FACE_CONIC: {  /* :1767-1798 */
        vec_t d = subvv_(p2, p1), a = p1, pt1;
        double a1;
        a.z -= f->p[1];
        a1 = intersect_conic(a, d, f->p[0], f->p[2]);
        pt1 = addvv_(a, multvs_(d, a1));
        if (is_base_ray && !point_inside(S, f->shape, pt1.x, pt1.y)) return -1;
        if (a1 > 1.0 || a1 < tol) return -1;
        return a1 * sep_(p1, p2);
    }
    case RPX_FACE_ASPHERIC: {  /* :1909-1976 */
        double a1, dz, fv, f_last;
        double tl = f->p[10] * f->p[10];
        vec_t d = subvv_(p2, p1), a = p1, pt1;
        aspheric_t A;
        int i, converged = 0;
        a.z -= f->p[1];
        a1 = intersect_conic(a, d, f->p[0], f->p[2]);
        A.R = -f->p[0];
        A.beta = 1 + f->p[2];
        A.A4 = f->p[3]; A.A6 = f->p[4]; A.A8 = f->p[5]; A.A10 = f->p[6];
        A.A12 = f->p[7]; A.A14 = f->p[8]; A.A16 = f->p[9];
        A.a = a;
        A.d = d;
        fv = f_last = eval_aspheric_impf(&A, a1);
        dz = -fv / eval_aspheric_grad(&A, a1);
        for (i = 0; i < 100; i++) {
            a1 += dz;
            if (dz * dz < tl) { converged = 1; break; }
            fv = eval_aspheric_impf(&A, a1);
            if (fabs(fv) > fabs(f_last)) return -1;
            f_last = fv;
            dz = -fv / eval_aspheric_grad(&A, a1);
        }
        if (!converged) return -1;
        pt1 = addvv_(a, multvs_(d, a1));
        if (is_base_ray && !point_inside(S, f->shape, pt1.x, pt1.y)) return -1;
        if (a1 > 1.0 || a1 < tol) return -1;
        return a1 * sep_(p1, p2);
    }
    default:
        return -1;
    }
}

static vec_t base_normal(const rpx_scene_desc_t *S, const rpx_face_t *f, vec_t p) {
    (void)S;
    switch (f->type) {
    case RPX_FACE_CIRCULAR: return V(0, 0, f->invert_normals ? 1 : -1);     /* :180-191 */
    case RPX_FACE_SHAPED_PLANAR: return V(0, 0, 1);                         /* :228-236 */
    case RPX_FACE_RECTANGULAR:
    case RPX_FACE_POLYGON: return V(0, 0, -1);                              /* :399, :1110 */
    case RPX_FACE_SPHERICAL:
    case RPX_FACE_SHAPED_SPHERICAL: {                                       /* :488-498 */
        int shaped = f->type == RPX_FACE_SHAPED_SPHERICAL;
        double curvature = shaped ? f->p[0] : f->p[1];
        double z_height = shaped ? f->p[1] : f->p[2];
        p.z -= (z_height - curvature);
        if (curvature < 0) {
            p.z = -p.z;
            p.y = -p.y;
            p.x = -p.x;
        }
        return norm_(p);
    }
    case RPX_FACE_EXTRUDED_PLANAR: return V(f->p[6], f->p[7], f->p[8]);     /* :701 */
    case RPX_FACE_CONIC: {                                                  /* :1800-1823 */
        double R = -f->p[0];
        double beta = 1 + f->p[2];
        int sign = f->invert_normals ? -1 : 1;
        vec_t g;
        p.z -= f->p[1];
        g.z = 2 * beta * (R - beta * p.z);
        g.x = -p.x * 2 * beta;
        g.y = -p.y * 2 * beta;
        if ((R * beta) < 0) sign *= -1;
        g.z *= sign;
        g.y *= sign;
        g.x *= sign;
        return norm_(g);
    }
    case RPX_FACE_ASPHERIC: {                                               /* :1978-2007 */
        double R = -f->p[0];
        double beta = 1 + f->p[2];
        int sign = f->invert_normals ? -1 : 1;
        double r2, root, df;
        vec_t g;
        p.z -= f->p[1];
        r2 = p.x * p.x + p.y * p.y;
        root = sqrt(1 - (beta * (r2) / (R * R)));
        df = 10 * f->p[6] * pow(r2, 4) + 8 * f->p[5] * pow(r2, 3) + 6 * f->p[4] * (r2 * r2) + 4 * f->p[3] * r2;
        df += 16 * f->p[9] * pow(r2, 7) + 14 * f->p[8] * pow(r2, 6) + 12 * f->p[7] * pow(r2, 5);
        df += 2 / (R * (1 + root));
        df += beta * (r2) / (pow(R, 3) * root * ((1 + root) * (1 + root)));
        g.z = 1.0;
        g.x = -df * p.x;
        g.y = -df * p.y;
        g.z *= sign;
        g.y *= sign;
        g.x *= sign;
        return norm_(g);
    }
    default: return p;
    }
}

static double distortion_intersect(const rpx_scene_desc_t *S, const rpx_face_t *self, vec_t p1, vec_t p2) {
    /* DistortionFace.intersect_c (cfaces.pyx:2339-2416) */
    const rpx_face_t *face = &S->faces[self->aux0];
    int zi = self->aux1;
    double z_shift = 0.0, tolerance, a1, a2, h = sep_(p2, p1);
    vec_t dxdyz, pt1, d, n, o, q1, q2;
    tolerance = self->p[0];
    a2 = base_intersect(S, face, p1, p2, 0);
    if (a2 > h || a2 < self->tolerance) return -1;
    d = subvv_(p2, p1);
    pt1 = addvv_(p1, multvs_(d, a2 / h));
    dxdyz = z_offset_and_gradient_c(S, zi, pt1.x, pt1.y);
    n = base_normal(S, face, pt1);
    pt1.z += dxdyz.z;
    n.x /= n.z;
    n.y /= n.z;
    n.x -= dxdyz.x;
    n.y -= dxdyz.y;
    o = subvv_(p1, pt1);
    a1 = -h * dotprod_(o, n) / dotprod_(d, n);
    if (a1 < fabs(dxdyz.z)) return -1;
    for (int i = 0; i < 20; i++) {
        pt1 = addvv_(p1, multvs_(d, a1 / h));
        if (fabs(a1 - a2) < tolerance) break;
        z_shift = z_offset_c(S, zi, pt1.x, pt1.y);
        q1 = p1;
        q2 = p2;
        q1.z -= z_shift;
        q2.z -= z_shift;
        a2 = base_intersect(S, face, q1, q2, 0);
        pt1 = addvv_(q1, multvs_(d, a2 / h));
        n = base_normal(S, face, pt1);
        dxdyz = z_offset_and_gradient_c(S, zi, pt1.x, pt1.y);
        pt1.z += dxdyz.z;
        n.x /= n.z;
        n.y /= n.z;
        n.x -= dxdyz.x;
        n.y -= dxdyz.y;
        o = subvv_(p1, pt1);
        a2 = a1;
        a1 = -h * dotprod_(o, n) / dotprod_(d, n);
    }
    if (!point_inside(S, self->shape, pt1.x, pt1.y)) return -1;
    return a1;
}

static double face_intersect(const rpx_scene_desc_t *S, const rpx_face_t *f, vec_t p1, vec_t p2, int is_base_ray) {
    if (f->type == RPX_FACE_DISTORTION) return distortion_intersect(S, f, p1, p2);
    return base_intersect(S, f, p1, p2, is_base_ray);
}

static vec_t face_normal(const rpx_scene_desc_t *S, const rpx_face_t *f, vec_t p) {
    if (f->type == RPX_FACE_DISTORTION) {  /* cfaces.pyx:2418-2431 */
        vec_t n, p1, dxdyz;
        dxdyz = z_offset_and_gradient_c(S, f->aux1, p.x, p.y);
        p1 = p;
        p1.z -= dxdyz.z;
        n = base_normal(S, &S->faces[f->aux0], p1);
        n.x /= n.z;
        n.y /= n.z;
        n.z = 1.0;
        n.x -= dxdyz.x;
        n.y -= dxdyz.y;
        return norm_(n);
    }
    return base_normal(S, f, p);
}

static vec_t face_tangent(const rpx_face_t *f) {  /* ctracer.pyx:1786, cfaces.pyx:704 */
    if (f->type == RPX_FACE_EXTRUDED_PLANAR) return V(0.0, 0.0, 1.0);
    return V(1.0, 0.0, 0.0);
}

/* FaceList.compute_orientation_c (ctracer.pyx:1939-1953) */
static orient_t compute_orientation(const rpx_scene_desc_t *S, const rpx_face_list_t *fl, const rpx_face_t *f,
                                    vec_t point) {
    orient_t out;
    point = transform_c(&fl->inv_trans, point);
    out.normal = face_normal(S, f, point);
    out.tangent = face_tangent(f);
    if (f->invert_normal) {
        out.normal = invert_(out.normal);
        out.tangent = invert_(out.tangent);
    }
    out.normal = rotate_c(&fl->trans, out.normal);
    out.tangent = rotate_c(&fl->trans, out.tangent);
    return out;
}

/* --------------------------------------------------------------- materials */
static double disp_curve(const rpx_dispersion_t *d, double wavelen) {  /* cmaterials.pyx:107-149 */
    const double *c = d->coefs;
    int n_coefs = d->n_coefs;
    switch (d->formula_id) {
    case 0: return c[0];
    case 1: {
        double n2 = 1.0, wl2 = wavelen * wavelen;
        n2 += c[0];
        for (int i = 0; i < (n_coefs - 1) / 2; i++) n2 += c[2 * i + 1] * wl2 / (wl2 - c[2 * i + 2] * c[2 * i + 2]);
        return sqrt(n2);
    }
    case 2: {
        double n2 = 1.0, wl2 = wavelen * wavelen;
        n2 += c[0];
        for (int i = 0; i < (n_coefs - 1) / 2; i++) n2 += c[2 * i + 1] * wl2 / (wl2 - c[2 * i + 2]);
        return sqrt(n2);
    }
    case 3: {
        double n2 = c[0];
        for (int i = 1; i < n_coefs - 1; i += 2) n2 += c[i] * pow(wavelen, c[i + 1]);
        return sqrt(n2);
    }
    default: return 0.0;  /* sellmeier_5 has no return statement */
    }
}

/* Float32-geometry emulation (measurement of a precision form, not part of
 * the reference): while on, refractive indices, hit points, face
 * orientations and every stored child field except phase and optical path
 * are rounded to float32 where that form would hold them in float32 (the
 * caller rounds the scene's face parameters).  Off by default; set only
 * between traces. */
static int g_f32 = 0;
void oracle_set_f32_geometry(int on) { g_f32 = on; }
static double r32(double x) { return g_f32 ? (double)(float)x : x; }
static vec_t r32v(vec_t v) { return g_f32 ? V((float)v.x, (float)v.y, (float)v.z) : v; }

static cplx disp_eval(const rpx_dispersion_t *d, double wvl) {  /* c_evaluate_n :206-228 */
    double n_imag = 0.0001 * d->absorption / (4 * M_PI);
    return CMPLX(r32(disp_curve(d, wvl)), r32(n_imag * wvl));
}

static oray_t convert_to_sp(oray_t ray, vec_t normal) {  /* cmaterials.pyx:49-91 */
    vec_t E2_vector, E1_vector, v, S_vector, P_vector;
    double A, B, Sr, Si, Pr, Pi;
    E2_vector = norm_(cross_(ray.direction, ray.E_vector));
    E1_vector = norm_(cross_(E2_vector, ray.direction));
    normal = norm_(normal);
    S_vector = cross_(ray.direction, normal);
    if (fabs(S_vector.x) < 1.0e-10 && fabs(S_vector.y) < 1.0e-10 && fabs(S_vector.z) < 1.0e-10) return ray;
    S_vector = norm_(S_vector);
    v = cross_(ray.direction, S_vector);
    P_vector = norm_(v);
    A = dotprod_(E1_vector, S_vector);
    B = dotprod_(E2_vector, S_vector);
    Sr = ray.e1_re * A + ray.e2_re * B;
    Si = ray.e1_im * A + ray.e2_im * B;
    B = dotprod_(E1_vector, P_vector);
    A = dotprod_(E2_vector, P_vector);
    Pr = ray.e1_re * B + ray.e2_re * A;
    Pi = ray.e1_im * B + ray.e2_im * A;
    ray.E_vector = S_vector;
    ray.e1_re = Sr;
    ray.e1_im = Si;
    ray.e2_re = Pr;
    ray.e2_im = Pi;
    return ray;
}

typedef struct {
    oray_t *rays;
    uint64_t n, cap;
    int overflow;
} sink_t;

static void add_ray(sink_t *s, const oray_t *r) {
    if (s->n < s->cap) {
        s->rays[s->n] = *r;
        if (g_f32) {
            oray_t *o = &s->rays[s->n];
            o->origin = r32v(o->origin);
            o->direction = r32v(o->direction);
            o->normal = r32v(o->normal);
            o->E_vector = r32v(o->E_vector);
            o->n_re = r32(o->n_re);
            o->n_im = r32(o->n_im);
            o->e1_re = r32(o->e1_re);
            o->e1_im = r32(o->e1_im);
            o->e2_re = r32(o->e2_re);
            o->e2_im = r32(o->e2_im);
            o->length = r32(o->length);
        }
    } else {
        s->overflow = 1;
    }
    s->n++;
}

static double wl_at(const double *wl, uint64_t n_wl, uint32_t i) { return (i < n_wl) ? wl[i] : NAN; }

static cplx cx_of(const double *p) { return CMPLX(p[0], p[1]); }

static void set_amp(double *re, double *im, cplx v) { *re = creal(v); *im = cimag(v); }

static void eval_child_ray(const rpx_scene_desc_t *S, const rpx_material_t *m, oray_t *in_ray, uint32_t idx,
                           vec_t point, orient_t orient, const double *wls, uint64_t n_wl, sink_t *new_rays) {
    (void)S;
    switch (m->type) {
    case RPX_MAT_TRANSPARENT: {  /* cmaterials.pyx:260-278 */
        vec_t normal = norm_(orient.normal);
        oray_t sp = convert_to_sp(*in_ray, normal);
        sp.accumulated_path += sp.length * sp.n_re;
        sp.origin = point;
        sp.normal = normal;
        sp.length = O_INF;
        sp.parent_idx = idx;
        sp.ray_type_id &= ~RPX_REFL_RAY;
        add_ray(new_rays, &sp);
        break;
    }
    case RPX_MAT_PEC: {  /* :285-319 */
        vec_t normal = norm_(orient.normal);
        oray_t sp = convert_to_sp(*in_ray, normal);
        double cosTheta = dotprod_(normal, in_ray->direction);
        vec_t cosThetaNormal = multvs_(normal, cosTheta);
        vec_t reflected = subvv_(in_ray->direction, multvs_(cosThetaNormal, 2));
        sp.accumulated_path += sp.length * sp.n_re;
        sp.origin = point;
        sp.normal = normal;
        sp.direction = reflected;
        sp.length = O_INF;
        sp.e1_re = -sp.e1_re;
        sp.e1_im = -sp.e1_im;
        sp.parent_idx = idx;
        sp.ray_type_id |= RPX_REFL_RAY;
        add_ray(new_rays, &sp);
        break;
    }
    case RPX_MAT_PARTIAL: {  /* :345-397 */
        vec_t normal = norm_(orient.normal);
        vec_t in_direction = norm_(in_ray->direction);
        oray_t sp = convert_to_sp(*in_ray, normal), sp2;
        double cosTheta = dotprod_(normal, in_direction), R, T;
        vec_t cosThetaNormal = multvs_(normal, cosTheta), reflected;
        sp2 = sp;
        sp.accumulated_path += sp.length * sp.n_re;
        sp2.accumulated_path = sp.accumulated_path;
        R = sqrt(m->reflectivity);
        T = sqrt(1 - m->reflectivity);
        reflected = subvv_(in_direction, multvs_(cosThetaNormal, 2));
        sp.origin = point;
        sp.normal = normal;
        sp.direction = reflected;
        sp.length = O_INF;
        set_amp(&sp.e1_re, &sp.e1_im, CMPLX(sp.e1_re, sp.e1_im) * RE(R));
        set_amp(&sp.e2_re, &sp.e2_im, CMPLX(sp.e2_re, sp.e2_im) * RE(R));
        sp.parent_idx = idx;
        sp.ray_type_id |= RPX_REFL_RAY;
        add_ray(new_rays, &sp);
        sp2.origin = point;
        sp2.normal = normal;
        sp2.direction = in_direction;
        sp2.length = O_INF;
        set_amp(&sp2.e1_re, &sp2.e1_im, CMPLX(sp2.e1_re, sp2.e1_im) * RE(T));
        set_amp(&sp2.e2_re, &sp2.e2_im, CMPLX(sp2.e2_re, sp2.e2_im) * RE(T));
        sp2.parent_idx = idx;
        sp2.ray_type_id &= ~RPX_REFL_RAY;
        add_ray(new_rays, &sp2);
        break;
    }
    case RPX_MAT_DIELECTRIC: {  /* :587-680 */
        vec_t normal = norm_(orient.normal);
        vec_t in_direction = norm_(in_ray->direction);
        oray_t sp = convert_to_sp(*in_ray, normal);
        double cosTheta, n1, n2, N2, cos1, N2_sin2, tan_mag_sq, c2, cos2, Two_n1_cos1, aspect, T_p, T_s;
        int flip;
        vec_t cosThetaNormal, reflected, tangent, tg2, transmitted;
        sp.accumulated_path += sp.length * sp.n_re;
        cosTheta = dotprod_(normal, in_direction);
        cos1 = fabs(cosTheta);
        if (cosTheta < 0.0) {
            n1 = m->n_outside[0];
            n2 = m->n_inside[0];
            sp.n_re = m->n_inside[0];
            sp.n_im = m->n_inside[1];
            flip = 1;
        } else {
            n1 = m->n_inside[0];
            n2 = m->n_outside[0];
            sp.n_re = m->n_outside[0];
            sp.n_im = m->n_outside[1];
            flip = -1;
        }
        N2 = (n2 / n1) * (n2 / n1);
        N2_sin2 = (cosTheta * cosTheta) + (N2 - 1);
        cosThetaNormal = multvs_(normal, cosTheta);
        if (N2_sin2 < 0.0) {
            reflected = subvv_(in_direction, multvs_(cosThetaNormal, 2));
            sp.origin = point;
            sp.normal = normal;
            sp.direction = reflected;
            sp.length = O_INF;
            sp.e1_re *= -1; sp.e1_im *= -1;
            sp.e2_re *= -1; sp.e2_im *= -1;
            sp.parent_idx = idx;
            sp.ray_type_id |= RPX_REFL_RAY;
        } else {
            tangent = subvv_(in_direction, cosThetaNormal);
            tg2 = multvs_(tangent, n1 / n2);
            tan_mag_sq = mag_sq_(tg2);
            c2 = sqrt(1 - tan_mag_sq);
            transmitted = subvv_(tg2, multvs_(normal, c2 * flip));
            cos2 = fabs(dotprod_(transmitted, normal));
            Two_n1_cos1 = (2 * n1) * cos1;
            aspect = sqrt(cos2 / cos1) * Two_n1_cos1;
            T_p = aspect / (n2 * cos1 + n1 * cos2);
            T_s = aspect / (n2 * cos2 + n1 * cos1);
            sp.origin = point;
            sp.normal = normal;
            sp.direction = transmitted;
            sp.length = O_INF;
            sp.e1_re *= T_s; sp.e1_im *= T_s;
            sp.e2_re *= T_p; sp.e2_im *= T_p;
            sp.parent_idx = idx;
            sp.ray_type_id &= ~RPX_REFL_RAY;
        }
        add_ray(new_rays, &sp);
        break;
    }
    case RPX_MAT_FULL_DIELECTRIC:
    case RPX_MAT_FULL_DISPERSIVE: {  /* :755-872, :896-1013 */
        vec_t normal = norm_(orient.normal);
        vec_t in_direction = norm_(in_ray->direction);
        oray_t sp = convert_to_sp(*in_ray, normal);
        cplx n1, n2, cos2, sin2, R_p, R_s, T_p, T_s, E1_amp, E2_amp, n_in, n_out;
        double cosTheta, cos1, sin1, P_in, tan_mag_sq, c2, aspect;
        int flip;
        vec_t cosThetaNormal, reflected, tangent, tg2, transmitted;
        if (m->type == RPX_MAT_FULL_DISPERSIVE) {
            double wl = wl_at(wls, n_wl, in_ray->wavelength_idx);
            n_in = disp_eval(&m->disp_inside, wl);
            n_out = disp_eval(&m->disp_outside, wl);
        } else {
            n_in = cx_of(m->n_inside);
            n_out = cx_of(m->n_outside);
        }
        sp.accumulated_path += sp.length * sp.n_re;
        E1_amp = CMPLX(sp.e1_re, sp.e1_im);
        E2_amp = CMPLX(sp.e2_re, sp.e2_im);
        cosTheta = dotprod_(normal, in_direction);
        cos1 = fabs(cosTheta);
        sin1 = sqrt(fabs(1 - cos1 * cos1));
        if (cosTheta < 0.0) { n1 = n_out; n2 = n_in; flip = 1; }
        else { n1 = n_in; n2 = n_out; flip = -1; }
        sin2 = (n1 * RE(sin1) / n2);
        cos2 = csqrt(RE(1) - sin2 * sin2);
        cosThetaNormal = multvs_(normal, cosTheta);
        P_in = creal(n1) * (creal(E1_amp) * creal(E1_amp) + cimag(E1_amp) * cimag(E1_amp) +
                            creal(E2_amp) * creal(E2_amp) + cimag(E2_amp) * cimag(E2_amp));
        if (P_in == 0.0) return;
        R_p = -(n2 * RE(cos1) - n1 * cos2) / (n2 * RE(cos1) + n1 * cos2);
        R_s = -(n2 * cos2 - n1 * RE(cos1)) / (n2 * cos2 + n1 * RE(cos1));
        R_s *= E1_amp;
        R_p *= E2_amp;
        {
            double a = cabs(R_s), b = cabs(R_p);
            if ((creal(n1) * (a * a + b * b) / P_in) > m->reflection_threshold) {
                reflected = subvv_(in_direction, multvs_(cosThetaNormal, 2));
                sp.origin = point;
                sp.normal = normal;
                sp.direction = reflected;
                sp.length = O_INF;
                set_amp(&sp.e1_re, &sp.e1_im, R_s);
                sp.e2_re = -creal(R_p);
                sp.e2_im = -cimag(R_p);
                sp.parent_idx = idx;
                sp.n_re = creal(n1);
                sp.n_im = cimag(n1);
                sp.ray_type_id |= RPX_REFL_RAY;
                add_ray(new_rays, &sp);
            }
        }
        tangent = subvv_(in_direction, cosThetaNormal);
        tg2 = multvs_(tangent, creal(n1) / creal(n2));
        tan_mag_sq = mag_sq_(tg2);
        c2 = sqrt(1 - tan_mag_sq);
        transmitted = subvv_(tg2, multvs_(normal, c2 * flip));
        aspect = sqrt(creal(cos2) / cos1);
        T_p = RE(aspect) * (RE(2.0 * cos1) * n1) / (n2 * RE(cos1) + n1 * cos2);
        T_s = RE(aspect) * (RE(2.0 * cos1) * n1) / (n2 * cos2 + n1 * RE(cos1));
        T_s *= E1_amp;
        T_p *= E2_amp;
        {
            double a = cabs(T_s), b = cabs(T_p);
            if ((creal(n2) * (a * a + b * b) / P_in) > m->transmission_threshold) {
                sp.origin = point;
                sp.normal = normal;
                sp.direction = transmitted;
                sp.length = O_INF;
                set_amp(&sp.e1_re, &sp.e1_im, T_s);
                set_amp(&sp.e2_re, &sp.e2_im, T_p);
                sp.parent_idx = idx;
                sp.n_re = creal(n2);
                sp.n_im = cimag(n2);
                sp.ray_type_id &= ~RPX_REFL_RAY;
                add_ray(new_rays, &sp);
            }
        }
        break;
    }
    case RPX_MAT_SINGLE_LAYER:
    case RPX_MAT_COATED_DISPERSIVE: {  /* :1026-1182, :1228-1391 */
        double wavelength = wl_at(wls, n_wl, in_ray->wavelength_idx);
        vec_t normal = norm_(orient.normal);
        vec_t in_direction = norm_(in_ray->direction);
        oray_t sp = convert_to_sp(*in_ray, normal);
        cplx n1, n2, n3, cos2, sin2, cos3, sin3, R_p, R_s, T_p, T_s, E1_amp, E2_amp, n_in, n_out, n_coat;
        cplx n1cos1, n2cos2, n3cos3, n1cos2, n2cos1, n2cos3, n3cos2, ep1, ep2, M00, M01, M10, M11, phi;
        double cosTheta, cos1, sin1, P_in, tan_mag_sq, c2, aspect, dwc, thickness;
        int flip;
        vec_t cosThetaNormal, reflected, tangent, tg2, transmitted;
        if (m->type == RPX_MAT_COATED_DISPERSIVE) {
            n_in = disp_eval(&m->disp_inside, wavelength);
            n_out = disp_eval(&m->disp_outside, wavelength);
            n_coat = disp_eval(&m->disp_coating, wavelength);
        } else {
            n_in = cx_of(m->n_inside);
            n_out = cx_of(m->n_outside);
            n_coat = cx_of(m->n_coating);
        }
        thickness = m->thickness;
        sp.accumulated_path += sp.length * sp.n_re;
        E1_amp = CMPLX(sp.e1_re, sp.e1_im);
        E2_amp = CMPLX(sp.e2_re, sp.e2_im);
        cosTheta = dotprod_(normal, in_direction);
        cos1 = fabs(cosTheta);
        sin1 = sqrt(fabs(1 - cos1 * cos1));
        n2 = n_coat;
        if (cosTheta < 0.0) { n1 = n_out; n3 = n_in; flip = 1; }
        else { n1 = n_in; n3 = n_out; flip = -1; }
        sin2 = (n1 * RE(sin1) / n2);
        cos2 = csqrt(RE(1) - sin2 * sin2);
        sin3 = (n1 * RE(sin1) / n3);
        cos3 = csqrt(RE(1) - sin3 * sin3);
        cosThetaNormal = multvs_(normal, cosTheta);
        P_in = creal(n1) * (creal(E1_amp) * creal(E1_amp) + cimag(E1_amp) * cimag(E1_amp) +
                            creal(E2_amp) * creal(E2_amp) + cimag(E2_amp) * cimag(E2_amp));
        if (P_in == 0.0) return;
        n1cos1 = n1 * RE(cos1);
        n2cos2 = n2 * cos2;
        n3cos3 = n3 * cos3;
        dwc = 2 * M_PI * thickness / wavelength;
        phi = CMPLX(-0.0, -1.0) * RE(dwc) * (n2 - sin2 * sin2) / cos2;
        ep1 = cexp(phi) / (RE(4) * n2cos2 * n3cos3);
        ep2 = cexp(RE(-2) * phi);
        M00 = -ep1 * ((n1cos1 - n2cos2) * (n2cos2 + n3cos3) + (n1cos1 + n2cos2) * (n2cos2 - n3cos3) * ep2);
        M01 = ep1 * ((n1cos1 - n2cos2) * (n2cos2 - n3cos3) * ep2 + (n1cos1 + n2cos2) * (n2cos2 + n3cos3));
        M10 = ep1 * ((n1cos1 - n2cos2) * (n2cos2 - n3cos3) + (n1cos1 + n2cos2) * (n2cos2 + n3cos3) * ep2);
        M11 = -ep1 * ((n1cos1 - n2cos2) * (n2cos2 + n3cos3) * ep2 + (n1cos1 + n2cos2) * (n2cos2 - n3cos3));
        R_s = -M00 / M01;
        T_s = M10 + M11 * R_s;
        n1cos2 = n1 * cos2;
        n2cos1 = n2 * RE(cos1);
        n2cos3 = n2 * cos3;
        n3cos2 = n3 * cos2;
        M00 = -ep1 * ((n1cos2 - n2cos1) * (n2cos3 + n3cos2) + (n1cos2 + n2cos1) * (n2cos3 - n3cos2) * ep2);
        M01 = ep1 * ((n1cos2 - n2cos1) * (n2cos3 - n3cos2) * ep2 + (n1cos2 + n2cos1) * (n2cos3 + n3cos2));
        M10 = ep1 * ((n1cos2 - n2cos1) * (n2cos3 - n3cos2) + (n1cos2 + n2cos1) * (n2cos3 + n3cos2) * ep2);
        M11 = -ep1 * ((n1cos2 - n2cos1) * (n2cos3 + n3cos2) * ep2 + (n1cos2 + n2cos1) * (n2cos3 - n3cos2));
        R_p = -M00 / M01;
        T_p = M10 + M11 * R_p;
        R_s *= E1_amp;
        R_p *= E2_amp;
        {
            double a = cabs(R_s), b = cabs(R_p);
            if ((creal(n1) * (a * a + b * b) / P_in) > m->reflection_threshold) {
                reflected = subvv_(in_direction, multvs_(cosThetaNormal, 2));
                sp.origin = point;
                sp.normal = normal;
                sp.direction = reflected;
                sp.length = O_INF;
                set_amp(&sp.e1_re, &sp.e1_im, R_s);
                sp.e2_re = -creal(R_p);
                sp.e2_im = -cimag(R_p);
                sp.parent_idx = idx;
                sp.n_re = creal(n1);
                sp.n_im = cimag(n1);
                sp.ray_type_id |= RPX_REFL_RAY;
                add_ray(new_rays, &sp);
            }
        }
        tangent = subvv_(in_direction, cosThetaNormal);
        tg2 = multvs_(tangent, creal(n1) / creal(n3));
        tan_mag_sq = mag_sq_(tg2);
        c2 = sqrt(1 - tan_mag_sq);
        transmitted = subvv_(tg2, multvs_(normal, c2 * flip));
        aspect = sqrt(creal(cos3) / cos1);
        T_s *= (E1_amp * RE(aspect));
        T_p *= (E2_amp * RE(aspect));
        {
            double a = cabs(T_s), b = cabs(T_p);
            if ((creal(n3) * (a * a + b * b) / P_in) > m->transmission_threshold) {
                sp.origin = point;
                sp.normal = normal;
                sp.direction = transmitted;
                sp.length = O_INF;
                set_amp(&sp.e1_re, &sp.e1_im, T_s);
                set_amp(&sp.e2_re, &sp.e2_im, T_p);
                sp.parent_idx = idx;
                sp.n_re = creal(n3);
                sp.n_im = cimag(n3);
                sp.ray_type_id &= ~RPX_REFL_RAY;
                add_ray(new_rays, &sp);
            }
        }
        break;
    }
    case RPX_MAT_GRATING: {  /* :1472-1542; several orders: each order's child in list order (this
                                build's multi-order grating, n_orders 2 or 3) */
        int n_orders = m->n_orders > 1 ? m->n_orders : 1;
        for (int o = 0; o < n_orders; o++) {
            int order = o == 0 ? m->order : (int)(int8_t)((uint32_t)m->more_orders >> (8 * (o - 1)));
            vec_t normal = norm_(orient.normal);
            vec_t tangent = norm_(orient.tangent);
            vec_t tangent2 = cross_(normal, tangent);
            double wavelen = wl_at(wls, n_wl, in_ray->wavelength_idx);
            double line_spacing = 1000.0 / m->lines_per_mm;
            vec_t reflected = norm_(in_ray->direction);
            double k_z = dotprod_(normal, reflected);
            double k_y = dotprod_(tangent2, reflected);
            double k_x = dotprod_(tangent, reflected);
            int sign = (k_z < 0.0) ? 1 : -1;
            oray_t sp;
            vec_t og = V(m->origin[0], m->origin[1], m->origin[2]);
            k_x = k_x - order * wavelen / (line_spacing * in_ray->n_re);
            k_z = 1 - (k_x * k_x) - (k_y * k_y);
            if (k_z < 0) continue;
            k_z = sign * sqrt(k_z);
            reflected = multvs_(tangent, k_x);
            reflected = addvv_(reflected, multvs_(tangent2, k_y));
            reflected = addvv_(reflected, multvs_(normal, k_z));
            sp = convert_to_sp(*in_ray, normal);
            sp.accumulated_path += sp.length * sp.n_re;
            sp.origin = point;
            sp.normal = normal;
            sp.direction = reflected;
            sp.e1_re = -sp.e1_re * m->efficiency;
            sp.e1_im = -sp.e1_im * m->efficiency;
            sp.e2_re = sp.e2_re * m->efficiency;
            sp.e2_im = sp.e2_im * m->efficiency;
            sp.parent_idx = idx;
            sp.ray_type_id |= RPX_REFL_RAY;
            sp.phase += 1000.0 * dotprod_(subvv_(og, point), tangent) * order * 2 * M_PI / line_spacing;
            add_ray(new_rays, &sp);
        }
        break;
    }
    default:
        break;
    }
}

static opara_t eval_parabasal_ray(const rpx_material_t *m, const oray_t *base_ray, vec_t direction, vec_t point,
                                  orient_t orient, uint32_t ray_type_id, const double *wls, uint64_t n_wl) {
    opara_t out;
    vec_t normal = norm_(orient.normal);
    switch (m->type) {
    case RPX_MAT_DIELECTRIC:
    case RPX_MAT_FULL_DIELECTRIC:
    case RPX_MAT_SINGLE_LAYER:
    case RPX_MAT_COATED_DISPERSIVE: {  /* cmaterials.pyx:683-724, :1393-1434 */
        double cosTheta, n1, n2, nin, nout;
        int flip;
        vec_t cosThetaNormal;
        direction = norm_(direction);
        cosTheta = dotprod_(normal, direction);
        cosThetaNormal = multvs_(normal, cosTheta);
        if (m->type == RPX_MAT_COATED_DISPERSIVE) {
            double wl = wl_at(wls, n_wl, base_ray->wavelength_idx);
            nin = creal(disp_eval(&m->disp_inside, wl));
            nout = creal(disp_eval(&m->disp_outside, wl));
        } else {
            nin = m->n_inside[0];
            nout = m->n_outside[0];
        }
        if (cosTheta < 0.0) { n1 = nout; n2 = nin; flip = 1; }
        else { n1 = nin; n2 = nout; flip = -1; }
        if (ray_type_id & RPX_REFL_RAY) {
            out.direction = subvv_(direction, multvs_(cosThetaNormal, 2));
        } else {
            vec_t tangent = subvv_(direction, cosThetaNormal);
            vec_t tg2 = multvs_(tangent, n1 / n2);
            double tan_mag_sq = mag_sq_(tg2);
            double c2 = sqrt(1 - tan_mag_sq);
            out.direction = subvv_(tg2, multvs_(normal, c2 * flip));
        }
        break;
    }
    case RPX_MAT_GRATING: {  /* :1544-1599 */
        vec_t tangent = norm_(orient.tangent);
        vec_t tangent2 = cross_(normal, tangent);
        double wavelen = wl_at(wls, n_wl, base_ray->wavelength_idx);
        double line_spacing = 1000.0 / m->lines_per_mm;
        vec_t reflected = norm_(direction);
        double k_z = dotprod_(normal, reflected);
        double k_y = dotprod_(tangent2, reflected);
        double k_x = dotprod_(tangent, reflected);
        int sign = (k_z < 0.0) ? 1 : -1;
        k_x = k_x - m->order * wavelen / (line_spacing * base_ray->n_re);
        k_z = 1 - (k_x * k_x) - (k_y * k_y);
        k_z = sign * sqrt(k_z);
        reflected = multvs_(tangent, k_x);
        reflected = addvv_(reflected, multvs_(tangent2, k_y));
        reflected = addvv_(reflected, multvs_(normal, k_z));
        out.direction = reflected;
        break;
    }
    default: {  /* InterfaceMaterial.eval_parabasal_ray_c ctracer.pyx:1588-1610 */
        if (ray_type_id & RPX_REFL_RAY) {
            double cosTheta = dotprod_(normal, direction);
            vec_t cosThetaNormal = multvs_(normal, cosTheta);
            out.direction = subvv_(direction, multvs_(cosThetaNormal, 2));
        } else {
            out.direction = direction;
        }
        break;
    }
    }
    out.origin = point;
    out.normal = normal;
    out.length = O_INF;
    return out;
}

/* ------------------------------------------------------------ trace loops */
static int face_list_of(const rpx_scene_desc_t *S, int rec) {
    for (int l = 0; l < S->n_face_lists; l++) {
        const rpx_face_list_t *fl = &S->face_lists[l];
        if (rec >= fl->first_face && rec < fl->first_face + fl->n_faces) return l;
    }
    return -1;
}

/* FaceList.intersect_c (ctracer.pyx:1882-1904); returns face record or -1 */
static int facelist_intersect(const rpx_scene_desc_t *S, int l, oray_t *ray, vec_t ray_end) {
    const rpx_face_list_t *fl = &S->face_lists[l];
    vec_t p1 = transform_c(&fl->inv_trans, ray->origin);
    vec_t p2 = transform_c(&fl->inv_trans, ray_end);
    int out = -1;
    for (int i = 0; i < fl->n_faces; i++) {
        const rpx_face_t *face = &S->faces[fl->first_face + i];
        double dist = face_intersect(S, face, p1, p2, 1);
        if (face->tolerance < dist && dist < ray->length) {
            ray->length = dist;
            out = fl->first_face + i;
            ray->end_face_idx = (uint32_t)face->global_idx;
        }
    }
    return out;
}

/* select_ray_intersections / select_gausslet_intersections for one
 * collection (ctracer.pyx:1981-2058): a copy of every ray (base ray of every
 * gausslet) whose segment origin -> origin + direction * length meets face
 * list `l` is appended to `out`, with the copy's length and end_face_idx
 * updated by FaceList.intersect_c and wavelength_idx replaced by
 * wl_map[wavelength_idx] (the np.unique re-indexing of the caller). */
uint64_t oracle_select_intersections(const rpx_scene_desc_t *S, int l, const void *in, uint64_t n, int gausslet,
                                     const uint32_t *wl_map, void *out) {
    uint64_t k = 0;
    const size_t rec = gausslet ? sizeof(ogausslet_t) : sizeof(oray_t);
    for (uint64_t i = 0; i < n; i++) {
        unsigned char buf[sizeof(ogausslet_t)];
        memcpy(buf, (const unsigned char *)in + i * rec, rec);
        oray_t *ray = gausslet ? &((ogausslet_t *)buf)->base_ray : (oray_t *)buf;
        vec_t point = addvv_(ray->origin, multvs_(ray->direction, ray->length));
        const int rec_hit = facelist_intersect(S, l, ray, point);
        /* it.face_idx is the idx of the face last accepted (ctracer.pyx:2007-2009) */
        if (rec_hit >= 0 && S->faces[rec_hit].global_idx >= 0) {
            ray->wavelength_idx = wl_map[ray->wavelength_idx];
            memcpy((unsigned char *)out + k * rec, buf, rec);
            k++;
        }
    }
    return k;
}

/* RayCollection._eval_neighbours (ctracer.pyx:1084-1106): rmap[p][t] is the
 * last child of parent p with ray_type_id & REFL_RAY == t; child i's
 * neighbour j is rmap[pnb[p_i][j]][t_i] where pnb[p_i][j] >= 0, else -1.
 * `rmap` is caller scratch of 2 * n_parent ints.  Returns 0, or -1 for an
 * index outside the parent generation (the reference raises IndexError). */
int oracle_eval_neighbours(const uint32_t *parent_idx, const uint32_t *ray_type_id, uint64_t n_child,
                           const int32_t *pnb, uint64_t n_parent, int k, int32_t *rmap, int32_t *nb) {
    for (uint64_t q = 0; q < 2 * n_parent; q++) rmap[q] = -1;
    for (uint64_t q = 0; q < n_child * (uint64_t)k; q++) nb[q] = -1;
    for (uint64_t i = 0; i < n_child; i++) {
        if (parent_idx[i] >= n_parent) return -1;
        rmap[2 * (uint64_t)parent_idx[i] + (ray_type_id[i] & 1u)] = (int32_t)i;
    }
    for (uint64_t i = 0; i < n_child; i++) {
        const uint64_t pidx = parent_idx[i];
        const uint32_t rtype = ray_type_id[i] & 1u;
        for (int j = 0; j < k; j++) {
            const int32_t child_nb = pnb[pidx * (uint64_t)k + j];
            if (child_nb >= 0) {
                if ((uint64_t)child_nb >= n_parent) return -1;
                nb[i * (uint64_t)k + j] = rmap[2 * (uint64_t)child_nb + rtype];
            }
        }
    }
    return 0;
}

/* GaussletCollection.config_parabasal_rays (ctracer.pyx:1430-1481). */
int oracle_config_parabasal_rays(ogausslet_t *gs, uint64_t n, const double *wavelength_list, uint64_t n_wl,
                                 double radius, double working_dist) {
    for (uint64_t i = 0; i < n; i++) {
        ogausslet_t *gc = gs + i;
        vec_t base_d = norm_(gc->base_ray.direction);
        vec_t o;
        if (base_d.x > base_d.y) o = V(0.0, 1.0, 0.0);
        else o = V(1.0, 0.0, 0.0);
        vec_t d1 = norm_(cross_(base_d, o));
        vec_t d2 = norm_(cross_(base_d, d1));
        if (gc->base_ray.wavelength_idx >= n_wl) return -1;
        double theta0 = wavelength_list[gc->base_ray.wavelength_idx] / (M_PI * radius * 1000.0);
        for (int j = 0; j < 6; j += 2) {
            double angle = (j * 2 * M_PI / 6);
            o = addvv_(multvs_(d1, radius * cos(angle)), multvs_(d2, radius * sin(angle)));
            o = addvv_(o, gc->base_ray.origin);
            o = addvv_(o, multvs_(base_d, working_dist));
            vec_t d = addvv_(multvs_(d1, -theta0 * sin(angle)), multvs_(d2, theta0 * cos(angle)));
            vec_t da = addvv_(base_d, d);
            vec_t db = subvv_(base_d, d);
            gc->para[j].direction = norm_(da);
            gc->para[j].origin = subvv_(o, multvs_(da, working_dist));
            gc->para[j + 1].direction = norm_(db);
            gc->para[j + 1].origin = subvv_(o, multvs_(db, working_dist));
            gc->para[j].normal = gc->base_ray.normal;
            gc->para[j + 1].normal = gc->base_ray.normal;
            gc->para[j].length = gc->base_ray.length;
            gc->para[j + 1].length = gc->base_ray.length;
        }
    }
    return 0;
}

/* GaussletCollection.project_to_plane (ctracer.pyx:1391-1420). */
void oracle_project_to_plane(ogausslet_t *gs, uint64_t n, const double *origin, const double *direction) {
    vec_t o = V(origin[0], origin[1], origin[2]);
    vec_t d = norm_(V(direction[0], direction[1], direction[2]));
    for (uint64_t i = 0; i < n; i++) {
        ogausslet_t *gc = gs + i;
        oray_t *ray = &gc->base_ray;
        double a = dotprod_(subvv_(o, ray->origin), d) / dotprod_(ray->direction, d);
        ray->origin = addvv_(ray->origin, multvs_(ray->direction, a));
        ray->accumulated_path += ray->n_re * a;
        for (int j = 0; j < 6; j++) {
            opara_t *para = &gc->para[j];  /* ctracer.pyx:1416-1419 */
            a = dotprod_(subvv_(o, para->origin), d) / dotprod_(para->direction, d);
            para->origin = addvv_(para->origin, multvs_(para->direction, a));
        }
    }
}

/* GaussletCollection.lagrange_invariant (ctracer.pyx:1347-1389).  Returns -1
 * if a wavelength_idx is outside the table. */
int oracle_lagrange_invariant(const ogausslet_t *gs, uint64_t n, const double *wavelen, uint64_t n_wl, double *out) {
    for (uint64_t k = 0; k < n; k++) {
        const ogausslet_t *g = gs + k;
        const oray_t *r = &g->base_ray;
        vec_t h[6], u[6];
        vec_t axis1 = norm_(r->E_vector);
        vec_t axis2 = cross_(axis1, r->direction);
        for (int i = 0; i < 6; i++) {
            const opara_t *p = &g->para[i];
            vec_t o = subvv_(p->origin, r->origin);
            h[i].x = dotprod_(o, axis1);
            h[i].y = dotprod_(o, axis2);
            h[i].z = 0.0;
            u[i].x = dotprod_(p->direction, axis1);
            u[i].y = dotprod_(p->direction, axis2);
            u[i].z = 0.0;
        }
        double v = 0, t;
        t = dotprod_(h[0], u[5]) - dotprod_(h[5], u[0]); v += t * t;
        t = dotprod_(h[1], u[2]) - dotprod_(h[2], u[1]); v += t * t;
        t = dotprod_(h[3], u[4]) - dotprod_(h[4], u[3]); v += t * t;
        t = dotprod_(h[1], u[4]) - dotprod_(h[4], u[1]); v += t * t;
        t = dotprod_(h[0], u[3]) - dotprod_(h[3], u[0]); v += t * t;
        t = dotprod_(h[2], u[5]) - dotprod_(h[5], u[2]); v += t * t;
        if (r->wavelength_idx >= n_wl) return -1;
        out[k] = 1000 * sqrt(v / 6) / wavelen[r->wavelength_idx];
    }
    return 0;
}

/* trace_segment_c (ctracer.pyx:2062-2118) / trace_one_face_segment_c (:2121-2170).
 * one_list < 0 selects the non-sequential form.  Returns the number of children
 * (written only up to `cap`). */
static void trace_segment_range(const rpx_scene_desc_t *S, int one_list, int one_face, oray_t *rays, uint64_t i0,
                                uint64_t i1, float max_length, const double *wls, uint64_t n_wl, sink_t *sink,
                                uint32_t *face_hits) {
    for (uint64_t i = i0; i < i1; i++) {
        oray_t *ray = rays + i;
        int nearest = -1, nearest_list = -1;
        vec_t point;
        ray->length = max_length;
        ray->end_face_idx = RPX_NO_FACE;
        point = addvv_(ray->origin, multvs_(ray->direction, max_length));
        if (one_list < 0) {
            for (int j = 0; j < S->n_face_lists; j++) {
                int it = facelist_intersect(S, j, ray, point);
                if (it >= 0) {
                    nearest = it;
                    nearest_list = j;
                }
            }
        } else {  /* FaceList.intersect_one_face_c (ctracer.pyx:1861-1879) */
            const rpx_face_list_t *fl = &S->face_lists[one_list];
            vec_t p1 = transform_c(&fl->inv_trans, ray->origin);
            vec_t p2 = transform_c(&fl->inv_trans, point);
            const rpx_face_t *face = &S->faces[fl->first_face + one_face];
            double dist = face_intersect(S, face, p1, p2, 1);
            if (face->tolerance < dist && dist < ray->length) {
                ray->length = dist;
                ray->end_face_idx = (uint32_t)face->global_idx;
                nearest = fl->first_face + one_face;
                nearest_list = one_list;
            }
        }
        if (nearest >= 0) {
            const rpx_face_t *face = &S->faces[nearest];
            orient_t orient;
            if (face_hits) face_hits[face->global_idx] += 1;
            point = r32v(addvv_(ray->origin, multvs_(ray->direction, ray->length)));
            orient = compute_orientation(S, &S->face_lists[nearest_list], face, point);
            orient.normal = r32v(orient.normal);
            orient.tangent = r32v(orient.tangent);
            if (face->material >= 0)
                eval_child_ray(S, &S->materials[face->material], ray, (uint32_t)i, point, orient, wls, n_wl, sink);
        }
    }
}

uint64_t oracle_trace_segment(const rpx_scene_desc_t *S, int one_list, int one_face, oray_t *rays, uint64_t n,
                              float max_length, const double *wls, uint64_t n_wl, oray_t *out, uint64_t cap,
                              uint32_t *face_hits) {
    sink_t sink = {out, 0, cap, 0};
    trace_segment_range(S, one_list, one_face, rays, 0, n, max_length, wls, n_wl, &sink, face_hits);
    return sink.n;
}

/* The same generation over `threads` OpenMP threads: each traces a
 * contiguous block of rays into its own part of `out` (max_children slots
 * per ray, so cap >= max_children * n), then the parts are moved down in
 * block order.  Output and hit counts equal oracle_trace_segment's.
 * Returns UINT64_MAX if cap is too small or a block overflowed. */
uint64_t oracle_trace_segment_mt(const rpx_scene_desc_t *S, int one_list, int one_face, oray_t *rays, uint64_t n,
                                 float max_length, const double *wls, uint64_t n_wl, oray_t *out, uint64_t cap,
                                 uint32_t *face_hits, uint64_t n_hits, int max_children, int threads) {
    if (threads < 1) threads = 1;
    if ((uint64_t)max_children * n > cap) return UINT64_MAX;
    uint64_t counts[256] = {0}, starts[257] = {0};
    int bad = 0;
    if (threads > 256) threads = 256;
    uint32_t *hits = calloc((size_t)threads * (n_hits ? n_hits : 1), sizeof(uint32_t));
    if (!hits) return UINT64_MAX;
    for (int t = 0; t <= threads; t++) starts[t] = n * (uint64_t)t / (uint64_t)threads;
#pragma omp parallel for num_threads(threads) schedule(static, 1) reduction(| : bad)
    for (int t = 0; t < threads; t++) {
        sink_t sink = {out + (uint64_t)max_children * starts[t], 0,
                       (uint64_t)max_children * (starts[t + 1] - starts[t]), 0};
        trace_segment_range(S, one_list, one_face, rays, starts[t], starts[t + 1], max_length, wls, n_wl, &sink,
                            hits + (size_t)t * n_hits);
        counts[t] = sink.n;
        bad |= sink.n > sink.cap;
    }
    uint64_t total = 0;
    if (!bad) {
        for (int t = 0; t < threads; t++) {
            memmove(out + total, out + (uint64_t)max_children * starts[t], counts[t] * sizeof(oray_t));
            total += counts[t];
        }
        for (int t = 0; t < threads; t++)
            for (uint64_t k = 0; k < n_hits; k++) face_hits[k] += hits[(size_t)t * n_hits + k];
    }
    free(hits);
    return bad ? UINT64_MAX : total;
}

/* trace_gausslet_c (ctracer.pyx:2214-2281) with trace_parabasal_rays (:2350-2385);
 * one_list >= 0 gives trace_one_face_gausslet_c (:2284-2347). */
uint64_t oracle_trace_gausslet(const rpx_scene_desc_t *S, int one_list, int one_face, ogausslet_t *gs, uint64_t n,
                               double max_length, const double *wls, uint64_t n_wl, ogausslet_t *out, uint64_t cap,
                               uint32_t *face_hits) {
    uint64_t n_out = 0;
    oray_t child_buf[4];
    for (uint64_t i = 0; i < n; i++) {
        ogausslet_t *g = gs + i;
        oray_t *ray = &g->base_ray;
        int nearest = -1, nearest_list = -1;
        vec_t point;
        ray->end_face_idx = RPX_NO_FACE;
        point = addvv_(ray->origin, multvs_(ray->direction, max_length));
        if (one_list < 0) {
            for (int j = 0; j < S->n_face_lists; j++) {
                int it = facelist_intersect(S, j, ray, point);
                if (it >= 0) {
                    nearest = it;
                    nearest_list = j;
                }
            }
        } else {
            const rpx_face_list_t *fl = &S->face_lists[one_list];
            vec_t p1 = transform_c(&fl->inv_trans, ray->origin);
            vec_t p2 = transform_c(&fl->inv_trans, point);
            const rpx_face_t *face = &S->faces[fl->first_face + one_face];
            double dist = face_intersect(S, face, p1, p2, 1);
            if (face->tolerance < dist && dist < ray->length) {
                ray->length = dist;
                ray->end_face_idx = (uint32_t)face->global_idx;
                nearest = fl->first_face + one_face;
                nearest_list = one_list;
            }
        }
        if (nearest < 0) continue;
        {
            const rpx_face_t *face = &S->faces[nearest];
            const rpx_face_list_t *fl = &S->face_lists[nearest_list];
            sink_t sink = {child_buf, 0, 4, 0};
            vec_t ppoint[6];
            orient_t porient[6];
            orient_t orient;
            int ok = 1;
            if (face_hits) face_hits[face->global_idx] += 1;
            point = addvv_(ray->origin, multvs_(ray->direction, ray->length));
            orient = compute_orientation(S, fl, face, point);
            if (face->material >= 0)
                eval_child_ray(S, &S->materials[face->material], ray, (uint32_t)i, point, orient, wls, n_wl, &sink);
            for (int j = 0; j < 6; j++) {  /* FaceList.intersect_para_c (ctracer.pyx:1915-1928) */
                opara_t *p = &g->para[j];
                vec_t ray_end = addvv_(p->origin, multvs_(p->direction, max_length));
                vec_t p1 = transform_c(&fl->inv_trans, p->origin);
                vec_t p2 = transform_c(&fl->inv_trans, ray_end);
                double dist = face_intersect(S, face, p1, p2, 0);
                if (face->tolerance < dist && dist < p->length) {
                    p->length = dist;
                } else {
                    ok = 0;
                    break;
                }
                ppoint[j] = addvv_(p->origin, multvs_(p->direction, p->length));
                porient[j] = compute_orientation(S, fl, face, ppoint[j]);
            }
            if (!ok) continue;
            for (uint64_t k = 0; k < sink.n && k < 4; k++) {
                ogausslet_t ng;
                ng.base_ray = child_buf[k];
                for (int j = 0; j < 6; j++)
                    ng.para[j] = eval_parabasal_ray(&S->materials[face->material], &child_buf[k], g->para[j].direction,
                                                    ppoint[j], porient[j], ng.base_ray.ray_type_id, wls, n_wl);
                /* new_gausslets.reset_length_c(max_length) (ctracer.pyx:2280) */
                ng.base_ray.length = max_length;
                for (int j = 0; j < 6; j++) ng.para[j].length = max_length;
                if (n_out < cap) out[n_out] = ng;
                n_out++;
            }
        }
    }
    return n_out;
}

/* ------------------------------------------------- unit-level entry points */
double oracle_face_intersect(const rpx_scene_desc_t *S, int rec, const double *p1, const double *p2, int is_base_ray) {
    return face_intersect(S, &S->faces[rec], V(p1[0], p1[1], p1[2]), V(p2[0], p2[1], p2[2]), is_base_ray);
}

void oracle_face_normal(const rpx_scene_desc_t *S, int rec, const double *p, double *out) {
    vec_t n = face_normal(S, &S->faces[rec], V(p[0], p[1], p[2]));
    out[0] = n.x;
    out[1] = n.y;
    out[2] = n.z;
}

int oracle_shape_inside(const rpx_scene_desc_t *S, int root, double x, double y) { return point_inside(S, root, x, y); }

void oracle_zernike(const rpx_scene_desc_t *S, int zi, const double *x, const double *y, uint64_t n, double *sag,
                    double *grad) {
    for (uint64_t i = 0; i < n; i++) {
        if (sag) sag[i] = z_offset_c(S, zi, x[i], y[i]);
        if (grad) {
            vec_t g = z_offset_and_gradient_c(S, zi, x[i], y[i]);
            grad[3 * i] = g.x;
            grad[3 * i + 1] = g.y;
            grad[3 * i + 2] = g.z;
        }
    }
}

/* InterfaceMaterial.eval_child_ray (ctracer.pyx:1620-1632) */
uint64_t oracle_eval_child_ray(const rpx_scene_desc_t *S, int mat, oray_t *in_ray, uint32_t idx, const double *point,
                               const double *normal, const double *tangent, const double *wls, uint64_t n_wl,
                               oray_t *out, uint64_t cap) {
    sink_t sink = {out, 0, cap, 0};
    orient_t o;
    o.normal = V(normal[0], normal[1], normal[2]);
    o.tangent = V(tangent[0], tangent[1], tangent[2]);
    eval_child_ray(S, &S->materials[mat], in_ray, idx, V(point[0], point[1], point[2]), o, wls, n_wl, &sink);
    return sink.n;
}

void oracle_dispersion(const rpx_dispersion_t *d, const double *wl, uint64_t n, double *out_complex) {
    for (uint64_t i = 0; i < n; i++) {
        cplx v = disp_eval(d, wl[i]);
        out_complex[2 * i] = creal(v);
        out_complex[2 * i + 1] = cimag(v);
    }
}

/* ------------------------------------------------------------- fields */
/* evaluate_one_mode (cfields.pyx:156-215) over rows of x, y, dx, dy. */
void oracle_evaluate_modes(const double *X, const double *Y, const double *DX, const double *DY, uint64_t n, int size,
                           double blending, double *out) {
    for (uint64_t r = 0; r < n; r++) {
        const double *x = X + r * size, *y = Y + r * size, *dx = DX + r * size, *dy = DY + r * size;
        double *o = out + 6 * r;
        double a00 = 0, a01 = 0, a02 = 0, a11 = 0, a12 = 0, a22 = 0, b0 = 0, b1 = 0, b2 = 0;
        double xi2, yi2, den;
        for (int i = 0; i < size; i++) {
            xi2 = x[i] * x[i];
            yi2 = y[i] * y[i];
            a00 += xi2 * xi2;
            a01 += 2 * xi2 * x[i] * y[i];
            a02 += xi2 * yi2;
            a11 += 4 * xi2 * yi2;
            a12 += 2 * x[i] * y[i] * yi2;
            a22 += yi2 * yi2;
            b0 += xi2;
            b1 += 2 * x[i] * y[i];
            b2 += yi2;
        }
        den = (-a00 * a11 * a22 + a00 * (a12 * a12) + (a01 * a01) * a22 - 2 * a01 * a02 * a12 + (a02 * a02) * a11);
        o[1] = blending * (-b0 * (a11 * a22 - (a12 * a12)) + b1 * (a01 * a22 - a02 * a12) - b2 * (a01 * a12 - a02 * a11)) / den;
        o[3] = -blending * (b0 * (a01 * a22 - a02 * a12) - b1 * (a00 * a22 - (a02 * a02)) + b2 * (a00 * a12 - a01 * a02)) / den;
        o[5] = blending * (-b0 * (a01 * a12 - a02 * a11) + b1 * (a00 * a12 - a01 * a02) - b2 * (a00 * a11 - (a01 * a01))) / den;
        a00 = a01 = a02 = a11 = a12 = a22 = b0 = b1 = b2 = 0;
        for (int i = 0; i < size; i++) {
            xi2 = x[i] * x[i];
            yi2 = y[i] * y[i];
            a00 += xi2;
            a01 += x[i] * y[i];
            a11 += xi2 + yi2;
            a22 += yi2;
            b0 += dx[i] * x[i];
            b1 += dx[i] * y[i] + dy[i] * x[i];
            b2 += dy[i] * y[i];
        }
        a12 = a01 * a01;
        den = (a00 * a12 - a00 * a11 * a22 + a12 * a22);
        o[0] = (-a12 * b2 + a01 * a22 * b1 + b0 * (a12 - a11 * a22)) / den;
        o[2] = -(a00 * a01 * b2 - a00 * a22 * b1 + a01 * a22 * b0) / den;
        o[4] = (a00 * a01 * b1 - a12 * b0 - b2 * (a00 * a11 - a12)) / den;
    }
}

/* evaluate_neighbours_gc (core/fields.py:114-136) on gausslet_t records:
 * x, y, dx, dy of shape [n][6]. */
void oracle_neighbours_gc(const ogausslet_t *g, uint64_t n, double *X, double *Y, double *DX, double *DY) {
    for (uint64_t i = 0; i < n; i++) {
        const oray_t *b = &g[i].base_ray;
        vec_t E = b->E_vector, d = b->direction;
        vec_t H = cross_(E, d);   /* numpy.cross(E, direction) */
        for (int j = 0; j < 6; j++) {
            const opara_t *p = &g[i].para[j];
            vec_t off = subvv_(p->origin, b->origin);
            vec_t nd = p->direction;
            double dz = (nd.x * d.x + nd.y * d.y) + nd.z * d.z;
            X[6 * i + j] = (off.x * E.x + off.y * E.y) + off.z * E.z;
            Y[6 * i + j] = (off.x * H.x + off.y * H.y) + off.z * H.z;
            DX[6 * i + j] = ((nd.x * E.x + nd.y * E.y) + nd.z * E.z) / dz;
            DY[6 * i + j] = ((nd.x * H.x + nd.y * H.y) + nd.z * H.z) / dz;
        }
    }
}

/* sum_gaussian_modes (cfields.pyx:52-115) with calc_mode_U (:118-151).
 * modes: complex128[n][3] interleaved; out: complex128[npt][3], overwritten. */
void oracle_sum_gaussian_modes(const oray_t *rays, uint64_t nray, const double *modes, const double *wls,
                               uint64_t n_wl, const double *points, uint64_t npt, double time_ps, double *out) {
    double c = 0.299792458;
    const cplx rootI = csqrt(CMPLX(0.0, 1.0));
    c *= time_ps;
    memset(out, 0, sizeof(double) * 6 * npt);
    for (uint64_t iray = 0; iray < nray; iray++) {
        oray_t ray = rays[iray];
        vec_t E = norm_(ray.E_vector);
        vec_t H = norm_(cross_(ray.direction, E));
        double k = 2000.0 * M_PI / wl_at(wls, n_wl, ray.wavelength_idx);
        double phase = ray.phase + (ray.accumulated_path * k) - (c * k / ray.n_re);
        cplx kz = CMPLX(ray.n_re, ray.n_im);
        cplx A = CMPLX(modes[6 * iray], modes[6 * iray + 1]);
        cplx B = CMPLX(modes[6 * iray + 2], modes[6 * iray + 3]);
        cplx C = CMPLX(modes[6 * iray + 4], modes[6 * iray + 5]);
        double invk, inv_root_area;
        cplx detG0;
        kz *= RE(k);
        invk = 2. / creal(kz);
        inv_root_area = sqrt(sqrt(cimag(A) * cimag(C) - (cimag(B) * cimag(B))) * (2.0 / M_PI));
        A = CMPLX(creal(A), cimag(A) * invk);
        B = CMPLX(creal(B), cimag(B) * invk);
        C = CMPLX(creal(C), cimag(C) * invk);
        detG0 = (A * C) - (B * B);
        for (uint64_t ipt = 0; ipt < npt; ipt++) {
            vec_t pt = V(points[3 * ipt], points[3 * ipt + 1], points[3 * ipt + 2]);
            double x, y, z;
            cplx denom, AA, CC, U, E1, E2;
            pt = subvv_(pt, ray.origin);
            x = dotprod_(pt, E);
            y = dotprod_(pt, H);
            z = dotprod_(pt, ray.direction);
            denom = (RE(1) + (RE(z) * (A + C)) + RE(z * z) * detG0) * RE(2);
            AA = (A + RE(z) * detG0) / denom;
            CC = (C + RE(z) * detG0) / denom;
            U = cexp(CMPLX(0.0, 1.0) * (RE(phase) + kz * (RE(z) + AA * RE(x * x) + B * RE(2 * x * y) / denom + CC * RE(y * y))));
            U /= csqrt(((RE(1) + RE(z) * A) * (RE(1) + RE(z) * C) - (RE(z) * B) * (RE(z) * B)) * CMPLX(0.0, 1.0));
            U *= RE(inv_root_area) * rootI;
            E1 = CMPLX(ray.e1_re, ray.e1_im) * U;
            E2 = CMPLX(ray.e2_re, ray.e2_im) * U;
            {
                cplx o0 = E1 * RE(E.x) + E2 * RE(H.x);
                cplx o1 = E1 * RE(E.y) + E2 * RE(H.y);
                cplx o2 = E1 * RE(E.z) + E2 * RE(H.z);
                out[6 * ipt + 0] += creal(o0);
                out[6 * ipt + 1] += cimag(o0);
                out[6 * ipt + 2] += creal(o1);
                out[6 * ipt + 3] += cimag(o1);
                out[6 * ipt + 4] += creal(o2);
                out[6 * ipt + 5] += cimag(o2);
            }
        }
    }
}

/* calc_mode_curvature (cfields.pyx:231-293). */
void oracle_calc_mode_curvature(const double *dx, const double *dy, const double *dx2, const double *dy2,
                                const double *dxdy, uint64_t N, double *A, double *B, double *C, double *x,
                                double *y, double *z) {
    vec_t a, b, c;
    for (uint64_t i = 0; i < N; i++) {
        double Fz;
        c.x = -dx[i];
        c.y = -dy[i];
        c.z = 1.0;
        c = norm_(c);
        b.x = 0.0;
        b.y = 0.0;
        b.z = 1.0;
        a = norm_(cross_(b, c));
        b = cross_(c, a);
        x[3 * i] = a.x; x[3 * i + 1] = a.y; x[3 * i + 2] = a.z;
        y[3 * i] = b.x; y[3 * i + 1] = b.y; y[3 * i + 2] = b.z;
        z[3 * i] = c.x; z[3 * i + 1] = c.y; z[3 * i + 2] = c.z;
        Fz = c.x * dx[i] + c.y * dy[i] + c.z;
        A[i] = -(a.x * a.x * dx2[i] + 2 * a.x * a.y * dxdy[i] + a.y * a.y * dy2[i]) / Fz;
        B[i] = -(a.x * b.x * dx2[i] + (a.x * b.y + b.x * a.y) * dxdy[i] + a.y * b.y * dy2[i]) / Fz;
        C[i] = -(b.x * b.x * dx2[i] + 2 * b.x * b.y * dxdy[i] + b.y * b.y * dy2[i]) / Fz;
    }
}

/* apply_mode_curvature (cfields.pyx:397-424). */
void oracle_apply_mode_curvature(ogausslet_t *gs, uint64_t n, const double *A, const double *B, const double *C) {
    for (uint64_t i = 0; i < n; i++) {
        ogausslet_t *g = gs + i;
        oray_t *ray = &g->base_ray;
        vec_t z = ray->direction, x = ray->E_vector, y = cross_(x, z);
        for (int j = 0; j < 6; j++) {
            opara_t *para = &g->para[j];
            vec_t r = subvv_(para->origin, ray->origin), shift;
            double px = dotprod_(r, x), py = dotprod_(r, y);
            shift = multvs_(x, (-A[i] * px - B[i] * py));
            shift = addvv_(shift, multvs_(y, (-C[i] * py - B[i] * px)));
            para->direction = norm_(addvv_(shift, para->direction));
        }
    }
}

static cplx calc_mode_U(cplx A, cplx B, cplx C, cplx detG0, vec_t pt, vec_t E, vec_t H, vec_t direction, cplx k,
                        double phase, double inv_root_area) {  /* cfields.pyx:118-151 */
    const cplx rootI = csqrt(CMPLX(0.0, 1.0));
    cplx denom, AA, CC, U;
    double x = dotprod_(pt, E), y = dotprod_(pt, H), z = dotprod_(pt, direction);
    denom = (RE(1) + (RE(z) * (A + C)) + RE(z * z) * detG0) * RE(2);
    AA = (A + RE(z) * detG0) / denom;
    CC = (C + RE(z) * detG0) / denom;
    U = cexp(CMPLX(0.0, 1.0) * (RE(phase) + k * (RE(z) + AA * RE(x * x) + B * RE(2 * x * y) / denom + CC * RE(y * y))));
    U /= csqrt(((RE(1) + RE(z) * A) * (RE(1) + RE(z) * C) - (RE(z) * B) * (RE(z) * B)) * CMPLX(0.0, 1.0));
    U *= RE(inv_root_area) * rootI;
    return U;
}

/* build_interaction_matrix (cfields.pyx:295-395) without its output size
 * estimate: returns the number of pairs; data (complex), xi, xj sized by
 * the caller for n * l pairs. */
uint64_t oracle_build_interaction_matrix(const double *rx, const double *ry, uint64_t N, const double *tx,
                                         const double *ty, uint64_t L, const double *A, const double *B,
                                         const double *C, const double *x, const double *y, const double *z,
                                         double wavelength, double spacing, double max_spacing, double blending,
                                         double *data, uint64_t *xi, uint64_t *xj) {
    uint64_t ct = 0;
    double phase = 0.0, k, inv_root_area, impart;
    k = 2000.0 * M_PI / wavelength;
    impart = 2 * (blending * blending) / (k * spacing * spacing);
    inv_root_area = 1.0;
    for (uint64_t i = 0; i < N; i++) {
        vec_t o = V(rx[i], ry[i], 0.0);
        vec_t a = V(x[3 * i], x[3 * i + 1], x[3 * i + 2]);
        vec_t b = V(y[3 * i], y[3 * i + 1], y[3 * i + 2]);
        vec_t c = V(z[3 * i], z[3 * i + 1], z[3 * i + 2]);
        cplx _A = CMPLX(A[i], impart), _B = CMPLX(B[i], 0), _C = CMPLX(C[i], impart);
        cplx detG0 = (_A * _C);
        for (uint64_t j = 0; j < L; j++) {
            vec_t pt = V(tx[j], ty[j], 0.0);
            pt = subvv_(pt, o);
            if (((pt.x * pt.x) + (pt.y * pt.y)) <= (max_spacing * max_spacing)) {
                cplx u = calc_mode_U(_A, _B, _C, detG0, pt, a, b, c, RE(k), phase, inv_root_area);
                xi[ct] = i;
                xj[ct] = j;
                data[2 * ct] = creal(u);
                data[2 * ct + 1] = cimag(u);
                ct += 1;
            }
        }
    }
    return ct;
}

int oracle_sizes(int which) {
    switch (which) {
    case 0: return (int)sizeof(oray_t);
    case 1: return (int)sizeof(opara_t);
    case 2: return (int)sizeof(ogausslet_t);
    default: return -1;
    }
}
