"""Zernike distortion: oracle pinned by the reference's tests and closed
forms; the compiled straight-line program (what the GPU runs) checked
bit for bit against the oracle's literal memoised recursion.

Reference tests restated: test/test_cdistortions.py:54-93 (ANSI table),
:34-51/:164-180 (radial closed forms), :99-142 (j7 against
SimpleTestZernikeJ7)."""
import math
from collections import Counter

import numpy as np
import pytest

import oracle
from raypier_optics_amd.cdistortions import ZernikeDistortion, eval_nmk, jnm_map, compile_zernike
from raypier_optics_amd._scene import SceneBuilder

RADIAL = {  # closed forms (test_cdistortions.py:34-51)
    (0, 0): lambda r: 1.0, (1, 1): lambda r: r, (2, 0): lambda r: 2 * r ** 2 - 1, (2, 2): lambda r: r ** 2,
    (3, 1): lambda r: 3 * r ** 3 - 2 * r, (3, 3): lambda r: r ** 3, (4, 0): lambda r: 6 * r ** 4 - 6 * r ** 2 + 1,
    (4, 2): lambda r: 4 * r ** 4 - 3 * r ** 2, (4, 4): lambda r: r ** 4,
    (5, 1): lambda r: 10 * r ** 5 - 12 * r ** 3 + 3 * r, (5, 3): lambda r: 5 * r ** 5 - 4 * r ** 3,
    (5, 5): lambda r: r ** 5, (6, 0): lambda r: 20 * r ** 6 - 30 * r ** 4 + 12 * r ** 2 - 1,
    (6, 2): lambda r: 15 * r ** 6 - 20 * r ** 4 + 6 * r ** 2, (6, 4): lambda r: 6 * r ** 6 - 5 * r ** 4,
    (6, 6): lambda r: r ** 6}


def tables_for(dist):
    b = SceneBuilder()
    zi = b.distortion_index(dist)
    return b.finish(), zi


def test_ansi_table():
    check = [(0, 0, 0), (1, 1, -1), (2, 1, 1), (3, 2, -2), (4, 2, 0), (5, 2, 2), (6, 3, -3), (7, 3, -1),
             (8, 3, 1), (9, 3, 3), (10, 4, -4), (11, 4, -2), (12, 4, 0), (13, 4, 2), (14, 4, 4), (15, 5, -5),
             (16, 5, -3), (17, 5, -1), (18, 5, 1), (19, 5, 3)]
    eval_nmk(19)
    assert [(i, n, m) for i, (n, m, k) in enumerate(jnm_map[:20])] == check
    assert [(i, n, m) for i, (n, m, k) in enumerate(eval_nmk(i) for i in range(20))] == check
    dups = {(n, abs(m)): k for n, m, k in jnm_map}
    assert max(Counter(dups.values()).values()) <= 1


def test_radial_polynomials_closed_form():
    """Single terms with m >= 0 at theta = 0: sag = N * value * R_n^m(r)."""
    for (n, m), f in RADIAL.items():
        j = (n * (n + 2) + m) // 2
        t, zi = tables_for(ZernikeDistortion([(j, 1.0)], unit_radius=1.0))
        r = np.linspace(0.0, 1.0, 11)
        sag, _ = oracle.zernike(t, zi, r, np.zeros_like(r))
        N = math.sqrt(n + 1) if m == 0 else math.sqrt(2 * (n + 1))
        assert np.allclose(sag / N, [f(x) for x in r], atol=1e-13), (n, m)


def j7_closed_form(x, y, R):
    """SimpleTestZernikeJ7 (cdistortions.pyx:39-81): Z7 = sqrt(8)(3 rho^2 - 2) y."""
    x, y = x / R, y / R
    root8 = math.sqrt(8)
    z = root8 * (3 * (x * x + y * y) - 2) * y
    gx = root8 * 6 * x * y / R
    gy = root8 * (3 * x * x + 9 * y * y - 2) / R
    return z, gx, gy


def test_j7_against_closed_form():
    """test_cdistortions.py:99-142, on the 4x4 grid and at the origin."""
    t, zi = tables_for(ZernikeDistortion(j7=1.0, unit_radius=10.0))
    xs = np.linspace(-5., 5., 4)
    x, y = (a.ravel() for a in np.meshgrid(xs, xs))
    x, y = np.append(x, 0.0), np.append(y, 0.0)
    sag, grad = oracle.zernike(t, zi, x, y)
    for i in range(len(x)):
        z, gx, gy = j7_closed_form(x[i], y[i], 10.0)
        assert abs(sag[i] - z) < 1e-13 and abs(grad[i, 2] - z) < 1e-13
        assert abs(grad[i, 0] - gx) < 1e-13 and abs(grad[i, 1] - gy) < 1e-13


def numdiff(t, zi, x, y, h=1e-6):
    sxp, _ = oracle.zernike(t, zi, x + h, y)
    sxm, _ = oracle.zernike(t, zi, x - h, y)
    syp, _ = oracle.zernike(t, zi, x, y + h)
    sym, _ = oracle.zernike(t, zi, x, y - h)
    return (sxp - sxm) / (2 * h), (syp - sym) / (2 * h)


# Terms whose gradient the reference computes without reading a shared
# derivative slot; every other term with n >= 4 reads a slot that holds a
# different (n, m) value (see the two tests below).
CLEAN_TERMS = list(range(1, 11)) + [12, 14]


def test_single_term_gradient_is_derivative_of_sag():
    rng = np.random.default_rng(5)
    x, y = rng.uniform(-4, 4, 30), rng.uniform(-4, 4, 30)
    for j in CLEAN_TERMS:
        t, zi = tables_for(ZernikeDistortion([(j, 0.7)], unit_radius=5.0))
        _, g = oracle.zernike(t, zi, x, y)
        gx, gy = numdiff(t, zi, x, y)
        assert np.allclose(g[:, 0], gx, atol=1e-7) and np.allclose(g[:, 1], gy, atol=1e-7), j


def test_single_42_term_reads_shared_slot():
    """For a lone (4,+-2) term, the R'(3,1) sub-recursion caches R'_{2,0} in
    slot (1, 2); the C-term of R'(4,2) then reads that slot expecting
    R'_{2,2}: the radial derivative is off by R'_{2,2} - R'_{2,0} = -2r."""
    rng = np.random.default_rng(8)
    x, y = rng.uniform(-4, 4, 30), rng.uniform(-4, 4, 30)
    ur = 5.0
    for j, m in ((11, -2), (13, 2)):
        t, zi = tables_for(ZernikeDistortion([(j, 0.7)], unit_radius=ur))
        _, g = oracle.zernike(t, zi, x, y)
        gx, gy = numdiff(t, zi, x, y)
        r, th = np.hypot(x / ur, y / ur), np.arctan2(y, x)
        PH = np.cos(m * th) if m >= 0 else -np.sin(m * th)
        q = math.sqrt(10) * 0.7 * (-2 * r) * PH
        assert np.allclose(g[:, 0] - gx, q * np.cos(th) / ur, atol=1e-7)
        assert np.allclose(g[:, 1] - gy, q * np.sin(th) / ur, atol=1e-7)


def test_derivative_cache_slot_sharing_reproduced():
    """The Noll 4-15 set: the reference's R' recursion for (4,2) reads
    R'_{2,0} = 4r from the slot where it expects R'_{2,2} = 2r (C-term index
    (h-1)*h + |m|, cdistortions.pyx:211), so terms j=11 and j=13 carry
    R'_{4,2} - 2r; all other terms of the set are exact derivatives.  The
    oracle must reproduce exactly that."""
    coefs = [(j, 0.01 * (j - 2)) for j in range(3, 15)]   # Noll 4-15 in ANSI order
    ur = 8.0
    t, zi = tables_for(ZernikeDistortion(coefs, unit_radius=ur))
    rng = np.random.default_rng(6)
    x, y = rng.uniform(-5, 5, 40), rng.uniform(-5, 5, 40)
    _, g = oracle.zernike(t, zi, x, y)
    gx, gy = numdiff(t, zi, x, y)
    xr, yr = x / ur, y / ur
    r, th = np.hypot(xr, yr), np.arctan2(yr, xr)
    quirk = np.zeros_like(r)
    for j, m in ((11, -2), (13, 2)):
        v = dict(coefs)[j]
        N = math.sqrt(2 * 5) * v
        PH = np.cos(m * th) if m >= 0 else -np.sin(m * th)
        quirk += N * (-2 * r) * PH
    assert np.allclose(g[:, 0] - gx, quirk * np.cos(th) / ur, atol=1e-7)
    assert np.allclose(g[:, 1] - gy, quirk * np.sin(th) / ur, atol=1e-7)
    assert np.abs(quirk).max() > 1e-3   # the effect is real, not rounding


# --------------------------------------------------------- compiled program
def run_program(insns, n_slots, r):
    slot = [float("nan")] * n_slots

    def val(ref):
        return slot[ref] if ref >= 0 else (0.0 if ref == -1 else 1.0)
    for ins in insns:
        a, b, c, d, e = (val(int(s)) for s in ins["src"])
        if ins["op"] == 0:
            v = r * (a + b)
            v -= c
        elif ins["op"] == 1:
            v = a + b
            v += r * (c + d)
            v -= e
        else:
            v = a + b
            v -= c
        slot[int(ins["dst"])] = v
    return val


def interpret(t, zi, x, y):
    """The compiled programs of the scene tables, run in Python floats with
    the reference's summation (libm calls are the same C functions the oracle
    uses): the form the polynomial tables the kernels evaluate are expanded
    from."""
    Z = t.zernikes[zi]
    coefs = t.zcoefs[Z["coef_begin"]: Z["coef_begin"] + Z["n_coefs"]]
    sag_p = t.zinsns[Z["sag_begin"]: Z["sag_begin"] + Z["sag_len"]]
    grad_p = t.zinsns[Z["grad_begin"]: Z["grad_begin"] + Z["grad_len"]]
    ur = float(Z["unit_radius"])
    out_s, out_g = [], []
    for X, Y in zip(x, y):
        X, Y = X / ur, Y / ur
        r = math.sqrt(X * X + Y * Y)
        th = math.atan2(Y, X)
        v = run_program(sag_p, int(Z["n_slots"]), r)
        z = 0.0
        for c in coefs:
            n, m = int(c["n"]), int(c["m"])
            N = math.sqrt(n + 1) if m == 0 else math.sqrt(2 * (n + 1))
            N *= float(c["value"])
            PH = math.cos(m * th) if m >= 0 else -math.sin(m * th)
            z += N * v(int(c["sag_ref"])) * PH
        out_s.append(z)
        v = run_program(grad_p, int(Z["n_slots"]), r)
        gx = gy = gz = 0.0
        st, ct = math.sin(th), math.cos(th)
        for c in coefs:
            n, m = int(c["n"]), int(c["m"])
            if m >= 0:
                PH, PHp = math.cos(m * th), -m * math.sin(m * th)
            else:
                PH, PHp = -math.sin(m * th), -m * math.cos(m * th)
            R, Rp, Ror = v(int(c["r_ref"])), v(int(c["rp_ref"])), v(int(c["ror_ref"]))
            N = math.sqrt(n + 1) if m == 0 else math.sqrt(2 * (n + 1))
            N *= float(c["value"])
            gz += N * R * PH
            gx += N * (Rp * ct * PH + Ror * (-st) * PHp)
            gy += N * (Rp * st * PH + Ror * (ct) * PHp)
        out_g.append((gx / ur, gy / ur, gz))
    return np.array(out_s), np.array(out_g)


@pytest.mark.parametrize("coefs", [
    [(j, 0.01 * (j - 2)) for j in range(3, 15)],                 # Noll 4-15
    [(7, 1.0)], [(4, 0.3), (13, -0.2)], [(12, 1.0), (24, 0.5), (4, -0.1)],
    [(j, 0.1 / (1 + j)) for j in range(0, 45)],                   # every term to n = 8
])
def test_compiled_program_matches_literal_recursion(coefs):
    t, zi = tables_for(ZernikeDistortion(coefs, unit_radius=7.5))
    rng = np.random.default_rng(7)
    x, y = rng.uniform(-7, 7, 25), rng.uniform(-7, 7, 25)
    x[0] = y[0] = 0.0
    s_ref, g_ref = oracle.zernike(t, zi, x, y)
    s, g = interpret(t, zi, x, y)
    assert np.array_equal(s, s_ref)
    assert np.array_equal(g, g_ref)


def test_program_independent_of_operand_order():
    """C leaves the evaluation order of f() + g() unspecified.  For the Noll
    4-15 set the recursion's cache contents do not depend on it: swapping the
    two recursive operands gives the same values."""
    coefs = [(j, 0.01 * (j - 2)) for j in range(3, 15)]
    kmax, cs, sag, grad, _ = compile_zernike(sorted(coefs), 1.0)
    import raypier_optics_amd.cdistortions as cd
    orig_R, orig_Rp = cd._ZProgram.R, cd._ZProgram.Rp

    def R_swapped(self, k, n, m):
        if n < m:
            return cd.CONST_ZERO
        if n == 0:
            return cd.CONST_ONE
        s = self._slot(0, k)
        if s in self.written:
            return s
        nA, mA = n - 1, abs(m - 1)
        h = cd._cdiv(nA, 2)
        kA, kB = h * (h + 1) + abs(mA), h * (h + 1) + abs(m + 1)
        h = cd._cdiv(n - 2, 2)
        kC = h * (h + 1) + abs(m)
        b = self.R(kB, nA, m + 1)
        a = self.R(kA, nA, mA)
        c = self.R(kC, n - 2, m)
        self._emit(0, s, [a, b, c])
        return s
    try:
        cd._ZProgram.R = R_swapped
        kmax2, cs2, sag2, grad2, _ = compile_zernike(sorted(coefs), 1.0)
    finally:
        cd._ZProgram.R = orig_R
    for r in np.linspace(0, 1, 9):
        v1, v2 = run_program(grad, 3 * kmax, r), run_program(grad2, 3 * kmax, r)
        for c1, c2 in zip(cs, cs2):
            for ref in ("r_ref", "rp_ref", "ror_ref"):
                assert v1(int(c1[ref])) == v2(int(c2[ref]))


# ------------------------------------------------------ polynomial tables
def horner(c, k, r):
    d = int(c["deg"][k])
    v = float(c["poly"][k][d])
    for i in range(d - 1, -1, -1):
        v = v * r + float(c["poly"][k][i])
    return v


@pytest.mark.parametrize("coefs", [
    [(j, 0.01 * (j - 2)) for j in range(3, 15)],                 # Noll 4-15
    [(7, 1.0)], [(4, 0.3), (13, -0.2)], [(12, 1.0), (24, 0.5), (4, -0.1)],
    [(j, 0.1 / (1 + j)) for j in range(0, 45)],                   # every term to n = 8
])
def test_polynomial_tables_equal_programs(coefs):
    """The kernels evaluate each value a coefficient uses from its integer
    polynomial in r: at every r it equals the compiled program's value (the
    literal recursion, slot sharing included) to rounding, and the sag and
    gradient built from it agree with the oracle's recursion to 1e-12."""
    t, zi = tables_for(ZernikeDistortion(coefs, unit_radius=7.5))
    Z = t.zernikes[zi]
    coefs_t = t.zcoefs[Z["coef_begin"]: Z["coef_begin"] + Z["n_coefs"]]
    sag_p = t.zinsns[Z["sag_begin"]: Z["sag_begin"] + Z["sag_len"]]
    grad_p = t.zinsns[Z["grad_begin"]: Z["grad_begin"] + Z["grad_len"]]
    for r in np.linspace(0.0, 1.05, 23):
        vs = run_program(sag_p, int(Z["n_slots"]), r)
        vg = run_program(grad_p, int(Z["n_slots"]), r)
        for c in coefs_t:
            scale = np.abs(c["poly"]).sum() + 1.0
            for k, (ref, v) in enumerate((("sag_ref", vs), ("r_ref", vg), ("rp_ref", vg), ("ror_ref", vg))):
                assert abs(horner(c, k, r) - v(int(c[ref]))) <= 4e-16 * scale * max(1.0, r) ** 9
    # the whole sag / gradient from the tables against the oracle
    rng = np.random.default_rng(11)
    x, y = rng.uniform(-7, 7, 25), rng.uniform(-7, 7, 25)
    x[0] = y[0] = 0.0
    s_ref, g_ref = oracle.zernike(t, zi, x, y)
    ur = float(Z["unit_radius"])
    for X, Y, s0, g0 in zip(x, y, s_ref, g_ref):
        X, Y = X / ur, Y / ur
        r, th = math.hypot(X, Y), math.atan2(Y, X)
        z = gx = gy = 0.0
        for c in coefs_t:
            m = int(c["m"])
            N = float(c["nval"])
            if m >= 0:
                PH, PHp = math.cos(m * th), -m * math.sin(m * th)
            else:
                PH, PHp = -math.sin(m * th), -m * math.cos(m * th)
            R, Rp, Ror = horner(c, 1, r), horner(c, 2, r), horner(c, 3, r)
            z += N * horner(c, 0, r) * PH
            gx += N * (Rp * math.cos(th) * PH - Ror * math.sin(th) * PHp)
            gy += N * (Rp * math.sin(th) * PH + Ror * math.cos(th) * PHp)
        tol = 1e-12 * (1.0 + np.abs(g_ref).max() + np.abs(s_ref).max())
        assert abs(z - s0) <= tol
        assert abs(gx / ur - g0[0]) <= tol and abs(gy / ur - g0[1]) <= tol


def test_polynomial_tables_textbook():
    """Closed forms: R_2^0 = 2r^2 - 1, R_3^1 = 3r^3 - 2r, R_4^0 = 6r^4 - 6r^2 + 1
    and their derivatives (n < 5, where the recursion has no quirk)."""
    kmax, cs, sag, grad, _ = compile_zernike([(4, 1.0), (8, 1.0), (12, 1.0)], 1.0)
    want = {(2, 0): [-1, 0, 2], (3, 1): [0, -2, 0, 3], (4, 0): [1, 0, -6, 0, 6]}
    for c in cs:
        p = want[(int(c["n"]), int(c["m"]))]
        d = len(p) - 1
        assert int(c["deg"][1]) == d and list(c["poly"][1][:d + 1]) == p
        assert list(c["poly"][0][:d + 1]) == p
        dp = [k * p[k] for k in range(1, d + 1)]
        assert int(c["deg"][2]) == d - 1 and list(c["poly"][2][:d]) == dp
