"""Host side of the lens adjoint's Zernike distortions: the Cartesian
monomial expansion of each ZernikeDistortion term against the radial
polynomial tables the trace kernels evaluate (compile_zernike: the
reference's recursion run on exact integer polynomials), and the mapping
between coefficients and the monomial table."""
import math

import numpy as np

from raypier_optics_amd import adjoint, _abi
from raypier_optics_amd.cdistortions import compile_zernike, eval_nmk


def _direct(js, coefs, x, y):
    """sum_j c_j N_j R(r) cos/sin(m theta) from compile_zernike's tables."""
    _, recs, _, _, _ = compile_zernike(sorted(zip(js, coefs)), 1.0)
    r = math.hypot(x, y)
    th = math.atan2(y, x)
    out = 0.0
    for c in recs:
        deg = int(c["deg"][0])
        R = sum(float(a) * r ** i for i, a in enumerate(c["poly"][0][:deg + 1]))
        m = int(c["m"])
        ph = math.cos(m * th) if m >= 0 else math.sin(-m * th)
        out += float(c["nval"]) * R * ph
    return out


def test_monomials_match_radial_tables():
    rng = np.random.default_rng(3)
    js = list(range(0, 28))
    coefs = rng.normal(0, 1, len(js))
    table = sum(c * adjoint.zernike_monomials(j) for j, c in zip(js, coefs))
    for _ in range(50):
        rr, th = math.sqrt(rng.random()), 2 * math.pi * rng.random()
        x, y = rr * math.cos(th), rr * math.sin(th)
        via = sum(t * x ** a * y ** b for t, (a, b) in zip(table, adjoint.monomials()))
        want = _direct(js, coefs, x, y)
        assert abs(via - want) <= 1e-12 * (1 + abs(want)), (x, y, via, want)


def test_known_terms():
    """ANSI j=4 (defocus) = sqrt(3)(2 r^2 - 1); j=3 = sqrt(6) r^2 sin(2 theta);
    j=12 (spherical) = sqrt(5)(6 r^4 - 6 r^2 + 1)."""
    mono = adjoint.monomials()
    z4 = dict(zip(mono, adjoint.zernike_monomials(4)))
    assert math.isclose(z4[(0, 0)], -math.sqrt(3)) and math.isclose(z4[(2, 0)], 2 * math.sqrt(3))
    assert math.isclose(z4[(0, 2)], 2 * math.sqrt(3))
    z3 = dict(zip(mono, adjoint.zernike_monomials(3)))
    assert math.isclose(z3[(1, 1)], 2 * math.sqrt(6)) and sum(abs(v) for v in z3.values()) == z3[(1, 1)]
    z12 = dict(zip(mono, adjoint.zernike_monomials(12)))
    s5 = math.sqrt(5)
    for ab, v in {(0, 0): s5, (2, 0): -6 * s5, (0, 2): -6 * s5, (4, 0): 6 * s5, (2, 2): 12 * s5,
                  (0, 4): 6 * s5}.items():
        assert math.isclose(z12[ab], v)
    assert eval_nmk(12)[:2] == (4, 0)


def test_table_and_gradient_mapping():
    lens = adjoint.SequentialLens([0.0, 5.0], [[0.02, 0, 0, 0, 0, 0], [-0.02, 0, 0, 0, 0, 0]], [5.0, 5.0],
                                  [0, 1], [1, 0], [None, None], 50.0,
                                  zernike=[(4.0, [(4, 1e-3), (12, -2e-4)]), None])
    assert lens.has_zernike and list(lens.zcoefs) == [1e-3, -2e-4]
    T = lens.zernike_table()
    assert T.shape == (2, _abi.LENS_ZTERMS) and not T[1].any()
    want = 1e-3 * adjoint.zernike_monomials(4) - 2e-4 * adjoint.zernike_monomials(12)
    assert np.allclose(T[0], want, rtol=0, atol=1e-18)
    # chain rule: d/dc_j = sum_k T_jk d/dC_k
    g = np.zeros((2, _abi.LENS_ZTERMS))
    g[0] = np.arange(_abi.LENS_ZTERMS)
    gz = lens.zernike_grad(g)
    assert np.allclose(gz, [adjoint.zernike_monomials(4) @ g[0], adjoint.zernike_monomials(12) @ g[0]])
