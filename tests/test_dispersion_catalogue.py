"""Dispersion formulas 1, 3 and 5 of BaseDispersionCurve (cmaterials.pyx:
104-149) in the oracle, pinned by rows of the reference's glass catalogue
(tests/golden/dispersion_rows.json, extracted from
raypier/material_data/glass_dispersion_database.db by
tests/golden/make_dispersion_rows.py).

* Formula 3 (polynomial, n^2 = c0 + sum c_i lambda^p_i): glasses that other
  makers sell as equivalents of a Schott glass (K-BK7 ~ N-BK7, E-SF5 ~ SF5,
  J-SF6 ~ SF6 ...) have the same catalogue n_d and Abbe number to the
  catalogue's rounding.  Evaluated with formula 3, their n_d must match the
  Schott glass's formula-2 n_d (itself pinned by the N-BK7 catalogue values)
  to 6e-4 and the Abbe number to 2% (most pairs agree to 1e-5 in n_d; the
  largest gaps are catalogue differences such as K-BK7 1.51633 against
  N-BK7 1.51680, or lead-free N-LAF7 against LAF7).  A sign flip of an exponent, a dropped
  term or a shifted coefficient pair misses by orders of magnitude more.
  (The CDGM "H-" glasses are near-equivalents with different indices and are
  left out.)
* Formula 1 (Sellmeier with the pole wavelengths squared at evaluation,
  n^2 = 1 + c0 + sum B_i lambda^2 / (lambda^2 - C_i^2)): a formula-2
  catalogue row rewritten with C_i -> sqrt(C_i) must give the formula-2
  index, and the catalogue's chalcogenide rows (IG2 ... AMTIR-3) must show
  normal dispersion with indices between 2.4 and 3.0 over their range.
* Formula 5: the reference's sellmeier_5 has no return statement, so the
  value a Cython `cdef double` function yields is 0.0; only the absorption
  (imaginary) part remains.
"""
import json
import os

import numpy as np

import oracle
from raypier_optics_amd import _abi
from raypier_optics_amd.cmaterials import BaseDispersionCurve

HERE = os.path.dirname(os.path.abspath(__file__))
ROWS = json.load(open(os.path.join(HERE, "golden", "dispersion_rows.json")))
LINES = [0.5875618, 0.4861327, 0.6562725]    # d, F, C


def curve(row, absorption=0.0):
    return BaseDispersionCurve(row["formula_id"], np.array(row["coefs"], np.float64), absorption,
                               row["wavelength_min"], row["wavelength_max"])


def n_of(row, wl, absorption=0.0):
    rec = np.zeros((), _abi.dispersion_dtype)
    curve(row, absorption)._fill(rec)
    return oracle.dispersion(rec, wl)


def equivalent_pairs():
    return [p for p in ROWS["formula_3_pairs"] if p["poly"]["book"] != "cdgm"]


def test_formula3_matches_schott_equivalents():
    pairs = equivalent_pairs()
    assert len(pairs) >= 100
    for p in pairs:
        a = n_of(p["poly"], LINES).real
        b = n_of(p["schott"], LINES).real
        assert abs(a[0] - b[0]) <= 6e-4, (p["poly"]["name"], a[0], b[0])
        va, vb = (a[0] - 1) / (a[1] - a[2]), (b[0] - 1) / (b[1] - b[2])
        assert abs(va / vb - 1) <= 0.02, (p["poly"]["name"], va, vb)


def test_formula1_is_formula2_with_squared_poles():
    for p in ROWS["formula_3_pairs"][:40]:
        row2 = p["schott"]
        c = np.array(row2["coefs"])
        row1 = dict(row2, formula_id=1, coefs=[c[0]] + [v for k in range(1, len(c), 2)
                                                          for v in (c[k], np.sqrt(c[k + 1]))])
        wl = np.linspace(row2["wavelength_min"], row2["wavelength_max"], 17)
        np.testing.assert_allclose(n_of(row1, wl).real, n_of(row2, wl).real, rtol=2e-15)


def test_formula1_catalogue_chalcogenides():
    assert len(ROWS["formula_1"]) == 8
    for row in ROWS["formula_1"]:
        wl = np.linspace(max(row["wavelength_min"], 2.0), row["wavelength_max"], 25)
        n = n_of(row, wl).real
        assert np.all((n > 2.4) & (n < 3.0)), row["name"]
        assert np.all(np.diff(n) < 0), row["name"]


def test_formula5_returns_zero_real_part():
    for row in ROWS["formula_5"]:
        n = n_of(row, [0.5875618, 1.0], absorption=2.0)
        assert np.all(n.real == 0.0)
        np.testing.assert_allclose(n.imag, 1e-4 * 2.0 / (4 * np.pi) * np.array([0.5875618, 1.0]), rtol=1e-15)
