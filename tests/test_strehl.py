"""Strehl ratio of traced rays at a reference point, against closed forms
(host tensors; no device work)."""
import numpy as np
import torch

from raypier_optics_amd import metrics


def _bundle(n, P, defocus_mm=0.0, wavelength_um=0.55, seed=3):
    """Rays whose last origins Q lie around a converging cone towards P,
    aimed at P, with optical path to P equal to 50 mm plus
    defocus_mm * rho^2 (rho^2 evenly spaced over [0, 1): a uniform pupil)."""
    rng = np.random.default_rng(seed)
    rho2 = (np.arange(n) + 0.5) / n
    th = rng.random(n) * 2 * np.pi
    r = 5.0 * np.sqrt(rho2)
    Q = np.stack([r * np.cos(th), r * np.sin(th), -20.0 + 0.3 * rng.standard_normal(n)], axis=1)
    v = np.asarray(P) - Q
    dist = np.linalg.norm(v, axis=1)
    d = v / dist[:, None]
    n_re = np.full(n, 1.0)
    acc = 50.0 - dist + defocus_mm * rho2
    phase = np.zeros(n)
    t = lambda a: torch.as_tensor(a, dtype=torch.float64)
    return t(Q), t(d), t(acc), t(phase), t(n_re)


def test_perfect_spherical_wave_is_one():
    P = (0.1, -0.2, 30.0)
    s = metrics.strehl_ratio(_bundle(20000, P), 0.55, P)
    assert abs(s - 1.0) < 1e-12


def test_defocus_closed_form():
    """W = a rho^2 over a uniform pupil: S = sinc^2(k a / 2)."""
    P = (0.0, 0.0, 25.0)
    lam = 0.55
    k = 2 * np.pi / (lam * 1e-3)
    for a in (0.2 * lam * 1e-3, 0.5 * lam * 1e-3, 1.3 * lam * 1e-3):
        x = k * a / 2
        want = (np.sin(x) / x) ** 2
        got = metrics.strehl_ratio(_bundle(200000, P, defocus_mm=a), lam, P)
        assert abs(got - want) < 1e-6, (a, got, want)


def test_phase_term_and_index():
    """A constant phase and a constant extra path leave S unchanged; a ray
    phase equal to -k W undoes the aberration."""
    P = (0.0, 0.0, 25.0)
    lam = 0.55
    k = 2 * np.pi / (lam * 1e-3)
    a = 0.7 * lam * 1e-3
    Q, d, acc, ph, n = _bundle(50000, P, defocus_mm=a)
    s0 = metrics.strehl_ratio((Q, d, acc, ph, n), lam, P)
    s1 = metrics.strehl_ratio((Q, d, acc + 3.0, ph + 1.234, n), lam, P)
    assert abs(s0 - s1) < 1e-9 and s0 < 0.9
    rho2 = (np.arange(50000) + 0.5) / 50000
    fixed = ph - torch.as_tensor(k * a * rho2)
    assert abs(metrics.strehl_ratio((Q, d, acc, fixed, n), lam, P) - 1.0) < 1e-9


def test_best_focus_finds_the_point():
    P = (0.0, 0.0, 25.0)
    paths = _bundle(20000, P)
    point, s = metrics.best_focus_strehl(paths, 0.55, (0.0, 0.0, 25.05), span_mm=0.1, steps=21)
    assert abs(point[2] - 25.0) < 1e-9 and abs(s - 1.0) < 1e-12
