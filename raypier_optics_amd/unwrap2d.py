"""2-D phase unwrapping (drop-in for raypier.core.unwrap2d).

unwrap2d(phase_array, anchor) takes a (N, M) array of wrapped phases and
returns (unwrapped, residues) like core/unwrap2d.py:8-28, computed here by
the unweighted least-squares method (Ghiglia and Romero, JOSA A 11, 107,
1994): the wrapped phase differences along both axes define a Poisson
equation for the unwrapped phase, solved with Neumann boundaries by a
type-II discrete cosine transform.  The least-squares surface is then made
congruent with the input: the surface is shifted to pass through the anchor
pixel's wrapped value, and each pixel is moved by the whole number of turns
that brings it nearest the surface (so the anchor keeps its wrapped value).

residues[i, j] is the circulation of the wrapped differences around the
2 x 2 cell with corner (i, j), in turns: 0 wherever the wrapped data come
from a continuous phase, +-1 at a phase singularity.  Small host arrays (the
decomposition grid); no device work.
"""
import numpy as np

TWO_PI = 2.0 * np.pi


def wrap(a):
    """Phase(s) reduced to [-pi, pi)."""
    return (np.asarray(a) + np.pi) % TWO_PI - np.pi


def _poisson_neumann(rho):
    """Solution phi (mean zero) of the discrete Laplacian phi = rho with
    reflecting boundaries, through the DCT-II eigenbasis."""
    from scipy.fft import dctn, idctn
    n, m = rho.shape
    lam = (2.0 * np.cos(np.pi * np.arange(n) / n) - 2.0)[:, None] + \
          (2.0 * np.cos(np.pi * np.arange(m) / m) - 2.0)[None, :]
    spec = dctn(rho, type=2, norm="ortho")
    lam[0, 0] = 1.0
    spec = spec / lam
    spec[0, 0] = 0.0
    return idctn(spec, type=2, norm="ortho")


def unwrap2d(phase_array, anchor=(0, 0)):
    psi = np.asarray(phase_array, dtype=np.float64)
    n, m = psi.shape
    g0 = wrap(np.diff(psi, axis=0))          # (n-1, m)
    g1 = wrap(np.diff(psi, axis=1))          # (n, m-1)
    # divergence of the wrapped gradient field, zero flux across the border
    rho = np.zeros_like(psi)
    rho[:-1, :] += g0
    rho[1:, :] -= g0
    rho[:, :-1] += g1
    rho[:, 1:] -= g1
    surface = _poisson_neumann(rho) if (n > 1 or m > 1) else np.zeros_like(psi)
    # the surface is fixed up to a constant: pin it to the anchor's wrapped
    # value, then move every pixel by the turns that bring it to the surface
    a, b = anchor
    surface += psi[a, b] - surface[a, b]
    out = psi + TWO_PI * np.round((surface - psi) / TWO_PI)
    residues = np.rint((g0[:, :-1] + g1[1:, :] - g0[:, 1:] - g1[:-1, :]) / TWO_PI)
    return out, residues
