"""2-D phase unwrapping (drop-in for raypier.core.unwrap2d).

unwrap2d (core/unwrap2d.py:8-28): Itoh's method along each axis in turn
(numpy.unwrap over axis 0, then axis 1), shifted by whole turns so that the
anchor pixel keeps its wrapped value, plus the residue of every 2x2 cell.
Small host arrays (the decomposition grid); no device work.
"""
import numpy as np


def unwrap2d(phase_array, anchor=(0, 0)):
    phase = np.asarray(phase_array)
    d0 = np.diff(phase, axis=0)
    d1 = np.diff(phase, axis=1)
    residuals = d0[:, :-1] - d0[:, 1:] - d1[:-1, :] + d1[1:, :]
    out = np.unwrap(np.unwrap(phase, axis=0), axis=1)
    a, b = anchor
    turns = int((out[a, b] - phase[a, b]) / (2 * np.pi))
    out -= turns * np.pi * 2
    return out, residuals
