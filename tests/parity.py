"""Comparison helpers shared by the tests.

Floating-point parity uses the normwise (max-norm) relative error
||a - b||_inf / ||b||_inf, evaluated per (P,R) node for local diffusion:
the PSD of one node spans ~15 decades in energy, and a banded solve with a
different (but equally stable) operation order carries an absolute error of
order eps*cond*||x|| on every component, so pointwise relative error on the
smallest components is not a meaningful bar for any FP64 implementation.
"""
import numpy as np


def rel_err(a, b):
    """||a - b||_inf / ||b||_inf (0 when both are zero)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if b.size == 0:
        return 0.0
    den = np.abs(b).max()
    num = np.abs(a - b).max()
    if den == 0:
        return float(num)
    return float(num / den)


def node_rel_err(a, b):
    """Largest per-(P,R)-node normwise error of two [P][R][V][K] arrays."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    num = np.abs(a - b).max(axis=(2, 3))
    den = np.abs(b).max(axis=(2, 3))
    den = np.where(den == 0, 1.0, den)
    return float((num / den).max())


def bit_equal(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return a.shape == b.shape and bool(np.all(a.view(np.uint64) == b.view(np.uint64)))
