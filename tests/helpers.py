"""Shared helpers for the tests."""
import numpy as np


def rel_err_per_var(a, b, idx):
    """max |a-b| / max |b| per variable over the points `idx`."""
    d = np.abs(a[:, idx] - b[:, idx]).max(axis=1)
    s = np.abs(b[:, idx]).max(axis=1)
    s = np.where(s == 0.0, 1.0, s)
    return d / s


TOL = 1e-11  # per-variable relative parity bound (max|dRHS| / max|RHS|)
