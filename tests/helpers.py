"""Shared helpers for the parity tests."""
import numpy as np


def rel_linf(x, ref):
    """Normalised L-infinity error ||x - ref||_inf / ||ref||_inf (0 if both are 0)."""
    x = np.asarray(x)
    ref = np.asarray(ref)
    den = np.max(np.abs(ref)) if ref.size else 0.0
    num = np.max(np.abs(x - ref)) if ref.size else 0.0
    if den == 0.0:
        return float(num)
    return float(num / den)
