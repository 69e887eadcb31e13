"""Comparison helpers shared by the parity tests."""
import numpy as np


def rel_norm_err(a, ref):
    """max_i |a_i - ref_i| / max_i |ref_i| (normwise relative error)."""
    a = np.asarray(a, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    scale = np.max(np.abs(ref))
    if scale == 0.0:
        return float(np.max(np.abs(a)))
    return float(np.max(np.abs(a - ref)) / scale)


def disp_err(gpu_disp, ref_disp):
    """Nodal displacement error: max_i |u_i - u_i^ref| / max_i |u_i^ref| over
    node vectors."""
    d = np.linalg.norm(np.asarray(gpu_disp) - np.asarray(ref_disp), axis=0)
    s = np.linalg.norm(np.asarray(ref_disp), axis=0)
    return float(d.max() / s.max())
