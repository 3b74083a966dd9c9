"""Helpers shared by the GPU tests (random rays, one context per session)."""
import numpy as np


def random_rays(rng, n, center, radius, tmin=1e-4, tmax=1e8):
    o = center + (rng.random((n, 3)) * 2 - 1) * radius
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, :3] = o
    r[:, 3:6] = d
    r[:, 6] = tmin
    r[:, 7] = tmax
    return r


def compare_hits(p_gpu, t_gpu, p_ref, t_ref):
    """First-hit parity: ids equal, or an exact/near tie (|dt| <= 4 ulp-ish relative)."""
    same = p_gpu == p_ref
    tol = 4e-7 * np.maximum(1.0, np.abs(t_ref))
    tie = (~same) & (p_gpu >= 0) & (p_ref >= 0) & (np.abs(t_gpu - t_ref) <= tol)
    bad = (~same) & (~tie)
    return same, tie, bad
