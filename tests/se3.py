"""SE(3) helpers for the differentiable-raycast tests (FP64)."""
import math

import numpy as np


def hat(w):
    return np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]], np.float64)


def se3_exp(xi):
    """exp of the twist xi = (omega, v): rotation by Rodrigues, translation
    through the left Jacobian V."""
    w, v = np.asarray(xi[:3], np.float64), np.asarray(xi[3:], np.float64)
    th = float(np.linalg.norm(w))
    Wh = hat(w)
    if th < 1e-12:
        R = np.eye(3) + Wh
        V = np.eye(3) + 0.5 * Wh
    else:
        A = math.sin(th) / th
        B = (1 - math.cos(th)) / th ** 2
        C = (th - math.sin(th)) / th ** 3
        R = np.eye(3) + A * Wh + B * Wh @ Wh
        V = np.eye(3) + B * Wh + C * Wh @ Wh
    T = np.eye(4)
    T[:3, :3] = R
    T[:3, 3] = V @ v
    return T


def sphere_tsdf(vol, spheres, trunc):
    """TSDF of a union of spheres, clamped to [-1, 1], reference layout."""
    d = vol.voxel_mm
    ax = [(np.arange(n) + s) * d + d / 2 for n, s in zip(vol.dims, vol.start)]
    X, Y, Z = np.meshgrid(*ax, indexing="ij")
    F = np.full(X.shape, np.inf)
    for cx, cy, cz, r in spheres:
        F = np.minimum(F, (np.sqrt((X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2) - r) / trunc)
    return np.clip(F, -1, 1).ravel()
