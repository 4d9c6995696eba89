"""Derived fields and error metrics of a state (host-side numpy).

A state is the solver's four node-sized arrays (Ax, Ay, Az, V), shape
(4, Nz, Ny, Nx); coefficients are the six node-sized Yee arrays (Mu_x, Mu_y,
Mu_z, Sigma_x, Sigma_y, Sigma_z) that ``Solver.get_coeffs`` returns."""
from __future__ import annotations

import numpy as np


def rel(a, b) -> float:
    """||a - b|| / ||b|| (the absolute difference when b = 0)."""
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / nb) if nb > 0 else float(np.linalg.norm(a - b))


def a_part(f):
    return np.asarray(f)[:3].ravel()


def sigma_grad_v(h, coef, fields):
    """sigma * grad V on the free Ax, Ay, Az edges: the V coupling of the A rows
    (QMS_3D.f90:595-607).  Well conditioned even where V itself carries a
    conductor's floating constant (a constant potential has no gradient)."""
    V = fields[3]
    hx, hy, hz = h
    sx, sy, sz = coef[3], coef[4], coef[5]
    gx = sx[:, :, :-1] * (V[:, :, 1:] - V[:, :, :-1]) / hx
    gy = sy[:, :-1, :] * (V[:, 1:, :] - V[:, :-1, :]) / hy
    gz = sz[:-1, :, :] * (V[1:, :, :] - V[:-1, :, :]) / hz
    return np.concatenate([gx[2:-2, 2:-2, 1:-1].ravel(), gy[2:-2, 1:-1, 2:-2].ravel(), gz[1:-1, 2:-2, 2:-2].ravel()])
