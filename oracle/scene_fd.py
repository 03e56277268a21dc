"""CPU restatement of the reprojection residuals whose Jacobian the reference
builds with Ceres (src/uncertaintyTE/snavely_reprojection_error.h:110-302,
composed by JacobianComposer::scene2Jacobian, JacobianComposer.cpp:84-194),
and their Jacobian by central finite differences.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the reference results for the
device Jacobian kernel (unc_scene_jacobian); the product package never imports it.

Pinning: the residual is written the way the reference's functors are (the
angle-axis rotation with its small-angle branch, :110-140); the derivatives
are not restated at all but differenced, so a wrong sign, index or chain-rule
term in the analytic kernel shows up as a mismatch well above the differencing
error (about 1e-8 relative with the steps below).
"""
from __future__ import annotations

import numpy as np

EPS = np.finfo(np.float64).eps
# free camera parameters per model (cp) and whether k2 enters the residual
MODELS = {0: (6, True), 1: (7, True), 2: (8, False), 3: (9, True)}


def rotate_rc(aa, pt):
    """AngleAxisRotatePointRC (snavely_reprojection_error.h:110-140), batched."""
    th2 = np.sum(aa * aa, -1)
    big = th2 > EPS
    th = np.sqrt(np.where(big, th2, 1.0))
    w = aa / th[..., None]
    cos, sin = np.cos(th)[..., None], np.sin(th)[..., None]
    wxp = np.cross(w, pt)
    tmp = np.sum(w * pt, -1)[..., None] * (1.0 - cos)
    full = pt * cos + wxp * sin + w * tmp
    first = pt + np.cross(aa, pt)
    return np.where(big[..., None], full, first)


def residual(params, X, model):
    """Predicted (u, v) of the reference's functor for this model:
    params = (aa, C, f, k1, k2) (..., 9), X (..., 3)."""
    aa, C = params[..., 0:3], params[..., 3:6]
    f, k1, k2 = params[..., 6], params[..., 7], params[..., 8]
    if not MODELS[model][1]:
        k2 = np.zeros_like(k2)
    p = rotate_rc(aa, X - C)
    x, y = p[..., 0] / p[..., 2], p[..., 1] / p[..., 2]
    d = x * x + y * y
    g = 1.0 + k1 * d + k2 * d * d
    return np.stack([f * x * g, f * y * g], -1)


def jacobian_fd(cams, points, obs_cam, obs_pt, model):
    """(n_obs, 2, cp + 3) by central differences, camera columns first."""
    cp = MODELS[model][0]
    P = cams[obs_cam].astype(np.float64)
    X = points[obs_pt].astype(np.float64)
    out = np.zeros((len(obs_cam), 2, cp + 3))
    for j in range(cp + 3):
        if j < cp:
            h = 1e-6 * np.maximum(1.0, np.abs(P[:, j]))
            Pp, Pm = P.copy(), P.copy()
            Pp[:, j] += h
            Pm[:, j] -= h
            d = residual(Pp, X, model) - residual(Pm, X, model)
        else:
            k = j - cp
            h = 1e-6 * np.maximum(1.0, np.abs(X[:, k]))
            Xp, Xm = X.copy(), X.copy()
            Xp[:, k] += h
            Xm[:, k] -= h
            d = residual(P, Xp, model) - residual(P, Xm, model)
        out[:, :, j] = d / (2 * h)[:, None]
    return out
