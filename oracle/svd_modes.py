"""CPU restatement of the SVD algorithm modes of computeCovariances
(src/uncertaintyTE/compute.cpp:483-495 -> svdInverse :185-249, camera output
removeScaleJ4Z :168-183 with a NULL scale).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the reference results for the
device SVD path; the product package never imports it.

It follows the reference literally: the dense, unscaled M = J^T J over every
camera and point, LAPACK's dgesvd (SVD_QR_ITERATION) or dgesdd
(SVD_DEVIDE_AND_CONQUER) through scipy, U diag(1/sigma) V^T over the kept
singular values (the largest N - svdRemoveN, or those above epsilon), and the
camera blocks packed as (X[j,k] + X[k,j]/2).  The reference then crashes in its
point step (NULL scale, compute.cpp:402,413); the completion used here is the
same formula as the TE path with unit scales and every point:
Sigma_p = V^-1 + V^-1 (sum_o W_o^T Sigma_c W_o) V^-1, which equals
pinv(A) B Sigma B^T pinv(A)^T (compute.cpp:516-527) -- checked literally in the
tests.

Pinning: tests/test_oracle_svd.py checks the Penrose conditions of the
pseudo-inverse on the kept subspace, the agreement of the two LAPACK drivers
with an eigen-decomposition of M, and the point formula against the literal
pinv(A) B Sigma B^T pinv(A)^T form.
"""
from __future__ import annotations

import numpy as np
import scipy.linalg


def dense_jacobian(jac) -> np.ndarray:
    J = np.zeros((jac.num_rows, jac.num_cols))
    rows = np.asarray(jac.rows)
    for r in range(jac.num_rows):
        for e in range(rows[r], rows[r + 1]):
            J[r, jac.cols[e]] += jac.vals[e]
    return J


def pinv_svd(M: np.ndarray, algorithm: int, svd_remove_n: int = -1, epsilon: float = 1e-10):
    """svdInverse (compute.cpp:185-249): U diag(kept 1/sigma) V^T."""
    driver = "gesvd" if algorithm == 0 else "gesdd"
    U, s, Vt = scipy.linalg.svd(M, lapack_driver=driver)
    n = M.shape[0]
    if svd_remove_n != -1:
        keep = np.arange(n) < (n - svd_remove_n)
    else:
        eps = 1e-10 if epsilon < 0 else epsilon
        keep = s > eps
    w = np.where(keep, 1.0 / np.where(keep, s, 1.0), 0.0)
    return (U * w[None, :]) @ Vt, s, keep


def pinv_eigh(M: np.ndarray, svd_remove_n: int = -1, epsilon: float = 1e-10):
    """The same pseudo-inverse from LAPACK's symmetric eigen-decomposition
    (dsyevd): a second algorithm, whose distance to pinv_svd measures how far
    two correct FP64 algorithms drift on this M."""
    w, V = np.linalg.eigh(M)
    order = np.argsort(-np.abs(w), kind="stable")
    keep = np.zeros(len(w), bool)
    if svd_remove_n != -1:
        keep[order[: max(0, len(w) - svd_remove_n)]] = True
    else:
        keep = np.abs(w) > (1e-10 if epsilon < 0 else epsilon)
    return (V[:, keep] / w[keep]) @ V[:, keep].T


def svd_mode(jac, algorithm: int, svd_remove_n: int = -1, epsilon: float = 1e-10,
             use_eigh: bool = False):
    """(cam (n_cams, cp(cp+1)/2), pt (n_points, 6), iZ, singular values);
    use_eigh swaps the LAPACK SVD for the symmetric eigen-decomposition."""
    cp, nc, np_ = jac.cam_params, jac.num_cams, jac.num_points
    J = dense_jacobian(jac)
    M = J.T @ J
    iZ, s, keep = pinv_svd(M, algorithm, svd_remove_n, epsilon)
    if use_eigh:
        iZ = pinv_eigh(M, svd_remove_n, epsilon)
    iu = np.triu_indices(cp)
    cam = np.zeros((nc, cp * (cp + 1) // 2))
    Zc = np.zeros((nc, cp, cp))
    for a in range(nc):
        B = iZ[a * cp:(a + 1) * cp, a * cp:(a + 1) * cp]
        Zc[a] = B
        cam[a] = (B + B.T / 2)[iu]
    N = nc * cp
    rl = cp + 3
    nobs = jac.num_rows // 2
    V = np.zeros((np_, 3, 3))
    S = np.zeros((np_, 3, 3))
    vals = np.asarray(jac.vals).reshape(nobs, 2, rl)
    cols = np.asarray(jac.cols).reshape(nobs, 2, rl)
    for o in range(nobs):
        c = cols[o, 0, 0] // cp
        p = (cols[o, 0, cp] - N) // 3
        Jc, Jp = vals[o, :, :cp], vals[o, :, cp:]
        V[p] += Jp.T @ Jp
        W = Jc.T @ Jp
        S[p] += W.T @ Zc[c] @ W
    Vi = np.linalg.inv(V)
    P = Vi + Vi @ S @ Vi
    pt = P[:, [0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2]]
    return cam, pt, iZ, s
