"""CPU restatement of the cov::Uncertainty accessors
(src/uncertaintyTE/uncertainty.cpp:48-185).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the reference results for the
device eigen kernels (unc_eig_cameras / unc_eig_points); the product package never
imports it.

Pinning:
  - unpack_cameras / unpack_points and the eigenvalue functions are checked
    against the arrays of the reference's own accessor tests
    (src/test_eigValues/covarianceEigenValues_test.cpp:81-137, fixture
    tests/golden/eig_accessors_daliborka.json) at the reference's tolerances.
  - The eigenvalue functions follow the reference's method (real parts of the
    general eigenvalues, uncertainty.cpp:84-86 and :155-157) and are also held
    to LAPACK's symmetric solver on the same blocks.
  - center_eigvecs / point_eigvecs: the reference returns Eigen's EigenSolver
    vectors in EigenSolver's own order and signs (uncertainty.cpp:100-113,
    :173-183); Eigen is absent here, so that order and those signs are parity
    unpinned.  These functions return the convention of the device kernels
    (vectors by descending eigenvalue, largest-magnitude component positive),
    and the tests pin the vectors themselves by the eigen equation A v = l v
    and orthonormality.
"""
from __future__ import annotations

import numpy as np


def unpack_cameras(cam: np.ndarray, cp: int) -> np.ndarray:
    """uncertainty.cpp:48-72: row-major packed upper triangle -> (n, cp, cp)."""
    per = cp * (cp + 1) // 2
    cam = np.asarray(cam, np.float64).reshape(-1, per)
    out = np.zeros((cam.shape[0], cp, cp))
    t = 0
    for i in range(cp):
        for j in range(i, cp):
            out[:, i, j] = cam[:, t]
            out[:, j, i] = cam[:, t]
            t += 1
    return out


def unpack_points(pt: np.ndarray) -> np.ndarray:
    """uncertainty.cpp:120-144: (xx, xy, xz, yy, yz, zz) -> (n, 3, 3)."""
    p = np.asarray(pt, np.float64).reshape(-1, 6)
    idx = [[0, 1, 2], [1, 3, 4], [2, 4, 5]]
    return p[:, np.array(idx)]


def camera_eigenvalues(cam: np.ndarray, cp: int) -> np.ndarray:
    """uncertainty.cpp:74-89: real parts of the eigenvalues, descending."""
    ev = np.linalg.eigvals(unpack_cameras(cam, cp)).real
    return -np.sort(-ev, axis=1)


def point_eigenvalues(pt: np.ndarray) -> np.ndarray:
    """uncertainty.cpp:146-162: real parts of the eigenvalues, descending."""
    ev = np.linalg.eigvals(unpack_points(pt)).real
    return -np.sort(-ev, axis=1)


def _vectors_convention(M: np.ndarray) -> np.ndarray:
    """(n, 9): eigenvectors of symmetric 3x3 blocks by descending eigenvalue,
    each with its largest-magnitude component positive, vector k at [3k:3k+3]."""
    w, V = np.linalg.eigh(M)
    V = V[:, :, ::-1]
    k = np.argmax(np.abs(V), axis=1)
    sg = np.sign(np.take_along_axis(V, k[:, None, :], axis=1))
    sg[sg == 0] = 1.0
    V = V * sg
    return np.transpose(V, (0, 2, 1)).reshape(-1, 9)


def center_eigvecs(cam: np.ndarray, cp: int) -> np.ndarray:
    """uncertainty.cpp:91-115: eigenvectors of A(3:5, 3:5) of each camera."""
    return _vectors_convention(unpack_cameras(cam, cp)[:, 3:6, 3:6])


def point_eigvecs(pt: np.ndarray) -> np.ndarray:
    """uncertainty.cpp:164-185: eigenvectors of each point block."""
    return _vectors_convention(unpack_points(pt))
