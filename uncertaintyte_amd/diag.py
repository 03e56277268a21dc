"""Kernel-level diagnostics: ctypes binding of `_build/libuncertainty_amd_diag.so`
(declared in csrc/diag.h), a separate build of the same kernels used by the
parity tests to pin one kernel at a time.  The product library
(`libuncertainty_amd.so`) does not export these.  Device 0, host arrays,
numpy [row, col] convention."""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

from . import HERE, UncError, lib as _product_lib

DIAG_PATH = os.path.join(HERE, "_build", "libuncertainty_amd_diag.so")

_dlib = None


def dlib():
    global _dlib
    if _dlib is not None:
        return _dlib
    if not os.path.exists(DIAG_PATH):
        raise RuntimeError(f"{DIAG_PATH} is missing: build it with `python -m uncertaintyte_amd.build`")
    _product_lib()  # torch's HIP runtime first, as for the product library
    L = ctypes.CDLL(DIAG_PATH)
    vp, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)
    i = ctypes.c_int
    L.unc_diag_gemm_f64.argtypes = [i, i, i, ctypes.c_double, vp, i, vp, i, ctypes.c_double, vp, i, dp]
    L.unc_diag_gemm_f64_split.argtypes = [i, i, i, ctypes.c_double, vp, i, vp, i, vp, i, dp]
    L.unc_diag_gemm_f64_sym.argtypes = [i, ctypes.c_double, vp, vp, vp, dp, i]
    L.unc_diag_gemm_f64_sym_batch.argtypes = [i, i, vp, vp, vp, dp, i, dp]
    L.unc_diag_inverse.argtypes = [i, vp, vp, i]
    L.unc_diag_residual.argtypes = [i, vp, vp, vp, i]
    L.unc_diag_residual_crt.argtypes = [i, vp, vp, vp, i, i, dp]
    L.unc_diag_chol_inverse.argtypes = [i, vp, vp, vp]
    L.unc_diag_gemm_ex.argtypes = [i, i, i, ctypes.c_double, vp, i, vp, i, i, ctypes.c_double,
                                   vp, i, i, i, i, dp]
    L.unc_diag_gemm_bench.argtypes = [i, i, i, i, dp]
    L.unc_diag_residual_bench.argtypes = [i, i, i, dp]
    L.unc_diag_leaf_bench.argtypes = [i, i, dp]
    _dlib = L
    return L


def _check(rc: int, what: str):
    if rc != 0:
        msg = _product_lib().unc_error_string(rc).decode()
        raise UncError(rc, f"{what}: {msg}")


def gemm(A: np.ndarray, B: np.ndarray, C: Optional[np.ndarray] = None, alpha=1.0, beta=0.0,
         split: bool = False):
    """(C, max|C|) with C = alpha A B + beta C on the device; split=True takes the
    split-K form (beta must be 0; UncError UNC_ERR_UNSUPPORTED where the library
    would not split the shape)."""
    M, K = A.shape
    K2, N = B.shape
    assert K == K2
    Af = np.asfortranarray(A, dtype=np.float64)
    Bf = np.asfortranarray(B, dtype=np.float64)
    Cf = np.asfortranarray(np.zeros((M, N)) if C is None else C, dtype=np.float64).copy(order="F")
    m = ctypes.c_double()
    if split:
        if beta != 0.0:
            raise ValueError("the split-K form computes C = alpha A B")
        _check(dlib().unc_diag_gemm_f64_split(M, N, K, alpha, Af.ctypes.data, M, Bf.ctypes.data, K,
                                              Cf.ctypes.data, M, ctypes.byref(m)),
               "unc_diag_gemm_f64_split")
    else:
        _check(dlib().unc_diag_gemm_f64(M, N, K, alpha, Af.ctypes.data, M, Bf.ctypes.data, K, beta,
                                        Cf.ctypes.data, M, ctypes.byref(m)), "unc_diag_gemm_f64")
    return np.asarray(Cf), m.value


def gemm_sym(A: np.ndarray, B: np.ndarray, alpha: float = 1.0, parts: int = 0):
    """(C, max|C|) by the symmetric-product form (lower tiles, mirrored) with
    `parts` K chunks (0: the library's rule)."""
    n = A.shape[0]
    assert A.shape == (n, n) and B.shape == (n, n)
    Af = np.asfortranarray(A, dtype=np.float64)
    Bf = np.asfortranarray(B, dtype=np.float64)
    Cf = np.zeros((n, n), order="F")
    m = ctypes.c_double()
    _check(dlib().unc_diag_gemm_f64_sym(n, alpha, Af.ctypes.data, Bf.ctypes.data, Cf.ctypes.data,
                                        ctypes.byref(m), parts), "unc_diag_gemm_f64_sym")
    return np.asarray(Cf), m.value


def gemm_sym_batch(A: np.ndarray, Bs, reps: int = 0):
    """([C_b], [max|C_b|], avg_ms) with C_b = A B_b for up to four symmetric B_b
    by the Taylor chain's batched lower-tile kernel (one launch); avg_ms over
    `reps` further launches (None when reps == 0)."""
    n = A.shape[0]
    nb = len(Bs)
    Af = np.asfortranarray(A, dtype=np.float64)
    Bf = np.concatenate([np.asfortranarray(B, dtype=np.float64).reshape(-1, order="F") for B in Bs])
    Cf = np.zeros(nb * n * n)
    mx = (ctypes.c_double * nb)()
    ms = ctypes.c_double()
    _check(dlib().unc_diag_gemm_f64_sym_batch(n, nb, Af.ctypes.data, Bf.ctypes.data, Cf.ctypes.data,
                                              mx, reps, ctypes.byref(ms) if reps else None),
           "unc_diag_gemm_f64_sym_batch")
    Cs = [Cf[b * n * n:(b + 1) * n * n].reshape(n, n, order="F") for b in range(nb)]
    return Cs, list(mx), (ms.value if reps else None)


GEMM_A_LOWER, GEMM_B_LOWER, GEMM_BT_LOWER, GEMM_C_LOWER = 1, 2, 4, 8


def gemm_ex(A: np.ndarray, B: np.ndarray, C: Optional[np.ndarray] = None, alpha=1.0, beta=0.0,
            trans_b: bool = False, tri: int = 0, nmod: int = 0, reps: int = 0):
    """(C, avg_ms) with C = alpha A op(B) + beta C (op(B) = B.T when trans_b) under the
    operand flags `tri` (GEMM_*): nmod = 0 takes the FP64 MFMA product, 6..20 the int8
    CRT product with that many moduli; avg_ms of `reps` more calls (None if
    reps = 0)."""
    M, K = A.shape
    N = B.shape[0] if trans_b else B.shape[1]
    assert (B.shape[1] if trans_b else B.shape[0]) == K
    Af = np.asfortranarray(A, dtype=np.float64)
    Bf = np.asfortranarray(B, dtype=np.float64)
    Cf = np.asfortranarray(np.zeros((M, N)) if C is None else C, dtype=np.float64).copy(order="F")
    ms = ctypes.c_double(0.0)
    _check(dlib().unc_diag_gemm_ex(M, N, K, alpha, Af.ctypes.data, M, Bf.ctypes.data,
                                   Bf.shape[0], int(trans_b), beta, Cf.ctypes.data, M, tri, nmod,
                                   reps, ctypes.byref(ms) if reps else None),
           "unc_diag_gemm_ex")
    return np.ascontiguousarray(Cf), (ms.value if reps else None)


def inverse(A: np.ndarray, refine: bool = True) -> np.ndarray:
    n = A.shape[0]
    Af = np.asfortranarray(A, dtype=np.float64)
    X = np.zeros((n, n), order="F")
    _check(dlib().unc_diag_inverse(n, Af.ctypes.data, X.ctypes.data, 1 if refine else 0),
           "unc_diag_inverse")
    return np.asarray(X)


def chol_inverse(A: np.ndarray):
    """(L, X): X = L^-1 for the Cholesky factor L of SPD A; L holds the factor in
    its diagonal leaf blocks only (the recursion keeps L21 in X's block)."""
    n = A.shape[0]
    Af = np.asfortranarray(A, dtype=np.float64)
    L = np.zeros((n, n), order="F")
    X = np.zeros((n, n), order="F")
    _check(dlib().unc_diag_chol_inverse(n, Af.ctypes.data, L.ctypes.data, X.ctypes.data),
           "unc_diag_chol_inverse")
    return np.tril(np.asarray(L)), np.asarray(X)


def residual(A: np.ndarray, X: np.ndarray, tile_cols: int = 0) -> np.ndarray:
    """R = I - A X with the refinement's residual kernel (error-free extraction);
    tile_cols 16 / 32 forces the 64 x 16 / 64 x 32 tiles."""
    n = A.shape[0]
    Af = np.asfortranarray(A, dtype=np.float64)
    Xf = np.asfortranarray(X, dtype=np.float64)
    R = np.zeros((n, n), order="F")
    _check(dlib().unc_diag_residual(n, Af.ctypes.data, Xf.ctypes.data, R.ctypes.data, tile_cols),
           "unc_diag_residual")
    return np.asarray(R)


def residual_crt(A: np.ndarray, X: np.ndarray, blocks: int = 1, reps: int = 0):
    """(R, avg_ms) with R = I - A X for symmetric A by the int8 CRT residual (A in
    `blocks` row blocks); avg_ms over `reps` further calls (None when reps == 0)."""
    n = A.shape[0]
    Af = np.asfortranarray(A, dtype=np.float64)
    Xf = np.asfortranarray(X, dtype=np.float64)
    R = np.zeros((n, n), order="F")
    ms = ctypes.c_double()
    _check(dlib().unc_diag_residual_crt(n, Af.ctypes.data, Xf.ctypes.data, R.ctypes.data,
                                        blocks, reps, ctypes.byref(ms) if reps else None),
           "unc_diag_residual_crt")
    return np.asarray(R), (ms.value if reps else None)


def gemm_bench(n: int, reps: int = 10, form: int = 0, parts: int = 0) -> float:
    """Average ms of one n^3 product: form 0 plain, 1 split-K, 2 symmetric."""
    ms = ctypes.c_double()
    _check(dlib().unc_diag_gemm_bench(n, reps, form, parts, ctypes.byref(ms)), "unc_diag_gemm_bench")
    return ms.value


def residual_bench(n: int, reps: int = 5, tile_cols: int = 0) -> float:
    ms = ctypes.c_double()
    _check(dlib().unc_diag_residual_bench(n, reps, tile_cols, ctypes.byref(ms)),
           "unc_diag_residual_bench")
    return ms.value


def leaf_bench(reps: int = 100, variant: int = 0) -> float:
    us = ctypes.c_double()
    _check(dlib().unc_diag_leaf_bench(reps, variant, ctypes.byref(us)), "unc_diag_leaf_bench")
    return us.value
