"""ctypes bindings for the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by `__graft_entry__.smoke()` and
by the cpu_baseline leg of bench.py, to compare results against.  The product package
(uncertaintyte_amd) never imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

ORC_OK, ORC_ERR_BAD_INPUT, ORC_ERR_NO_FIXED_POINTS, ORC_ERR_SINGULAR, ORC_ERR_OOM = range(5)
ORC_ERR_UNSUPPORTED = 6
INV_REFINED, INV_REFERENCE = 0, 1

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class Opts(ctypes.Structure):
    _fields_ = [("inverse_mode", ctypes.c_int), ("nthreads", ctypes.c_int),
                ("lambda_", ctypes.c_double)]


class Stats(ctypes.Structure):
    _fields_ = [("lambda_", ctypes.c_double), ("n_terms", ctypes.c_int),
                ("cycle_change", ctypes.c_double * 20), ("fixed", ctypes.c_int * 3),
                ("n_obs_kept", ctypes.c_int), ("n_pts_compact", ctypes.c_int)]


class Debug(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in ("U_s", "Vinv_s", "d", "Z", "X0", "X")]


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-C", HERE, "-s"] + (["-B"] if force else []), check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_fixpts.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, _i32p, _f64p, ctypes.c_int,
                                 ctypes.c_int, _i32p, ctypes.POINTER(ctypes.c_int),
                                 ctypes.POINTER(ctypes.c_int), _i32p, _i32p, _f64p]
        L.orc_pipeline.argtypes = [ctypes.POINTER(Opts), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   _i32p, ctypes.c_int, ctypes.c_int, _i32p, _i32p, _f64p,
                                   _f64p, _f64p, ctypes.POINTER(Stats), ctypes.POINTER(Debug)]
        L.orc_taylor.argtypes = [ctypes.POINTER(Opts), ctypes.c_int, _f64p, ctypes.c_double,
                                 ctypes.c_void_p, _f64p, ctypes.POINTER(Stats)]
        L.orc_points_from_X.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p,
                                        ctypes.c_int, _i32p, _f64p, _f64p, _f64p]
        L.orc_points_from_blocks.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p,
                                             ctypes.c_int, _i32p, _f64p, _f64p,
                                             ctypes.c_char_p, _f64p, ctypes.c_char_p]
        L.orc_jtj_dense.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, _i32p, _f64p, _f64p]
        L.orc_inv_lu.argtypes = [ctypes.c_int, _f64p, _f64p]
        L.orc_inv_equilibrated.argtypes = [ctypes.c_int, _f64p, _f64p]
        L.orc_residual_dd.argtypes = [ctypes.c_int, _f64p, _f64p, _f64p]
        L.orc_gemm.argtypes = [ctypes.c_int, _f64p, _f64p, _f64p]
        L.orc_front.argtypes = [ctypes.POINTER(Opts), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                _i32p, ctypes.c_int, ctypes.c_int, _i32p, _i32p, _f64p, _f64p,
                                _f64p]
        L.orc_factorial32.argtypes = [ctypes.c_int]
        L.orc_factorial32.restype = ctypes.c_int
        L.orc_auto_lambda.argtypes = [ctypes.c_int]
        L.orc_auto_lambda.restype = ctypes.c_double
        _lib = L
    return _lib


class OracleError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__(f"{what}: oracle error {code}")
        self.code = code


def fixpts(jac, fix=None):
    """fixPts restated; returns (nrows, ncols, rows, cols, vals, fix_after_sort)."""
    pts = np.array(fix if fix is not None else jac.fix_array(), dtype=np.int32).copy()
    nnz = int(jac.rows[-1])
    orow = np.zeros(jac.num_rows + 1, np.int32)
    ocol = np.zeros(max(nnz, 1), np.int32)
    oval = np.zeros(max(nnz, 1), np.float64)
    nr, nc = ctypes.c_int(), ctypes.c_int()
    rc = lib().orc_fixpts(jac.num_rows, jac.num_cols, np.ascontiguousarray(jac.rows, np.int32),
                          np.ascontiguousarray(jac.cols, np.int32),
                          np.ascontiguousarray(jac.vals, np.float64), jac.cam_params,
                          jac.num_cams, pts, ctypes.byref(nr), ctypes.byref(nc), orow, ocol, oval)
    if rc:
        raise OracleError(rc, "orc_fixpts")
    n = nr.value
    m = n * (jac.cam_params + 3)
    return n, nc.value, orow[:n + 1].copy(), ocol[:m].copy(), oval[:m].copy(), tuple(pts)


def pipeline(jac, inverse_mode: int = INV_REFINED, nthreads: int = 0, lam: Optional[float] = None,
             fix=None, debug: bool = False):
    """Run the whole TAYLOR_EXPANSION restatement.

    Returns dict(cam=(nC, per), pt=(nP, 6), lambda, n_terms, cycle_change, fixed,
    and, with debug=True, U_s, Vinv_s, d, Z, X0, X).
    """
    L = lib()
    nc, cp, npts = jac.num_cams, jac.cam_params, jac.num_points
    per = cp * (cp + 1) // 2
    N = nc * cp
    cam = np.zeros(nc * per, np.float64)
    pt = np.zeros(npts * 6, np.float64)
    st = Stats()
    o = Opts(inverse_mode, nthreads, float(jac.lam if lam is None else lam))
    fx = np.array(fix if fix is not None else jac.fix_array(), dtype=np.int32)
    dbg = None
    arrs = {}
    if debug:
        npc = npts - 3
        arrs = dict(U_s=np.zeros(nc * cp * cp), Vinv_s=np.zeros(npc * 9), d=np.zeros(N + 3 * npc),
                    Z=np.zeros(N * N), X0=np.zeros(N * N), X=np.zeros(N * N))
        dbg = Debug(**{k: v.ctypes.data for k, v in arrs.items()})
    rc = L.orc_pipeline(ctypes.byref(o), nc, cp, npts, fx, jac.num_rows, jac.num_cols,
                        np.ascontiguousarray(jac.rows, np.int32),
                        np.ascontiguousarray(jac.cols, np.int32),
                        np.ascontiguousarray(jac.vals, np.float64), cam, pt, ctypes.byref(st),
                        ctypes.byref(dbg) if dbg is not None else None)
    if rc:
        raise OracleError(rc, "orc_pipeline")
    out = dict(cam=cam.reshape(nc, per), pt=pt.reshape(npts, 6), lambda_=st.lambda_,
               n_terms=st.n_terms, cycle_change=list(st.cycle_change[:st.n_terms]),
               fixed=tuple(st.fixed), n_obs_kept=st.n_obs_kept)
    if debug:
        out["U_s"] = arrs["U_s"].reshape(nc, cp, cp)
        out["Vinv_s"] = arrs["Vinv_s"].reshape(-1, 3, 3)
        out["d"] = arrs["d"]
        # column-major N x N -> numpy arrays indexed [row, col]
        for k in ("Z", "X0", "X"):
            out[k] = arrs[k].reshape(N, N).T.copy()
    return out


def taylor(Z: np.ndarray, lam: float, inverse_mode: int = INV_REFINED, nthreads: int = 0):
    """Taylor chain on a given Z ([row, col] numpy) -> (X0, X, stats)."""
    N = Z.shape[0]
    zc = np.ascontiguousarray(Z.T).reshape(-1)
    X0 = np.zeros(N * N)
    X = np.zeros(N * N)
    st = Stats()
    o = Opts(inverse_mode, nthreads, lam)
    rc = lib().orc_taylor(ctypes.byref(o), N, zc, lam, X0.ctypes.data, X, ctypes.byref(st))
    if rc:
        raise OracleError(rc, "orc_taylor")
    return X0.reshape(N, N).T.copy(), X.reshape(N, N).T.copy(), \
        dict(n_terms=st.n_terms, cycle_change=list(st.cycle_change[:st.n_terms]))


def jtj_dense(nrows, ncols, rows, cols, vals) -> np.ndarray:
    out = np.zeros(ncols * ncols)
    lib().orc_jtj_dense(nrows, ncols, np.ascontiguousarray(rows, np.int32),
                        np.ascontiguousarray(cols, np.int32),
                        np.ascontiguousarray(vals, np.float64), out)
    return out.reshape(ncols, ncols).T.copy()


def _square(fn, A: np.ndarray) -> np.ndarray:
    n = A.shape[0]
    X = np.zeros(n * n)
    rc = fn(n, np.ascontiguousarray(A.T).reshape(-1), X)
    if rc:
        raise OracleError(rc, fn.__name__)
    return X.reshape(n, n).T.copy()


def inv_lu(A):
    return _square(lib().orc_inv_lu, A)


def inv_equilibrated(A):
    return _square(lib().orc_inv_equilibrated, A)


def residual_dd(A, X):
    n = A.shape[0]
    R = np.zeros(n * n)
    lib().orc_residual_dd(n, np.ascontiguousarray(A.T).reshape(-1),
                          np.ascontiguousarray(X.T).reshape(-1), R)
    return R.reshape(n, n).T.copy()


def gemm(A, B):
    n = A.shape[0]
    C = np.zeros(n * n)
    lib().orc_gemm(n, np.ascontiguousarray(A.T).reshape(-1), np.ascontiguousarray(B.T).reshape(-1), C)
    return C.reshape(n, n).T.copy()


def factorial32(n: int) -> int:
    return lib().orc_factorial32(n)


def auto_lambda(N: int) -> float:
    return lib().orc_auto_lambda(N)


def front(jac, fix=None, nthreads: int = 0):
    """fixPts + assembly + Schur only: (Z [row, col], d) -- the scaled Schur
    complement the TE path inverts and the column scales."""
    nc, cp, npts = jac.num_cams, jac.cam_params, jac.num_points
    N = nc * cp
    Z = np.zeros(N * N)
    d = np.zeros(N + 3 * (npts - 3))
    fx = np.array(fix if fix is not None else jac.fix_array(), dtype=np.int32)
    o = Opts(INV_REFINED, nthreads, -1.0)
    rc = lib().orc_front(ctypes.byref(o), nc, cp, npts, fx, jac.num_rows, jac.num_cols,
                         np.ascontiguousarray(jac.rows, np.int32),
                         np.ascontiguousarray(jac.cols, np.int32),
                         np.ascontiguousarray(jac.vals, np.float64), Z, d)
    if rc:
        raise OracleError(rc, "orc_front")
    return Z.reshape(N, N).T, d


def zplus_reference(jac, fix=None, nthreads: int = 0):
    """FP64 gauge-fixed covariance: X = pinv_7(Z_s) (LAPACK eigen-decomposition
    of the oracle's Schur complement, the 7 smallest eigenvalues dropped), camera
    output = upper triangles of D X D per camera, points propagated from X's
    diagonal blocks with the same formula as the TAYLOR path.
    Returns dict(cam, pt, X, eig) -- the reference for the Z^+ precision modes."""
    Z, d = front(jac, fix=fix, nthreads=nthreads)
    w, Q = np.linalg.eigh(Z)
    inv = np.zeros_like(w)
    inv[7:] = 1.0 / w[7:]
    X = (Q * inv) @ Q.T
    cp, nc = jac.cam_params, jac.num_cams
    iu = np.triu_indices(cp)
    cam = np.empty((nc, cp * (cp + 1) // 2))
    for a in range(nc):
        s = slice(a * cp, (a + 1) * cp)
        B = X[s, s] * d[s][:, None] * d[s][None, :]
        cam[a] = B[iu]
    nr, ncl, rows, cols, vals, fx = fixpts(jac, fix)
    pt = np.zeros(jac.num_points * 6)
    Xc = np.ascontiguousarray(X.T).reshape(-1)  # column-major
    rc = lib().orc_points_from_X(nc, cp, jac.num_points,
                                 np.array(fix if fix is not None else jac.fix_array(), np.int32),
                                 nr, np.ascontiguousarray(cols, np.int32),
                                 np.ascontiguousarray(vals, np.float64), Xc, pt)
    if rc:
        raise OracleError(rc, "orc_points_from_X")
    return dict(cam=cam, pt=pt.reshape(-1, 6), X=X, eig=w)


def points_from_blocks(jac, blocks: dict, fix=None):
    """Point covariances (the TAYLOR point step, compute.cpp:516-527) from the
    camera blocks X_cc of the scaled problem given in `blocks` {camera: cp x cp},
    for the points all of whose observing cameras are in `blocks`.
    Returns (pt (nP, 6), done mask (nP,) of the points computed)."""
    nc, cp = jac.num_cams, jac.cam_params
    Xd = np.zeros((nc, cp, cp))
    ok = np.zeros(nc, np.uint8)
    for a, B in blocks.items():
        Xd[a] = np.asarray(B).T  # column-major per block
        ok[a] = 1
    nr, ncl, rows, cols, vals, fx = fixpts(jac, fix)
    pt = np.zeros(jac.num_points * 6)
    done = np.zeros(jac.num_points, np.uint8)
    rc = lib().orc_points_from_blocks(nc, cp, jac.num_points,
                                      np.array(fix if fix is not None else jac.fix_array(), np.int32),
                                      nr, np.ascontiguousarray(cols, np.int32),
                                      np.ascontiguousarray(vals, np.float64),
                                      np.ascontiguousarray(Xd.reshape(-1)), ok.tobytes(), pt,
                                      done.ctypes.data_as(ctypes.c_char_p))
    if rc:
        raise OracleError(rc, "orc_points_from_blocks")
    return pt.reshape(-1, 6), done.astype(bool)


def zplus_sampled(jac, cams, fix=None, nthreads: int = 0, log=None, blocks: bool = False):
    """Camera outputs of D pinv_7(Z_s) D for the listed cameras only, for scenes
    too large for a dense eigen-decomposition.  Uses the identity
    pinv_7(Z) = (Z + N N^T)^-1 - N N^T with N the eigenvectors of the 7 smallest
    eigenvalues, found by shift-invert Lanczos (ARPACK) on a LAPACK Cholesky of
    Z + sigma I; the sampled columns come from a Cholesky solve with Z + N N^T.
    Returns (cam rows for `cams`, the 7 smallest eigenvalues), and with
    blocks=True also {camera: cp x cp block of pinv_7(Z_s)} (scaled, for
    points_from_blocks).  `log(msg)` is called between the phases (long scenes
    take minutes)."""
    import scipy.linalg as sla
    from scipy.sparse.linalg import LinearOperator, eigsh
    log = log or (lambda msg: None)
    Z, d = front(jac, fix=fix, nthreads=nthreads)
    log("oracle: Schur complement done")
    n = Z.shape[0]
    cp = jac.cam_params
    sigma = 1e-10 * np.trace(Z) / n
    A = np.array(Z, order="F")
    A[np.diag_indices(n)] += sigma
    cf = sla.cho_factor(A, lower=True, overwrite_a=True, check_finite=False)
    log("oracle: shifted Cholesky done")
    op = LinearOperator((n, n), matvec=lambda v: sla.cho_solve(cf, v, check_finite=False),
                        dtype=np.float64)
    w, Nb = eigsh(Z, k=7, sigma=-sigma, OPinv=op, which="LM", tol=1e-14,
                  v0=np.linspace(1.0, 2.0, n))
    del cf, A
    log("oracle: gauge basis done")
    G = np.array(Z, order="F")
    G += Nb @ Nb.T
    cg = sla.cho_factor(G, lower=True, overwrite_a=True, check_finite=False)
    log("oracle: gauge-fixed Cholesky done")
    idx = np.concatenate([np.arange(a * cp, (a + 1) * cp) for a in cams])
    E = np.zeros((n, len(idx)))
    E[idx, np.arange(len(idx))] = 1.0
    cols = sla.cho_solve(cg, E, check_finite=False) - Nb @ Nb[idx].T
    iu = np.triu_indices(cp)
    out = np.empty((len(cams), cp * (cp + 1) // 2))
    blk = {}
    for q, a in enumerate(cams):
        s = slice(a * cp, (a + 1) * cp)
        X = cols[s, q * cp:(q + 1) * cp]
        X = 0.5 * (X + X.T)
        blk[int(a)] = X
        B = X * d[s][:, None] * d[s][None, :]
        out[q] = B[iu]
    if blocks:
        return out, np.sort(w), blk
    return out, np.sort(w)
