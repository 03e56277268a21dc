"""MI355X (gfx950) camera + point covariance pipeline of uncertaintyTE.

Python host mirror of the reference interface.  The work runs in the HIP
library `_build/libuncertainty_amd.so` (C ABI: include/uncertainty_amd.h); this
module binds it with ctypes and raises if the library is missing -- there is
no CPU fallback.

    from uncertaintyte_amd import jacob, compute_covariances
    J = jacob.read_jacob("scene.jacob")
    res = compute_covariances(J)          # res.cam (nC, 45), res.pt (nP, 6)

Mirrors:
  compute_covariances  <- computeCovariances / getCovariances
                          (compute.cpp:458-540, include/uncertainty.h:48-61)
  Uncertainty          <- cov::Uncertainty accessors (uncertainty.cpp:36-185)
"""
from __future__ import annotations

import ctypes
import dataclasses
import os
from typing import Optional

import numpy as np

from . import jacob  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libuncertainty_amd.so")
CLI_PATH = os.path.join(HERE, "_build", "uncertainty")

UNC_OK = 0
UNC_ERR_BAD_INPUT = 1
UNC_ERR_NO_FIXED_POINTS = 2
UNC_ERR_SINGULAR = 3
UNC_ERR_OOM = 4
UNC_ERR_DEVICE = 5
UNC_ERR_UNSUPPORTED = 6
UNC_ERR_NUMERIC = 7
UNC_ERR_TOO_LARGE = 8
NOALLOC_SCRATCH_BYTES = 64 << 20
PREC_FP64, PREC_MIXED, PREC_ZPLUS_FP64 = 0, 1, 2
SVD_QR_ITERATION, SVD_DEVIDE_AND_CONQUER, TAYLOR_EXPANSION = 0, 1, 2


class Options(ctypes.Structure):
    _fields_ = [("algorithm", ctypes.c_int), ("epsilon", ctypes.c_double),
                ("lambda_", ctypes.c_double), ("num_cams", ctypes.c_int),
                ("cam_params", ctypes.c_int), ("num_points", ctypes.c_int),
                ("num_obs", ctypes.c_int), ("svd_remove_n", ctypes.c_int),
                ("max_iter_te", ctypes.c_int), ("pts2fix", ctypes.c_int * 3),
                ("precision", ctypes.c_int), ("refine", ctypes.c_int),
                ("device", ctypes.c_int), ("flags", ctypes.c_int), ("crt_moduli", ctypes.c_int)]


class Stats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "time_create_j", "time_fix_j", "time_norm_j", "time_multiply_jj", "time_split_jj",
        "time_inv_v", "time_compose_z", "time_inv_z", "time_te", "time_pts_unc", "time_all",
        "lambda_")] + [("fixed_pts", ctypes.c_int * 3), ("num_terms", ctypes.c_int),
                       ("cycle_change", ctypes.c_double * 20), ("gemm_ms", ctypes.c_double),
                       ("gemm_flops", ctypes.c_double), ("gemm_count", ctypes.c_int),
                       ("num_obs_kept", ctypes.c_int), ("stage_ms", ctypes.c_double * 12),
                       ("batch_ms", ctypes.c_double), ("batch_flops", ctypes.c_double),
                       ("batch_count", ctypes.c_int), ("zplus_moduli", ctypes.c_int),
                       ("zplus_cond_lb", ctypes.c_double)]


class Debug(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in ("U_s", "Vinv_s", "d", "Z", "X0", "X")]


class UncError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{msg} (code {code})")
        self.code = code


_lib = None


def lib():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m uncertaintyte_amd.build`")
    # One HIP runtime per process: when torch is installed it ships its own
    # libamdhip64.so.7, and whichever copy is loaded first serves the process.
    # Loading torch first keeps torch usable next to this library.
    try:
        import torch  # noqa: F401
    except Exception:
        pass
    L = ctypes.CDLL(LIB_PATH)
    vp, ip, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
    L.unc_default_options.argtypes = [ctypes.POINTER(Options)]
    L.unc_default_options.restype = None
    L.unc_context_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
    L.unc_context_destroy.argtypes = [vp]
    L.unc_context_destroy.restype = None
    L.unc_compute.argtypes = [vp, ctypes.POINTER(Options), ctypes.c_int, ctypes.c_int, vp, vp, vp,
                              vp, vp, ctypes.POINTER(Stats)]
    L.unc_compute_device.argtypes = [vp, ctypes.POINTER(Options), ctypes.c_int, ctypes.c_int, vp,
                                     vp, vp, vp, vp, ctypes.POINTER(Stats), vp]
    L.unc_compute_device_noalloc.argtypes = [vp, ctypes.POINTER(Options), ctypes.c_int,
                                             ctypes.c_int, vp, vp, vp, vp, vp,
                                             ctypes.POINTER(Stats), vp]
    L.unc_compute_debug.argtypes = [vp, ctypes.POINTER(Options), ctypes.c_int, ctypes.c_int, vp, vp,
                                    vp, vp, vp, ctypes.POINTER(Stats), ctypes.POINTER(Debug)]
    L.unc_workspace_bytes.argtypes = [ctypes.POINTER(Options), ctypes.c_int]
    L.unc_workspace_bytes.restype = ctypes.c_ulonglong
    L.unc_error_string.argtypes = [ctypes.c_int]
    L.unc_error_string.restype = ctypes.c_char_p
    L.unc_stage_name.argtypes = [ctypes.c_int]
    L.unc_stage_name.restype = ctypes.c_char_p
    L.unc_abi_version.restype = ctypes.c_int
    L.unc_host_alloc.argtypes = [ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)]
    L.unc_host_free.argtypes = [ctypes.c_void_p]
    L.unc_host_free.restype = None
    L.unc_server_default_config.argtypes = [ctypes.POINTER(ServerConfig)]
    L.unc_server_default_config.restype = None
    L.unc_server_create.argtypes = [ctypes.POINTER(ServerConfig), ctypes.POINTER(ctypes.c_void_p)]
    L.unc_server_submit.argtypes = [ctypes.c_void_p, ctypes.POINTER(Options), ctypes.c_int,
                                    ctypes.c_int, vp, vp, vp, vp, vp, ctypes.POINTER(Stats),
                                    ctypes.POINTER(ctypes.c_longlong)]
    L.unc_server_wait.argtypes = [ctypes.c_void_p, ctypes.c_longlong, ctypes.POINTER(ctypes.c_int),
                                  ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    L.unc_server_query.argtypes = [ctypes.c_void_p, ctypes.POINTER(ServerStats)]
    L.unc_server_destroy.argtypes = [ctypes.c_void_p]
    L.unc_server_destroy.restype = None
    L.unc_device_lds_bytes.argtypes = [ctypes.c_int]
    L.unc_scene_read_colmap.argtypes = [ctypes.c_char_p, ctypes.POINTER(vp)]
    L.unc_scene_read_bal.argtypes = [ctypes.c_char_p, ctypes.POINTER(vp)]
    L.unc_scene_counts.argtypes = [vp, ip, ip, ip]
    L.unc_scene_arrays.argtypes = [vp, vp, vp, vp, vp]
    L.unc_scene_destroy.argtypes = [vp]
    L.unc_scene_destroy.restype = None
    L.unc_find_pts2fix.argtypes = [ctypes.c_int, vp, ip]
    L.unc_scene_jacobian.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     vp, vp, vp, vp, vp, vp, vp]
    L.unc_scene_jacobian_device.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, vp, vp, vp, vp, vp, vp, vp, vp]
    L.unc_eig_cameras.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, vp, vp]
    L.unc_eig_points.argtypes = [vp, ctypes.c_int, vp, vp, vp]
    L.unc_eig_cameras_device.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, vp, vp, vp]
    L.unc_eig_points_device.argtypes = [vp, ctypes.c_int, vp, vp, vp, vp]
    _lib = L
    return L


def error_string(code: int) -> str:
    return lib().unc_error_string(code).decode()


def _check(code: int, what: str):
    if code != UNC_OK:
        raise UncError(code, f"{what}: {error_string(code)}")


class ServerConfig(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("slots", ctypes.c_int), ("hbm_budget", ctypes.c_size_t),
                ("bypass_ms", ctypes.c_double), ("flags", ctypes.c_int)]


class ServerStats(ctypes.Structure):
    _fields_ = [("queued", ctypes.c_longlong), ("completed", ctypes.c_longlong),
                ("admitted_out_of_order", ctypes.c_longlong), ("held_bytes", ctypes.c_size_t),
                ("peak_held_bytes", ctypes.c_size_t), ("max_concurrent", ctypes.c_int)]


class _PinnedBlock:
    """Owner of one unc_host_alloc block; freed with the last array viewing it."""

    def __init__(self, nbytes: int):
        self.ptr = ctypes.c_void_p()
        # held: module globals (lib, ctypes) may be gone when this runs at exit
        self._free = lib().unc_host_free
        _check(lib().unc_host_alloc(max(nbytes, 1), ctypes.byref(self.ptr)), "unc_host_alloc")
        self.nbytes = nbytes

    def __del__(self):
        p, self.ptr = self.ptr, None
        if p:
            self._free(p)


def pinned_empty(shape, dtype) -> np.ndarray:
    """A numpy array in page-locked host memory (unc_host_alloc): uploads from it and
    downloads into it run at the DMA engines' rate."""
    dt = np.dtype(dtype)
    n = int(np.prod(shape)) if np.ndim(shape) else int(shape)
    blk = _PinnedBlock(n * dt.itemsize)
    buf = (ctypes.c_char * max(n * dt.itemsize, 1)).from_address(blk.ptr.value)
    buf._owner = blk  # keeps the block alive as long as any view of the buffer
    return np.frombuffer(buf, dtype=dt, count=n).reshape(shape)


def pinned_copy(a: np.ndarray) -> np.ndarray:
    out = pinned_empty(a.shape, a.dtype)
    out[...] = a
    return out


def make_options(jac, precision: int = PREC_FP64, refine: bool = True, device: int = 0,
                 fix=None, lam: Optional[float] = None, algorithm: int = TAYLOR_EXPANSION,
                 svd_remove_n: int = -1, epsilon: float = 1e-10, moduli: int = 0) -> Options:
    o = Options()
    lib().unc_default_options(ctypes.byref(o))
    o.algorithm = algorithm
    o.svd_remove_n = svd_remove_n
    o.epsilon = epsilon
    o.lambda_ = float(jac.lam if lam is None else lam)
    o.num_cams, o.cam_params, o.num_points, o.num_obs = (jac.num_cams, jac.cam_params,
                                                         jac.num_points, jac.num_obs)
    fx = fix if fix is not None else (jac.fix if jac.fix is not None else (-1, -1, -1))
    for i in range(3):
        o.pts2fix[i] = int(fx[i])
    o.precision = precision
    # True -> automatic number of Newton steps, False -> none, int -> that many
    o.refine = -1 if refine is True else (0 if refine is False else int(refine))
    o.device = device
    o.crt_moduli = moduli
    return o


@dataclasses.dataclass
class Result:
    cam: np.ndarray              # (num_cams, cp(cp+1)/2)
    pt: np.ndarray               # (num_points, 6)
    lambda_: float
    n_terms: int
    cycle_change: list
    fixed: tuple
    stats: Stats
    debug: Optional[dict] = None


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def compute_covariances(jac, precision: int = PREC_FP64, refine: bool = True, device: int = 0,
                        fix=None, lam: Optional[float] = None, debug: bool = False,
                        ctx=None, algorithm: int = TAYLOR_EXPANSION, svd_remove_n: int = -1,
                        epsilon: float = 1e-10, moduli: int = 0) -> Result:
    """Run the pipeline on a host-resident Jacobian (jacob.Jacobian).  algorithm
    SVD_QR_ITERATION / SVD_DEVIDE_AND_CONQUER selects the dense pseudo-inverse
    modes (small problems only) with svd_remove_n / epsilon as in cov::Options.
    moduli: PREC_MIXED's int8 CRT moduli (6..20, 0 = the library's default)."""
    L = lib()
    o = make_options(jac, precision, refine, device, fix, lam, algorithm, svd_remove_n, epsilon,
                     moduli)
    cp = jac.cam_params
    per = cp * (cp + 1) // 2
    cam = np.zeros(jac.num_cams * per, np.float64)
    pt = np.zeros(jac.num_points * 6, np.float64)
    rows = np.ascontiguousarray(jac.rows, np.int32)
    cols = np.ascontiguousarray(jac.cols, np.int32)
    vals = np.ascontiguousarray(jac.vals, np.float64)
    st = Stats()
    dbg_arrays = None
    if debug:
        N = jac.num_cams * cp
        npc = jac.num_points - 3
        dbg_arrays = dict(U_s=np.zeros(jac.num_cams * cp * cp), Vinv_s=np.zeros(npc * 9),
                          d=np.zeros(N + 3 * npc), Z=np.zeros(N * N), X0=np.zeros(N * N),
                          X=np.zeros(N * N))
        dbg = Debug(**{k: v.ctypes.data for k, v in dbg_arrays.items()})
        rc = L.unc_compute_debug(ctx, ctypes.byref(o), jac.num_rows, jac.num_cols, _ptr(rows),
                                 _ptr(cols), _ptr(vals), _ptr(cam), _ptr(pt), ctypes.byref(st),
                                 ctypes.byref(dbg))
    else:
        rc = L.unc_compute(ctx, ctypes.byref(o), jac.num_rows, jac.num_cols, _ptr(rows),
                           _ptr(cols), _ptr(vals), _ptr(cam), _ptr(pt), ctypes.byref(st))
    _check(rc, "unc_compute")
    out = None
    if debug:
        N = jac.num_cams * cp
        out = dict(U_s=dbg_arrays["U_s"].reshape(jac.num_cams, cp, cp),
                   Vinv_s=dbg_arrays["Vinv_s"].reshape(-1, 3, 3), d=dbg_arrays["d"])
        for k in ("Z", "X0", "X"):
            out[k] = dbg_arrays[k].reshape(N, N).T.copy()
    return Result(cam.reshape(jac.num_cams, per), pt.reshape(jac.num_points, 6), st.lambda_,
                  st.num_terms, list(st.cycle_change[:st.num_terms]), tuple(st.fixed_pts), st, out)


class Context:
    """A device context (stream + reusable workspace), see unc_context_create."""

    def __init__(self, device: int = 0):
        self.handle = ctypes.c_void_p()
        _check(lib().unc_context_create(device, ctypes.byref(self.handle)), "unc_context_create")

    def close(self):
        if self.handle:
            lib().unc_context_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def compute_device(opts: Options, num_rows: int, num_cols: int, d_rows: int, d_cols: int,
                   d_vals: int, d_cam: int, d_pt: int, ctx=None, stream: int = 0) -> Stats:
    """Device-resident call (integer device addresses, e.g. torch .data_ptr())."""
    st = Stats()
    h = ctx.handle if isinstance(ctx, Context) else ctx
    rc = lib().unc_compute_device(h, ctypes.byref(opts), num_rows, num_cols, d_rows, d_cols,
                                  d_vals, d_cam, d_pt, ctypes.byref(st), stream or None)
    _check(rc, "unc_compute_device")
    return st


def workspace_bytes(opts: Options, num_rows: int) -> int:
    return int(lib().unc_workspace_bytes(ctypes.byref(opts), num_rows))


# --------------------------------------------------------------------------
# cov::Uncertainty accessors (src/uncertaintyTE/uncertainty.cpp:36-185)
# --------------------------------------------------------------------------

class Uncertainty:
    """Output container with the reference's unpack and eigen accessors
    (cov::Uncertainty, src/uncertaintyTE/uncertainty.h:84-143).  The unpack
    accessors are host-side index work, as in the reference; the eigen
    accessors run on the device (unc_eig_cameras / unc_eig_points) and raise
    UncError when no GPU is present."""

    def __init__(self, cam_params: int, cam_unc: np.ndarray, pt_unc: np.ndarray, ctx=None):
        self.cam_params = cam_params
        self.per = cam_params * (cam_params + 1) // 2
        self.cam = np.asarray(cam_unc, np.float64).reshape(-1, self.per)
        self.pt = np.asarray(pt_unc, np.float64).reshape(-1, 6)
        self.ctx = ctx

    def getCamerasUncRaw(self) -> np.ndarray:
        return self.cam.reshape(-1).copy()

    def getPointsUncRaw(self) -> np.ndarray:
        return self.pt.reshape(-1).copy()

    def getCameraUncMatrix(self, i: int) -> np.ndarray:
        # uncertainty.cpp:48-61: row-major upper triangle mirrored
        cp = self.cam_params
        M = np.zeros((cp, cp))
        iu = np.triu_indices(cp)
        M[iu] = self.cam[i]
        M.T[iu] = self.cam[i]
        return M

    def getCamerasUncMatrices(self) -> np.ndarray:
        cp = self.cam_params
        M = np.zeros((len(self.cam), cp, cp))
        iu = np.triu_indices(cp)
        M[:, iu[0], iu[1]] = self.cam
        M[:, iu[1], iu[0]] = self.cam
        return M

    def getPointUncMatrix(self, i: int) -> np.ndarray:
        # uncertainty.cpp:120-134
        a = self.pt[i]
        return np.array([[a[0], a[1], a[2]], [a[1], a[3], a[4]], [a[2], a[4], a[5]]])

    def getPointsUncMatrices(self) -> np.ndarray:
        a = self.pt
        return np.stack([a[:, 0], a[:, 1], a[:, 2], a[:, 1], a[:, 3], a[:, 4], a[:, 2], a[:, 4],
                         a[:, 5]], 1).reshape(-1, 3, 3)

    def getCamerasUncEigenValues(self) -> np.ndarray:
        """(num_cams, cp) eigenvalues, descending (uncertainty.cpp:74-89)."""
        return eig_cameras(self.cam, self.cam_params, values=True, vectors=False, ctx=self.ctx)[0]

    def getCamerasCentersUncEigVec(self) -> np.ndarray:
        """(num_cams, 9): eigenvectors of the camera-centre block, vector k at
        [3k:3k+3] (uncertainty.cpp:91-115), ordered by descending eigenvalue."""
        return eig_cameras(self.cam, self.cam_params, values=False, vectors=True, ctx=self.ctx)[1]

    def getPointsUncEigenValues(self) -> np.ndarray:
        """(num_points, 3) eigenvalues, descending (uncertainty.cpp:146-162)."""
        return eig_points(self.pt, values=True, vectors=False, ctx=self.ctx)[0]

    def getPointsUncEigVectors(self) -> np.ndarray:
        """(num_points, 9) eigenvectors, vector k at [3k:3k+3]
        (uncertainty.cpp:164-185), ordered by descending eigenvalue."""
        return eig_points(self.pt, values=False, vectors=True, ctx=self.ctx)[1]


def _handle(ctx):
    return ctx.handle if isinstance(ctx, Context) else ctx


def eig_cameras(cam_unc: np.ndarray, cam_params: int, values: bool = True, vectors: bool = True,
                ctx=None):
    """(eigenvalues (n, cp) descending or None, centre-block eigenvectors (n, 9)
    or None) of the packed camera covariances, on the device."""
    per = cam_params * (cam_params + 1) // 2
    cam = np.ascontiguousarray(cam_unc, np.float64).reshape(-1, per)
    n = cam.shape[0]
    ev = np.zeros((n, cam_params)) if values else None
    vec = np.zeros((n, 9)) if vectors else None
    _check(lib().unc_eig_cameras(_handle(ctx), n, cam_params, _ptr(cam),
                                 None if ev is None else _ptr(ev),
                                 None if vec is None else _ptr(vec)), "unc_eig_cameras")
    return ev, vec


def eig_points(pt_unc: np.ndarray, values: bool = True, vectors: bool = True, ctx=None):
    """(eigenvalues (n, 3) descending or None, eigenvectors (n, 9) or None) of
    the packed point covariances, on the device."""
    pt = np.ascontiguousarray(pt_unc, np.float64).reshape(-1, 6)
    n = pt.shape[0]
    ev = np.zeros((n, 3)) if values else None
    vec = np.zeros((n, 9)) if vectors else None
    _check(lib().unc_eig_points(_handle(ctx), n, _ptr(pt), None if ev is None else _ptr(ev),
                                None if vec is None else _ptr(vec)), "unc_eig_points")
    return ev, vec


def eig_cameras_device(num_cams: int, cam_params: int, d_cam: int, d_vals: int = 0,
                       d_vecs: int = 0, ctx=None, stream: int = 0):
    """Enqueue the camera eigen kernel on device addresses (e.g. torch .data_ptr()).
    stream 0 means the context's stream (not the null stream); returns at once."""
    _check(lib().unc_eig_cameras_device(_handle(ctx), num_cams, cam_params, d_cam, d_vals or None,
                                        d_vecs or None, stream or None), "unc_eig_cameras_device")


def eig_points_device(num_points: int, d_pt: int, d_vals: int = 0, d_vecs: int = 0, ctx=None,
                      stream: int = 0):
    """Enqueue the point eigen kernel on device addresses (e.g. torch .data_ptr()).
    stream 0 means the context's stream (not the null stream); returns at once."""
    _check(lib().unc_eig_points_device(_handle(ctx), num_points, d_pt, d_vals or None,
                                       d_vecs or None, stream or None), "unc_eig_points_device")


# --------------------------------------------------------------------------
# reconstruction front-end (ColmapIO.cpp:24-146, JacobianComposer.cpp:63-194)
# --------------------------------------------------------------------------

CAM_SIMPLE_PINHOLE, CAM_PINHOLE, CAM_SIMPLE_RADIAL, CAM_RADIAL = range(4)
CAM_MODELS = {"SIMPLE_PINHOLE": CAM_SIMPLE_PINHOLE, "PINHOLE": CAM_PINHOLE,
              "SIMPLE_RADIAL": CAM_SIMPLE_RADIAL, "RADIAL": CAM_RADIAL}
CAM_PARAMS = {CAM_SIMPLE_PINHOLE: 6, CAM_PINHOLE: 7, CAM_SIMPLE_RADIAL: 8, CAM_RADIAL: 9}


def _read_scene(reader: str, path: str):
    L = lib()
    h = ctypes.c_void_p()
    _check(getattr(L, reader)(os.fsencode(path), ctypes.byref(h)), reader)
    try:
        nc, npt, no = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(L.unc_scene_counts(h, ctypes.byref(nc), ctypes.byref(npt), ctypes.byref(no)),
               "unc_scene_counts")
        cams = np.zeros((nc.value, 9))
        pts = np.zeros((npt.value, 3))
        oc = np.zeros(no.value, np.int32)
        op = np.zeros(no.value, np.int32)
        _check(L.unc_scene_arrays(h, _ptr(cams), _ptr(pts), _ptr(oc), _ptr(op)),
               "unc_scene_arrays")
    finally:
        L.unc_scene_destroy(h)
    return cams, pts, oc, op


def read_colmap(path: str):
    """(cams (n_cams, 9) [aa, C, f, k1, k2], points (n_points, 3), obs_cam,
    obs_pt) of a COLMAP text model, in the reference's parameter-block order."""
    return _read_scene("unc_scene_read_colmap", path)


def read_bal(path: str):
    """(cams (n_cams, 9) [aa, C, f, k1, k2], points (n_points, 3), obs_cam,
    obs_pt) of a BAL problem file, in file order (unc_scene_read_bal); the
    Jacobian of these is scene_jacobian(..., CAM_RADIAL)."""
    return _read_scene("unc_scene_read_bal", path)


def find_pts2fix(points: np.ndarray) -> tuple:
    """JacobianComposer::findPts2Fix with the library's fixed seed."""
    P = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
    out = (ctypes.c_int * 3)()
    _check(lib().unc_find_pts2fix(P.shape[0], _ptr(P), out), "unc_find_pts2fix")
    return tuple(out)


def scene_jacobian(cams: np.ndarray, points: np.ndarray, obs_cam: np.ndarray,
                   obs_pt: np.ndarray, model: int = CAM_RADIAL, fix=None, ctx=None):
    """jacob.Jacobian of a scene by the device kernel (unc_scene_jacobian);
    fixed points from find_pts2fix unless given."""
    from . import jacob as _jacob
    cams = np.ascontiguousarray(cams, np.float64).reshape(-1, 9)
    pts = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
    oc = np.ascontiguousarray(obs_cam, np.int32)
    op = np.ascontiguousarray(obs_pt, np.int32)
    cp = CAM_PARAMS[model]
    nobs = oc.shape[0]
    rows = np.zeros(2 * nobs + 1, np.int32)
    cols = np.zeros(2 * nobs * (cp + 3), np.int32)
    vals = np.zeros(2 * nobs * (cp + 3))
    _check(lib().unc_scene_jacobian(_handle(ctx), model, cams.shape[0], pts.shape[0], nobs,
                                    _ptr(cams), _ptr(pts), _ptr(oc), _ptr(op), _ptr(rows),
                                    _ptr(cols), _ptr(vals)), "unc_scene_jacobian")
    if fix is None:
        fix = find_pts2fix(pts)
    return _jacob.Jacobian(-1.0, cams.shape[0], cp, pts.shape[0], nobs, tuple(fix), 2 * nobs,
                           cams.shape[0] * cp + 3 * pts.shape[0], rows, cols, vals)


class Server:
    """Request loop (unc_server_*): submit() returns a ticket at once, wait()
    returns that request's Result plus its latency and queueing time (ms).
    The Jacobian arrays are kept referenced until the ticket is waited for."""

    def __init__(self, device: int = 0, slots: int = 2, hbm_budget: int = 64 << 30,
                 bypass_ms: float = 50.0):
        L = lib()
        cfg = ServerConfig()
        L.unc_server_default_config(ctypes.byref(cfg))
        cfg.device, cfg.slots, cfg.hbm_budget, cfg.bypass_ms = device, slots, hbm_budget, bypass_ms
        self.handle = ctypes.c_void_p()
        _check(L.unc_server_create(ctypes.byref(cfg), ctypes.byref(self.handle)),
               "unc_server_create")
        self._live = {}

    def submit(self, jac, precision: int = PREC_MIXED, refine=True, fix=None) -> int:
        o = make_options(jac, precision, refine, 0, fix, None)
        cp = jac.cam_params
        cam = np.zeros(jac.num_cams * (cp * (cp + 1) // 2), np.float64)
        pt = np.zeros(jac.num_points * 6, np.float64)
        rows = np.ascontiguousarray(jac.rows, np.int32)
        cols = np.ascontiguousarray(jac.cols, np.int32)
        vals = np.ascontiguousarray(jac.vals, np.float64)
        st = Stats()
        t = ctypes.c_longlong()
        _check(lib().unc_server_submit(self.handle, ctypes.byref(o), jac.num_rows, jac.num_cols,
                                       _ptr(rows), _ptr(cols), _ptr(vals), _ptr(cam), _ptr(pt),
                                       ctypes.byref(st), ctypes.byref(t)), "unc_server_submit")
        self._live[t.value] = (jac, o, rows, cols, vals, cam, pt, st)
        return t.value

    def wait(self, ticket: int):
        status = ctypes.c_int()
        lat = ctypes.c_double()
        q = ctypes.c_double()
        _check(lib().unc_server_wait(self.handle, ticket, ctypes.byref(status), ctypes.byref(lat),
                                     ctypes.byref(q)), "unc_server_wait")
        jac, o, rows, cols, vals, cam, pt, st = self._live.pop(ticket)
        _check(status.value, "request")
        cp = jac.cam_params
        res = Result(cam.reshape(jac.num_cams, cp * (cp + 1) // 2), pt.reshape(jac.num_points, 6),
                     st.lambda_, st.num_terms, list(st.cycle_change[:st.num_terms]),
                     tuple(st.fixed_pts), st, None)
        return res, lat.value, q.value

    def stats(self) -> ServerStats:
        s = ServerStats()
        _check(lib().unc_server_query(self.handle, ctypes.byref(s)), "unc_server_query")
        return s

    def close(self):
        if self.handle:
            lib().unc_server_destroy(self.handle)
            self.handle = ctypes.c_void_p()
            self._live.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
