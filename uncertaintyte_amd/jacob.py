"""`.jacob` Jacobian reader/writer and `.cov` covariance writer/reader.

Formats follow the reference:

* `.jacob` text (JacobianIO.cpp:32-63, documented in
  test/input_jacobians/template_of_jacobian.txt:4-14): a whitespace-separated
  token stream ``lambda numCams camParams numPoints numObs`` / three fixed
  point ids (``-1 -1 -1`` = none) / ``num_rows num_cols nnz`` / ``rows[num_rows+1]``
  / ``cols[nnz]`` / ``values[nnz]`` of a ceres::CRSMatrix.
* `.cov` text (IO.cpp:15-60): ``#``-prefixed header, then one line per camera
  with cp(cp+1)/2 upper-triangle values and one line per point with 6 values,
  each value printed with the iostream default (6 significant digits, ``%g``)
  and followed by a space.

A binary sidecar (``.jacob.npz``, plain arrays, no pickle) carries the same
fields for scenes whose text form would be gigabytes.
"""
from __future__ import annotations

import dataclasses
import os
from typing import Optional, Sequence

import numpy as np

ALGORITHMS = {0: "SVD_QR_ITERATION", 1: "SVD_DEVIDE_AND_CONQUER", 2: "TAYLOR_EXPANSION"}
TAYLOR_EXPANSION = 2


@dataclasses.dataclass
class Jacobian:
    """A ceres::CRSMatrix plus the `.jacob` header (cov::Options fields)."""

    lam: float
    num_cams: int
    cam_params: int
    num_points: int
    num_obs: int
    fix: Optional[tuple]  # three point ids, or None when the header holds -1
    num_rows: int
    num_cols: int
    rows: np.ndarray  # int32[num_rows+1]
    cols: np.ndarray  # int32[nnz]
    vals: np.ndarray  # float64[nnz]

    @property
    def nnz(self) -> int:
        return int(self.rows[-1]) if len(self.rows) else 0

    def pinned(self) -> "Jacobian":
        """The same Jacobian with its arrays in page-locked host memory
        (uncertaintyte_amd.pinned_copy), for repeated fast uploads."""
        from . import pinned_copy
        return dataclasses.replace(self, rows=pinned_copy(np.ascontiguousarray(self.rows, np.int32)),
                                   cols=pinned_copy(np.ascontiguousarray(self.cols, np.int32)),
                                   vals=pinned_copy(np.ascontiguousarray(self.vals, np.float64)))

    def fix_array(self) -> np.ndarray:
        return np.array(self.fix if self.fix is not None else (-1, -1, -1), dtype=np.int32)


def _parse_tokens(text: str) -> np.ndarray:
    # Comment lines (template_of_jacobian.txt) are not part of real files but
    # are tolerated: everything after '#' on a line is dropped.
    if "#" in text:
        text = "\n".join(line.split("#", 1)[0] for line in text.splitlines())
    return np.array(text.split(), dtype=np.float64) if len(text) < (1 << 20) else \
        np.fromstring(text, dtype=np.float64, sep=" ")  # noqa: NPY201 (text mode is supported)


def read_jacob(path: str) -> Jacobian:
    """Read a `.jacob` file (or its `.npz` sidecar)."""
    if path.endswith(".npz"):
        return read_jacob_npz(path)
    with open(path, "r") as f:
        text = f.read()
    t = _parse_tokens(text)
    if t.size < 11:
        raise ValueError(f"{path}: truncated .jacob header")
    lam, nc, cp, npts, nobs = float(t[0]), int(t[1]), int(t[2]), int(t[3]), int(t[4])
    fx = tuple(int(v) for v in t[5:8])
    nr, ncl, nnz = int(t[8]), int(t[9]), int(t[10])
    need = 11 + (nr + 1) + 2 * nnz
    if t.size < need:
        raise ValueError(f"{path}: truncated .jacob body ({t.size} tokens, need {need})")
    o = 11
    rows = t[o:o + nr + 1].astype(np.int32); o += nr + 1
    cols = t[o:o + nnz].astype(np.int32); o += nnz
    vals = np.ascontiguousarray(t[o:o + nnz], dtype=np.float64)
    fix = fx if (fx[0] >= 0 and fx[1] >= 0 and fx[2] >= 0) else None  # JacobianIO.cpp:47
    return Jacobian(lam, nc, cp, npts, nobs, fix, nr, ncl, rows, cols, vals)


def _fmt_g(v: float) -> str:
    return "%.17g" % v


def write_jacob(path: str, jac: Jacobian, precision: int = 17) -> None:
    """Write the text form; values with `precision` significant digits."""
    fx = jac.fix if jac.fix is not None else (-1, -1, -1)
    lam = jac.lam
    lam_s = "-1" if lam == -1 else repr(float(lam))
    with open(path, "w") as f:
        f.write(f"{lam_s} {jac.num_cams} {jac.cam_params} {jac.num_points} {jac.num_obs}\n")
        f.write(f"{fx[0]} {fx[1]} {fx[2]}\n")
        f.write(f"{jac.num_rows} {jac.num_cols} {jac.nnz}\n")
        np.savetxt(f, jac.rows[None, :], fmt="%d", delimiter=" ")
        np.savetxt(f, jac.cols[None, :], fmt="%d", delimiter=" ")
        np.savetxt(f, jac.vals[None, :], fmt=f"%.{precision}g", delimiter=" ")


def write_jacob_npz(path: str, jac: Jacobian) -> None:
    hdr = np.array([jac.num_cams, jac.cam_params, jac.num_points, jac.num_obs,
                    jac.num_rows, jac.num_cols], dtype=np.int64)
    np.savez(path, hdr=hdr, lam=np.array([jac.lam]), fix=jac.fix_array(),
             rows=jac.rows.astype(np.int32), cols=jac.cols.astype(np.int32),
             vals=jac.vals.astype(np.float64))


def read_jacob_npz(path: str) -> Jacobian:
    z = np.load(path, allow_pickle=False)
    h = z["hdr"]
    fx = tuple(int(v) for v in z["fix"])
    fix = fx if min(fx) >= 0 else None
    return Jacobian(float(z["lam"][0]), int(h[0]), int(h[1]), int(h[2]), int(h[3]), fix,
                    int(h[4]), int(h[5]), z["rows"].astype(np.int32),
                    z["cols"].astype(np.int32), z["vals"].astype(np.float64))


# --------------------------------------------------------------------------
# .cov
# --------------------------------------------------------------------------

def _g(v: float) -> str:
    """iostream default formatting of a double (precision 6, %g rules)."""
    return "%g" % v


@dataclasses.dataclass
class CovHeader:
    num_cams: int
    cam_params: int
    num_points: int
    num_obs: int
    algorithm: int = TAYLOR_EXPANSION
    fixed: Optional[Sequence[int]] = None
    lam: float = 0.0
    # Stage times in the reference's printed unit (milliseconds * 10e-3, i.e.
    # ten times the true seconds: compute.cpp:33).
    time_create_j: float = 0.0
    time_norm_j: float = 0.0
    time_multiply_jj: float = 0.0
    time_split_jj: float = 0.0
    time_inv_v: float = 0.0
    time_compose_z: float = 0.0
    time_inv_z: float = 0.0
    time_te: float = 0.0
    time_pts_unc: float = 0.0
    time_all: float = 0.0
    cycle_change: Sequence[float] = ()


def write_cov(path: str, h: CovHeader, cam_unc: np.ndarray, pt_unc: np.ndarray) -> None:
    """saveResults (IO.cpp:15-60), byte-compatible layout."""
    per = h.cam_params * (h.cam_params + 1) // 2
    out = []
    out.append("# ---- Covariance v0.1 ----\n")
    out.append(f"# Number of cameras: {h.num_cams}\n")
    out.append(f"# Number of camera parameters: {h.cam_params}\n")
    out.append(f"# Number of points in 3D: {h.num_points}\n")
    out.append(f"# Number of observations: {h.num_obs}\n")
    out.append(f"# Used algorithm: {ALGORITHMS.get(h.algorithm, 'not defined')}\n")
    if h.algorithm == TAYLOR_EXPANSION:
        if h.fixed is not None:
            out.append(f"# Fixed points: {h.fixed[0]}, {h.fixed[1]}, {h.fixed[2]}\n")
        out.append(f"# Used lambda: {_g(h.lam)}\n")
        out.append(f"# Loading jacobian time: {_g(h.time_create_j)}s\n")
        out.append(f"# Normalization of jacobian time: {_g(h.time_norm_j)}s\n")
        out.append(f"# Compose information matrix time: {_g(h.time_multiply_jj)}s\n")
        out.append(f"# Split infromation matrix time: {_g(h.time_split_jj)}s\n")
        out.append(f"# Inverse V time: {_g(h.time_inv_v)}s\n")
        out.append(f"# Compose Z time: {_g(h.time_compose_z)}s\n")
        out.append(f"# Inverse Z time: {_g(h.time_inv_z)}s\n")
        out.append(f"# Taylor expansion time: {_g(h.time_te)}s\n")
        out.append(f"# Point uncertainty time: {_g(h.time_pts_unc)}s\n")
        out.append(f"# TE number of iterations: {len(h.cycle_change)}\n")
        out.append("# TE cycle change: " + "".join(_g(c) + " " for c in h.cycle_change) + "\n")
    out.append(f"# Time of the algorithm: {_g(h.time_all)}s\n")
    out.append("# -----------------------\n")
    cam = np.asarray(cam_unc, dtype=np.float64).reshape(h.num_cams, per)
    for r in cam:
        out.append("".join(_g(v) + " " for v in r) + "\n")
    pts = np.asarray(pt_unc, dtype=np.float64).reshape(h.num_points, 6)
    for r in pts:
        out.append("".join(_g(v) + " " for v in r) + "\n")
    with open(path, "w") as f:
        f.write("".join(out))


def read_cov(path: str, num_cams: Optional[int] = None):
    """Parse a `.cov` file -> (header dict, cam values (nC, per), point values (nP, 6))."""
    hdr = {}
    cams, pts = [], []
    with open(path) as f:
        lines = f.read().splitlines()
    for line in lines:
        if line.startswith("#"):
            if ":" in line:
                k, v = line[1:].split(":", 1)
                hdr[k.strip()] = v.strip()
            continue
        if len(line) <= 1:
            continue
        vals = [float(x) for x in line.split()]
        nc = num_cams if num_cams is not None else int(hdr.get("Number of cameras", "0"))
        (cams if len(cams) < nc else pts).append(vals)
    return hdr, np.array(cams, dtype=np.float64), np.array(pts, dtype=np.float64)


def file_size(path: str) -> int:
    return os.path.getsize(path)
