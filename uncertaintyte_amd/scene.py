"""Seeded synthetic reconstructions and their analytic reprojection Jacobians.

Camera model: RADIAL of snavely_reprojection_error.h:265-302 -- parameters per
camera [angle-axis (3), centre C (3), f, k1, k2] in the parameter-block order
of JacobianComposer.cpp:131-157:

    p = R(aa) (X - C);  x = p0/p2, y = p1/p2;  r2 = x^2 + y^2
    u = f x (1 + k1 r2 + k2 r2^2),  v = f y (1 + k1 r2 + k2 r2^2)

The Jacobian is evaluated at the true parameters (pixel noise only moves the
observations, which a `.jacob` does not carry).  Rows are written camera-major
(observations sorted by camera, then point), two rows per observation, each
row holding the camera's cp columns followed by the point's three, which is
the layout of the reference's input files (template_of_jacobian.txt:4-12).

Workloads (sizes as named in the measurement plan):
    ring_scene(10, 1000, ...)          toy: 10 cameras on a ring, 1k points, 2-6 views
    rings_scene(200, 100_000, m=4)     200 cameras on 4 rings, 100k points
    walk_scene(1500, 1_000_000, m=5)   random-walk trajectory, 10% short-baseline points
"""
from __future__ import annotations

import dataclasses
import os
from typing import Optional

import numpy as np

from .jacob import Jacobian


# --------------------------------------------------------------------------
# rotations
# --------------------------------------------------------------------------

def _skew(v):
    z = np.zeros(v.shape[:-1])
    return np.stack([np.stack([z, -v[..., 2], v[..., 1]], -1),
                     np.stack([v[..., 2], z, -v[..., 0]], -1),
                     np.stack([-v[..., 1], v[..., 0], z], -1)], -2)


def aa_to_R(w):
    """Rodrigues: rotation matrices for angle-axis vectors w (..., 3)."""
    th = np.linalg.norm(w, axis=-1)
    small = th < 1e-8
    ths = np.where(small, 1.0, th)
    K = _skew(w)
    a = np.where(small, 1.0, np.sin(ths) / ths)
    b = np.where(small, 0.5, (1 - np.cos(ths)) / ths ** 2)
    I = np.broadcast_to(np.eye(3), K.shape)
    return I + a[..., None, None] * K + b[..., None, None] * (K @ K)


def R_to_aa(R):
    """Rotation matrices -> angle-axis via quaternions (Shepperd's method)."""
    R = np.asarray(R)
    tr = np.trace(R, axis1=-2, axis2=-1)
    q = np.zeros(R.shape[:-2] + (4,))
    c0 = tr > 0
    s = np.sqrt(np.maximum(tr + 1.0, 1e-300)) * 2
    q[..., 0] = np.where(c0, 0.25 * s, 0)
    q[..., 1] = np.where(c0, (R[..., 2, 1] - R[..., 1, 2]) / s, 0)
    q[..., 2] = np.where(c0, (R[..., 0, 2] - R[..., 2, 0]) / s, 0)
    q[..., 3] = np.where(c0, (R[..., 1, 0] - R[..., 0, 1]) / s, 0)
    for i, (j, k) in enumerate([(1, 2), (2, 0), (0, 1)]):
        ci = (~c0) & (R[..., i, i] >= R[..., j, j]) & (R[..., i, i] >= R[..., k, k])
        si = np.sqrt(np.maximum(1.0 + R[..., i, i] - R[..., j, j] - R[..., k, k], 1e-300)) * 2
        q[..., 0] = np.where(ci, (R[..., k, j] - R[..., j, k]) / si, q[..., 0])
        q[..., 1 + i] = np.where(ci, 0.25 * si, q[..., 1 + i])
        q[..., 1 + j] = np.where(ci, (R[..., j, i] + R[..., i, j]) / si, q[..., 1 + j])
        q[..., 1 + k] = np.where(ci, (R[..., k, i] + R[..., i, k]) / si, q[..., 1 + k])
        c0 = c0 | ci
    q *= np.where(q[..., :1] < 0, -1.0, 1.0)
    vn = np.linalg.norm(q[..., 1:], axis=-1)
    ang = 2 * np.arctan2(vn, q[..., 0])
    scale = np.where(vn < 1e-12, 2.0, ang / np.maximum(vn, 1e-300))
    return q[..., 1:] * scale[..., None]


def look_at(C, target, up=(0.0, 0.0, 1.0)):
    """World->camera rotations for cameras at C looking at target (optical axis +z)."""
    z = target - C
    z /= np.linalg.norm(z, axis=-1, keepdims=True)
    up = np.broadcast_to(np.asarray(up, dtype=np.float64), z.shape)
    x = np.cross(up, z)
    bad = np.linalg.norm(x, axis=-1) < 1e-9
    if np.any(bad):
        x[bad] = np.cross(np.array([1.0, 0.0, 0.0]), z[bad])
    x /= np.linalg.norm(x, axis=-1, keepdims=True)
    y = np.cross(z, x)
    return np.stack([x, y, z], axis=-2)


# --------------------------------------------------------------------------
# projection + analytic Jacobian
# --------------------------------------------------------------------------

def project(cam, X):
    """u, v for camera params cam (..., 9) and points X (..., 3)."""
    R = aa_to_R(cam[..., 0:3])
    p = np.einsum("...ij,...j->...i", R, X - cam[..., 3:6])
    x, y = p[..., 0] / p[..., 2], p[..., 1] / p[..., 2]
    r2 = x * x + y * y
    g = 1 + cam[..., 7] * r2 + cam[..., 8] * r2 * r2
    return np.stack([cam[..., 6] * x * g, cam[..., 6] * y * g], -1)


def jacobian_blocks(cam, X):
    """Analytic d(u,v)/d(cam) (..., 2, 9) and d(u,v)/dX (..., 2, 3)."""
    w = cam[..., 0:3]
    R = aa_to_R(w)
    pc = X - cam[..., 3:6]
    p = np.einsum("...ij,...j->...i", R, pc)
    f, k1, k2 = cam[..., 6], cam[..., 7], cam[..., 8]
    x, y = p[..., 0] / p[..., 2], p[..., 1] / p[..., 2]
    r2 = x * x + y * y
    g = 1 + k1 * r2 + k2 * r2 * r2
    gr = k1 + 2 * k2 * r2                     # dg/d(r2)
    gx, gy = 2 * x * gr, 2 * y * gr
    duv_dxy = np.stack([np.stack([f * (g + x * gx), f * x * gy], -1),
                        np.stack([f * y * gx, f * (g + y * gy)], -1)], -2)
    iz = 1.0 / p[..., 2]
    zz = np.zeros_like(iz)
    dxy_dp = np.stack([np.stack([iz, zz, -p[..., 0] * iz * iz], -1),
                       np.stack([zz, iz, -p[..., 1] * iz * iz], -1)], -2)
    duv_dp = duv_dxy @ dxy_dp                               # (..., 2, 3)
    # d(R v)/dw = -R [v]x J_r(w)  (right Jacobian of SO(3))
    th = np.linalg.norm(w, axis=-1)
    small = th < 1e-6
    ths = np.where(small, 1.0, th)
    c1 = np.where(small, 0.5 - th ** 2 / 24, (1 - np.cos(ths)) / ths ** 2)
    c2 = np.where(small, 1.0 / 6 - th ** 2 / 120, (ths - np.sin(ths)) / ths ** 3)
    Kw = _skew(w)
    Jr = np.eye(3) - c1[..., None, None] * Kw + c2[..., None, None] * (Kw @ Kw)
    dp_dw = -R @ _skew(pc) @ Jr
    J_aa = duv_dp @ dp_dw
    J_C = -(duv_dp @ R)
    J_X = duv_dp @ R
    J_f = np.stack([x * g, y * g], -1)[..., None]
    J_k1 = np.stack([f * x * r2, f * y * r2], -1)[..., None]
    J_k2 = np.stack([f * x * r2 * r2, f * y * r2 * r2], -1)[..., None]
    Jc = np.concatenate([J_aa, J_C, J_f, J_k1, J_k2], -1)
    return Jc, J_X


# --------------------------------------------------------------------------
# scenes
# --------------------------------------------------------------------------

@dataclasses.dataclass
class Scene:
    cams: np.ndarray      # (nC, 9)
    points: np.ndarray    # (nP, 3)
    obs_cam: np.ndarray   # (nObs,) camera of each observation, camera-major order
    obs_pt: np.ndarray    # (nObs,)
    fix: tuple            # three fixed point ids, ascending
    seed: int
    name: str = ""

    @property
    def num_cams(self):
        return self.cams.shape[0]

    @property
    def num_points(self):
        return self.points.shape[0]

    @property
    def num_obs(self):
        return self.obs_cam.shape[0]


def _camera_params(rng, C, R, f0=1000.0):
    n = C.shape[0]
    aa = R_to_aa(R)
    f = f0 * (1 + 0.02 * rng.standard_normal(n))
    k = 0.01 * rng.standard_normal((n, 2))
    return np.concatenate([aa, C, f[:, None], k], 1)


def _tracks_uniform(rng, n_pts, candidates, m, lo=2, hi=None):
    """For each point choose max(lo, Poisson(m)) (or U{lo..hi}) distinct cameras
    among `candidates` (n_pts, K) candidate camera ids (-1 = none)."""
    K = candidates.shape[1]
    if hi is not None:
        n = rng.integers(lo, hi + 1, n_pts)
    else:
        n = np.maximum(lo, rng.poisson(m, n_pts))
    valid = candidates >= 0
    n = np.minimum(n, valid.sum(1))
    keys = rng.random(candidates.shape)
    keys[~valid] = 2.0
    order = np.argsort(keys, axis=1)
    sel = np.take_along_axis(candidates, order, 1)
    take = np.arange(K)[None, :] < n[:, None]
    pts = np.broadcast_to(np.arange(n_pts)[:, None], sel.shape)[take]
    cams = sel[take]
    return cams, pts


def _finish(rng, name, seed, cams, points, oc, op):
    # drop observations with the point behind / too close to the camera
    R = aa_to_R(cams[oc, 0:3])
    p = np.einsum("nij,nj->ni", R, points[op] - cams[oc, 3:6])
    ok = p[:, 2] > 1e-2
    oc, op = oc[ok], op[ok]
    # points need >= 2 views
    cnt = np.bincount(op, minlength=points.shape[0])
    keep_pt = cnt >= 2
    remap = -np.ones(points.shape[0], np.int64)
    remap[keep_pt] = np.arange(keep_pt.sum())
    sel = keep_pt[op]
    oc, op = oc[sel], remap[op[sel]]
    points = points[keep_pt]
    # every camera must observe something
    used = np.bincount(oc, minlength=cams.shape[0]) > 0
    if not np.all(used):
        cremap = -np.ones(cams.shape[0], np.int64)
        cremap[used] = np.arange(used.sum())
        cams = cams[used]
        oc = cremap[oc]
    order = np.lexsort((op, oc))
    oc, op = oc[order].astype(np.int32), op[order].astype(np.int32)
    # deterministic "most distant triple" (findPts2Fix, JacobianComposer.cpp:63-83, seeded)
    nP = points.shape[0]
    cand = rng.choice(nP, size=min(nP, 512), replace=False)
    P = points[cand]
    D = np.linalg.norm(P[:, None, :] - P[None, :, :], axis=-1)
    i, j = np.unravel_index(np.argmax(D), D.shape)
    k = np.argmax(D[i] + D[j])
    fix = tuple(sorted(int(cand[t]) for t in (i, j, k)))
    if len(set(fix)) < 3:
        fix = (0, 1, 2)
    return Scene(cams, points, oc, op, fix, seed, name)


def ring_scene(n_cams=10, n_pts=1000, seed=1, radius=3.0, views=(2, 6)):
    """Cameras on a ring in z=0 around the unit cube, points uniform in the cube."""
    rng = np.random.default_rng(seed)
    ang = 2 * np.pi * np.arange(n_cams) / n_cams + 0.05 * rng.standard_normal(n_cams)
    C = np.stack([radius * np.cos(ang), radius * np.sin(ang), 0.05 * rng.standard_normal(n_cams)], 1)
    R = look_at(C, np.zeros((n_cams, 3)) + 0.02 * rng.standard_normal((n_cams, 3)))
    cams = _camera_params(rng, C, R)
    X = rng.uniform(-0.5, 0.5, (n_pts, 3))
    cand = np.broadcast_to(np.arange(n_cams), (n_pts, n_cams))
    oc, op = _tracks_uniform(rng, n_pts, cand, 0, lo=views[0], hi=views[1])
    return _finish(rng, f"ring{n_cams}x{n_pts}", seed, cams, X, oc, op)


def rings_scene(n_cams=200, n_pts=100_000, m=4.0, seed=2, n_rings=4):
    """n_rings rings of cameras (different radius and height) around the unit cube."""
    rng = np.random.default_rng(seed)
    per = n_cams // n_rings
    Cs = []
    for r in range(n_rings):
        k = per if r < n_rings - 1 else n_cams - per * (n_rings - 1)
        ang = 2 * np.pi * (np.arange(k) + 0.5 * r) / k
        rad = 3.0 + 0.8 * r
        h = -0.9 + 0.6 * r
        Cs.append(np.stack([rad * np.cos(ang), rad * np.sin(ang), np.full(k, h)], 1))
    C = np.concatenate(Cs) + 0.02 * rng.standard_normal((n_cams, 3))
    R = look_at(C, 0.05 * rng.standard_normal((n_cams, 3)))
    cams = _camera_params(rng, C, R)
    X = rng.uniform(-0.5, 0.5, (n_pts, 3))
    # every camera sees the cube; choose among 24 random candidates per point
    K = min(n_cams, 24)
    cand = np.argsort(rng.random((n_pts, n_cams)), axis=1)[:, :K] if n_cams <= 256 else \
        rng.integers(0, n_cams, (n_pts, K))
    if n_cams > 256:  # remove duplicate candidates
        s = np.sort(cand, 1)
        dup = np.zeros_like(s, bool)
        dup[:, 1:] = s[:, 1:] == s[:, :-1]
        s[dup] = -1
        cand = s
    oc, op = _tracks_uniform(rng, n_pts, cand, m)
    return _finish(rng, f"rings{n_cams}x{n_pts}", seed, cams, X, oc, op)


def walk_scene(n_cams=1500, n_pts=1_000_000, m=5.0, seed=3, short_frac=0.10, window=40):
    """Forward-looking cameras on a smooth random walk; points ahead of an anchor
    camera, seen by cameras of a window behind it.  `short_frac` of the points
    are far away and seen from a few consecutive cameras (parallax < 1 deg)."""
    rng = np.random.default_rng(seed)
    step = 0.1
    heading = np.cumsum(0.03 * rng.standard_normal(n_cams))
    pitch = 0.02 * np.cumsum(0.02 * rng.standard_normal(n_cams))
    d = np.stack([np.cos(heading) * np.cos(pitch), np.sin(heading) * np.cos(pitch), np.sin(pitch)], 1)
    C = np.cumsum(step * d, 0)
    fwd = d + 0.02 * rng.standard_normal((n_cams, 3))
    R = look_at(C, C + fwd)
    cams = _camera_params(rng, C, R)
    n_short = int(round(short_frac * n_pts))
    n_norm = n_pts - n_short
    # normal points: anchor camera s, depth 2..12 ahead, within the field of view,
    # seen from the window of cameras s - window + 1 .. s (those that exist)
    s = rng.integers(1, n_cams, n_norm)
    depth = rng.uniform(2.0, 12.0, n_norm)
    lat = rng.uniform(-0.35, 0.35, (n_norm, 2)) * depth[:, None]
    Rs = R[s]
    X1 = C[s] + depth[:, None] * Rs[:, 2] + lat[:, :1] * Rs[:, 0] + lat[:, 1:] * Rs[:, 1]
    cand1 = s[:, None] - np.arange(window)[None, :]
    cand1[cand1 < 0] = -1
    # short-baseline points: far (depth 40..80), seen by 2..4 consecutive cameras
    s2 = rng.integers(1, n_cams, n_short)
    depth2 = rng.uniform(40.0, 80.0, n_short)
    lat2 = rng.uniform(-0.3, 0.3, (n_short, 2)) * depth2[:, None]
    Rs2 = R[s2]
    X2 = C[s2] + depth2[:, None] * Rs2[:, 2] + lat2[:, :1] * Rs2[:, 0] + lat2[:, 1:] * Rs2[:, 1]
    cand2 = -np.ones((n_short, window), np.int64)
    cand2[:, :4] = s2[:, None] - np.arange(4)[None, :]
    cand2[cand2 < 0] = -1
    X = np.concatenate([X1, X2])
    oc1, op1 = _tracks_uniform(rng, n_norm, cand1, m)
    oc2, op2 = _tracks_uniform(rng, n_short, cand2[:, :4], 0, lo=2, hi=4)
    oc = np.concatenate([oc1, oc2])
    op = np.concatenate([op1, op2 + n_norm])
    return _finish(rng, f"walk{n_cams}x{n_pts}", seed, cams, X, oc, op)


def scene_jacobian(sc: Scene, lam: float = -1.0, cp: int = 9) -> Jacobian:
    """CRS Jacobian of a scene in the `.jacob` layout (camera-major rows)."""
    nC, nP, nO = sc.num_cams, sc.num_points, sc.num_obs
    N = nC * cp
    rl = cp + 3
    vals = np.empty((nO, 2, rl), np.float64)
    chunk = 1 << 20
    for o0 in range(0, nO, chunk):
        sl = slice(o0, min(nO, o0 + chunk))
        Jc, Jx = jacobian_blocks(sc.cams[sc.obs_cam[sl]], sc.points[sc.obs_pt[sl]])
        vals[sl, :, :cp] = Jc[..., :cp]
        vals[sl, :, cp:] = Jx
    cols = np.empty((nO, 2, rl), np.int32)
    cc = sc.obs_cam.astype(np.int64)[:, None] * cp + np.arange(cp)[None, :]
    pc = N + sc.obs_pt.astype(np.int64)[:, None] * 3 + np.arange(3)[None, :]
    cols[:, 0, :cp] = cc
    cols[:, 1, :cp] = cc
    cols[:, 0, cp:] = pc
    cols[:, 1, cp:] = pc
    rows = (np.arange(2 * nO + 1, dtype=np.int64) * rl).astype(np.int32)
    return Jacobian(lam, nC, cp, nP, nO, tuple(sc.fix), 2 * nO, N + 3 * nP, rows,
                    cols.reshape(-1), vals.reshape(-1))


WORKLOADS = {
    "toy": lambda seed=1: ring_scene(10, 1000, seed=seed),
    "rings200": lambda seed=2: rings_scene(200, 100_000, 4.0, seed=seed),
    "walk1500": lambda seed=3: walk_scene(1500, 1_000_000, 5.0, seed=seed),
    "walk4000": lambda seed=5: walk_scene(4000, 3_000_000, 6.0, seed=seed),
}


def make_workload(name: str, seed: Optional[int] = None) -> Jacobian:
    sc = WORKLOADS[name]() if seed is None else WORKLOADS[name](seed)
    return scene_jacobian(sc)


# --------------------------------------------------------------------------
# COLMAP text model writer (the input of ColmapIO::read, ColmapIO.cpp:24-146)
# --------------------------------------------------------------------------

def aa_to_quat(w):
    """Unit quaternions (qw, qx, qy, qz) of angle-axis vectors (..., 3)."""
    th = np.linalg.norm(w, axis=-1)
    ths = np.where(th > 0, th, 1.0)
    s = np.where(th > 0, np.sin(th / 2) / ths, 0.5)
    return np.concatenate([np.cos(th / 2)[..., None], w * s[..., None]], -1)


def write_colmap(path: str, sc: Scene, model: str = "RADIAL", image_id=None, point_id=None,
                 principal=(640.0, 480.0), extra_image: bool = False):
    """Write sc as cameras.txt / images.txt / points3D.txt (one COLMAP camera
    per image; observations are the projections plus the principal point).
    image_id / point_id map scene indices to file ids (default 1 + index);
    images are written in decreasing id and each image's observations in
    decreasing point id, so a reader has to sort.  extra_image adds an image
    whose only 2D point has no 3D point (POINT3D_ID -1)."""
    os.makedirs(path, exist_ok=True)
    nC, nP = sc.num_cams, sc.num_points
    iid = np.arange(1, nC + 1) if image_id is None else np.asarray(image_id)
    pid = np.arange(1, nP + 1) if point_id is None else np.asarray(point_id)
    cx, cy = principal
    with open(os.path.join(path, "cameras.txt"), "w") as f:
        f.write("# Camera list with one line of data per camera:\n")
        f.write("#   CAMERA_ID, MODEL, WIDTH, HEIGHT, PARAMS[]\n")
        for c in range(nC):
            fc, k1, k2 = sc.cams[c, 6:9]
            params = {"SIMPLE_PINHOLE": [fc, cx, cy], "PINHOLE": [fc, fc, cx, cy],
                      "SIMPLE_RADIAL": [fc, cx, cy, k1], "RADIAL": [fc, cx, cy, k1, k2]}[model]
            f.write(f"{c + 1} {model} {int(2 * cx)} {int(2 * cy)} "
                    + " ".join(repr(float(v)) for v in params) + "\n")
    uv = project(sc.cams[sc.obs_cam], sc.points[sc.obs_pt])
    q = aa_to_quat(sc.cams[:, 0:3])
    R = aa_to_R(sc.cams[:, 0:3])
    t = -np.einsum("cij,cj->ci", R, sc.cams[:, 3:6])
    with open(os.path.join(path, "images.txt"), "w") as f:
        f.write("# Image list with two lines of data per image:\n")
        f.write("#   IMAGE_ID, QW, QX, QY, QZ, TX, TY, TZ, CAMERA_ID, NAME\n")
        f.write("#   POINTS2D[] as (X, Y, POINT3D_ID)\n")
        if extra_image:
            f.write(f"{int(iid.max()) + 5} 1 0 0 0 0 0 0 1 extra.png\n100.5 200.5 -1\n")
        for c in sorted(range(nC), key=lambda c: -int(iid[c])):
            vals = [*q[c], *t[c]]
            f.write(f"{int(iid[c])} " + " ".join(repr(float(v)) for v in vals)
                    + f" {c + 1} img{c}.png\n")
            sel = np.nonzero(sc.obs_cam == c)[0]
            sel = sel[np.argsort(-pid[sc.obs_pt[sel]], kind="stable")]
            f.write(" ".join(f"{float(uv[o, 0] + cx)!r} {float(uv[o, 1] + cy)!r} "
                             f"{int(pid[sc.obs_pt[o]])}" for o in sel) + "\n")
    with open(os.path.join(path, "points3D.txt"), "w") as f:
        f.write("# 3D point list with one line of data per point:\n")
        f.write("#   POINT3D_ID, X, Y, Z, R, G, B, ERROR, TRACK[] as (IMAGE_ID, POINT2D_IDX)\n")
        for p in sorted(range(nP), key=lambda p: -int(pid[p])):
            X = sc.points[p]
            f.write(f"{int(pid[p])} {float(X[0])!r} {float(X[1])!r} {float(X[2])!r} "
                    "128 128 128 0.5\n")


def write_bal(path: str, sc: Scene, order: str = "point"):
    """Write sc as a BAL problem file (bal_problem.cpp:111-143 reads it):
    "nc np nobs", one "cam pt x y" line per observation, then the 9 numbers of
    each camera [aa, C, f, k1, k2] and the 3 of each point, one per line.  x, y
    are the projections.  order "point" lists the observations point by point
    (the order of the published BAL problems), "camera" camera by camera, and
    "shuffle" in a fixed random order."""
    uv = project(sc.cams[sc.obs_cam], sc.points[sc.obs_pt])
    if order == "point":
        sel = np.lexsort((sc.obs_cam, sc.obs_pt))
    elif order == "camera":
        sel = np.lexsort((sc.obs_pt, sc.obs_cam))
    elif order == "shuffle":
        sel = np.random.default_rng(17).permutation(sc.num_obs)
    else:
        raise ValueError(order)
    with open(path, "w") as f:
        f.write(f"{sc.num_cams} {sc.num_points} {sc.num_obs}\n")
        for o in sel:
            f.write(f"{int(sc.obs_cam[o])} {int(sc.obs_pt[o])} "
                    f"{float(uv[o, 0])!r} {float(uv[o, 1])!r}\n")
        for v in np.concatenate([sc.cams.reshape(-1), sc.points.reshape(-1)]):
            f.write(f"{float(v)!r}\n")
    return sel
