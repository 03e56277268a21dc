"""CPU: pins the Z^+ reference (oracle.zplus_reference / zplus_sampled) used by
the gauge-fixed precision modes.

The reference is D pinv_7(Z_s) D: the Moore-Penrose inverse of the scaled Schur
complement with its 7 smallest eigenvalues (the similarity gauge) dropped.
Checks: the Schur complement is the TE path's own (bit-identical), the gauge
really is 7-dimensional on the data (eigenvalue gap), the four Penrose
conditions on the range of Z, and agreement of two independent LAPACK/ARPACK
constructions (full eigh vs shift-invert Lanczos + Cholesky solve).
"""
import numpy as np
import pytest

import oracle
from uncertaintyte_amd import scene


@pytest.fixture(scope="module")
def dal_ref(daliborka):
    return oracle.zplus_reference(daliborka)


def test_front_is_the_te_paths_schur_complement(daliborka, daliborka_oracle):
    Z, d = oracle.front(daliborka)
    assert np.array_equal(Z, daliborka_oracle["Z"])
    assert np.array_equal(d, daliborka_oracle["d"])


def test_gauge_is_seven_dimensional(dal_ref):
    w = dal_ref["eig"]
    assert np.all(np.abs(w[:7]) < 1e-10 * w[-1])
    assert w[7] > 1e4 * np.abs(w[:7]).max()


def test_penrose_conditions(daliborka, dal_ref):
    Z, _ = oracle.front(daliborka)
    X = dal_ref["X"]
    nz, nx = np.linalg.norm(Z), np.linalg.norm(X)
    assert np.linalg.norm(X - X.T) <= 1e-12 * nx
    assert np.linalg.norm(Z @ X @ Z - Z) <= 1e-8 * nz           # rounding at cond ~ 2e7
    assert np.linalg.norm(X @ Z @ X - X) <= 1e-7 * nx           # cond(Z restricted) ~ 2e7
    P = Z @ X                                                    # projector onto range(Z)
    assert np.linalg.norm(P - P.T) <= 1e-6
    assert abs(np.trace(P) - (Z.shape[0] - 7)) < 1e-6


def test_camera_output_is_scaled_diagonal_block(daliborka, dal_ref):
    _, d = oracle.front(daliborka)
    X = dal_ref["X"]
    cp = daliborka.cam_params
    iu = np.triu_indices(cp)
    for a in (0, 17, 63):
        s = slice(a * cp, (a + 1) * cp)
        B = np.diag(d[s]) @ X[s, s] @ np.diag(d[s])
        assert np.allclose(dal_ref["cam"][a], B[iu], rtol=1e-14, atol=0)


@pytest.mark.parametrize("which", ["daliborka", "toy"])
def test_sampled_construction_agrees(which, daliborka):
    j = daliborka if which == "daliborka" else scene.make_workload("toy")
    ref = oracle.zplus_reference(j)
    cams = list(range(0, j.num_cams, max(1, j.num_cams // 7)))
    s, w = oracle.zplus_sampled(j, cams)
    rel = np.linalg.norm(s - ref["cam"][cams], axis=1) / np.linalg.norm(ref["cam"][cams], axis=1)
    assert rel.max() < 1e-9
    assert np.allclose(np.sort(w), ref["eig"][:7], atol=1e-14 * ref["eig"][-1])


def test_a_dropped_gauge_direction_is_detected(daliborka, dal_ref):
    """Planted mistake: pinv_6 (one gauge mode kept) must differ grossly."""
    Z, _ = oracle.front(daliborka)
    w, Q = np.linalg.eigh(Z)
    inv = np.zeros_like(w)
    inv[6:] = 1.0 / w[6:]
    X6 = (Q * inv) @ Q.T
    assert np.linalg.norm(X6 - dal_ref["X"]) > 1e3 * np.linalg.norm(dal_ref["X"])


def test_sampled_blocks_and_points_match_dense_reference():
    """zplus_sampled's camera blocks and the point step from them
    (points_from_blocks, the points seen only from the sampled cameras) agree
    with the dense eigh construction (zplus_reference) to the FP64 noise of two
    independent LAPACK/ARPACK constructions; a planted error in one camera
    block changes exactly the points that camera observes."""
    j = scene.scene_jacobian(scene.walk_scene(60, 20000, 5.0, seed=3))
    ref = oracle.zplus_reference(j)
    cams = list(range(10, 40))
    out, w, blk = oracle.zplus_sampled(j, cams, blocks=True)
    rel = np.linalg.norm(out - ref["cam"][cams], axis=1) / np.linalg.norm(ref["cam"][cams], axis=1)
    assert rel.max() < 1e-8
    pt, done = oracle.points_from_blocks(j, blk)
    m = done & (np.linalg.norm(ref["pt"], axis=1) > 0)
    assert m.sum() > 1000
    rp = np.linalg.norm(pt[m] - ref["pt"][m], axis=1) / np.linalg.norm(ref["pt"][m], axis=1)
    assert rp.max() < 1e-7
    assert np.all(pt[~done] == 0)
    bad = dict(blk)
    bad[20] = blk[20] * 1.01
    pt2, done2 = oracle.points_from_blocks(j, bad)
    changed = np.any(pt2 != pt, axis=1)
    rl = j.cam_params + 3
    obs_pt = (j.cols.reshape(-1, 2 * rl)[:, j.cam_params] - j.num_cams * j.cam_params) // 3
    obs_cam = j.cols.reshape(-1, 2 * rl)[:, 0] // j.cam_params
    seen20 = np.zeros(j.num_points, bool)
    seen20[obs_pt[obs_cam == 20]] = True
    assert np.array_equal(changed, seen20 & done)
