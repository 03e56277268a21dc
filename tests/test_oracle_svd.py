"""The oracle of the SVD algorithm modes (oracle/svd_modes.py; the reference's
svdInverse, compute.cpp:185-249) pinned on a small synthetic scene whose 7-dim
gauge is well separated from the rest of the spectrum of M = J^T J:
the two LAPACK drivers agree with each other and with an eigen-decomposition,
the result satisfies the Penrose conditions on the kept subspace, and the point
formula equals the literal pinv(A) B Sigma B^T pinv(A)^T of compute.cpp:516-527."""
import numpy as np
import pytest

from oracle import svd_modes
from uncertaintyte_amd import scene


@pytest.fixture(scope="module")
def small():
    return scene.scene_jacobian(scene.ring_scene(6, 250, seed=11))


def test_gauge_is_separated(small):
    J = svd_modes.dense_jacobian(small)
    ev = np.linalg.eigvalsh(J.T @ J)
    assert ev[7] > 1e4 * max(abs(ev[6]), 1e-300)


def test_drivers_agree_and_match_eigh(small):
    J = svd_modes.dense_jacobian(small)
    M = J.T @ J
    a, s, _ = svd_modes.pinv_svd(M, 0, svd_remove_n=7)
    b, _, _ = svd_modes.pinv_svd(M, 1, svd_remove_n=7)
    w, V = np.linalg.eigh(M)
    keep = np.argsort(-np.abs(w))[: M.shape[0] - 7]
    e = (V[:, keep] / w[keep]) @ V[:, keep].T
    scale = np.abs(e).max()
    # both LAPACK drivers reduce to the same bidiagonal problem; eigh is another
    # algorithm, so it agrees to cond(kept spectrum) * eps = 2.3e10 * 2.2e-16
    assert np.abs(a - b).max() <= 1e-9 * scale
    assert np.abs(a - e).max() <= 2e-5 * scale
    # epsilon mode with a threshold inside the gap keeps the same values
    c, _, keep_c = svd_modes.pinv_svd(M, 0, svd_remove_n=-1, epsilon=np.sqrt(s[-7] * s[-8]))
    assert keep_c.sum() == M.shape[0] - 7 and np.abs(c - a).max() <= 1e-12 * scale


def test_penrose_conditions(small):
    J = svd_modes.dense_jacobian(small)
    M = J.T @ J
    X, _, _ = svd_modes.pinv_svd(M, 1, svd_remove_n=7)
    # residuals at cond(kept spectrum) * eps (2.3e10 * 2.2e-16), as for LAPACK itself
    assert np.abs(M @ X @ M - M).max() <= 1e-11 * np.abs(M).max()
    assert np.abs(X @ M @ X - X).max() <= 2e-5 * np.abs(X).max()
    assert np.abs(X @ M - (X @ M).T).max() <= 5e-5
    P = X @ M                                     # projector onto the kept subspace
    assert abs(np.trace(P) - (M.shape[0] - 7)) < 1e-4


def test_point_formula_is_the_pinv_form(small):
    cam, pt, iZ, _ = svd_modes.svd_mode(small, 1, svd_remove_n=7)
    cp, N = small.cam_params, small.num_cams * small.cam_params
    rl = cp + 3
    vals = np.asarray(small.vals).reshape(-1, 2, rl)
    cols = np.asarray(small.cols).reshape(-1, 2, rl)
    for p in (0, 17, 133):
        obs = np.nonzero((cols[:, 0, cp] - N) // 3 == p)[0]
        A = np.concatenate([vals[o, :, cp:] for o in obs])          # 2n x 3
        k = len(obs)
        B = np.zeros((2 * k, k * (cp + 2)))
        Sig = np.zeros((k * (cp + 2), k * (cp + 2)))
        for t, o in enumerate(obs):
            c = cols[o, 0, 0] // cp
            B[2 * t:2 * t + 2, t * (cp + 2):t * (cp + 2) + cp] = vals[o, :, :cp]
            B[2 * t, t * (cp + 2) + cp] = -1
            B[2 * t + 1, t * (cp + 2) + cp + 1] = -1
            Sig[t * (cp + 2):t * (cp + 2) + cp, t * (cp + 2):t * (cp + 2) + cp] = \
                iZ[c * cp:(c + 1) * cp, c * cp:(c + 1) * cp]
            Sig[t * (cp + 2) + cp, t * (cp + 2) + cp] = 1
            Sig[t * (cp + 2) + cp + 1, t * (cp + 2) + cp + 1] = 1
        pA = np.linalg.pinv(A)
        S = pA @ B @ Sig @ B.T @ pA.T
        want = S[[0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2]]
        assert np.abs(pt[p] - want).max() <= 1e-9 * np.abs(want).max()


def test_planted_mistake_is_caught(small):
    """Keeping one singular value too many (pinv_6) must break the agreement."""
    J = svd_modes.dense_jacobian(small)
    M = J.T @ J
    a, _, _ = svd_modes.pinv_svd(M, 0, svd_remove_n=7)
    b, _, _ = svd_modes.pinv_svd(M, 0, svd_remove_n=6)
    assert np.abs(a - b).max() > 1e3 * np.abs(a).max()
