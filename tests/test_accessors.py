"""cov::Uncertainty accessors (src/uncertaintyTE/uncertainty.cpp:36-185) on the
CPU: the product's unpack accessors (host index work, as in the reference) and
the oracle's eigen restatement (oracle/accessors.py), pinned by the arrays of
the reference's accessor tests (src/test_eigValues/covarianceEigenValues_test.cpp:
81-137, fixture eig_accessors_daliborka.json) on the golden output.cov, with
the reference's tolerances (absolute 1e-5 for matrices, relative 1e-5 for
eigenvalues), and by the eigen equation for the vectors."""
import json
import os

import numpy as np
import pytest

from oracle import accessors as acc
from uncertaintyte_amd import Uncertainty, jacob


@pytest.fixture(scope="module")
def packed(golden_dir):
    _, cam, pt = jacob.read_cov(os.path.join(golden_dir, "04_Daliborka_output.cov"))
    return np.asarray(cam, np.float64), np.asarray(pt, np.float64)


@pytest.fixture(scope="module")
def unc(packed):
    return Uncertainty(9, packed[0], packed[1])


@pytest.fixture(scope="module")
def ref(golden_dir):
    return json.load(open(os.path.join(golden_dir, "eig_accessors_daliborka.json")))


def test_one_camera_matrix(unc, ref):
    got = unc.getCameraUncMatrix(4).reshape(-1)
    assert np.abs(got - np.array(ref["compose_one_camera_covariance_matrix"])).max() < 1e-5


def test_one_point_matrix(unc, ref):
    got = unc.getPointUncMatrix(142).reshape(-1)
    assert np.abs(got - np.array(ref["compose_one_point_covariance_matrix"])).max() < 1e-5


def test_all_matrices(unc, ref, packed):
    assert np.abs(unc.getCamerasUncMatrices().reshape(-1) -
                  np.array(ref["compose_camera_covariance_matrices"])).max() < 1e-5
    assert np.abs(unc.getPointsUncMatrices().reshape(-1) -
                  np.array(ref["compose_points_covariance_matrices"])).max() < 1e-5
    # the oracle's unpack (written as the reference's loops) agrees exactly
    assert np.array_equal(acc.unpack_cameras(packed[0], 9), unc.getCamerasUncMatrices())
    assert np.array_equal(acc.unpack_points(packed[1]), unc.getPointsUncMatrices())
    assert np.array_equal(unc.getCamerasUncRaw(), packed[0].reshape(-1))


def test_oracle_camera_eigenvalues(packed, ref):
    got = acc.camera_eigenvalues(packed[0], 9).reshape(-1)
    want = np.array(ref["compute_camera_eigenvalues"])
    assert np.all(np.abs(got - want) <= np.abs(got) * 1e-5 + 1e-12)


def test_oracle_point_eigenvalues(packed, ref):
    got = acc.point_eigenvalues(packed[1]).reshape(-1)
    want = np.array(ref["compute_points_eigenvalues"])
    tol = np.maximum(np.abs(got) * 1e-5, 1e-10)
    assert np.all(np.abs(got - want) <= tol)


def test_oracle_eigenvalues_match_symmetric_solver(packed):
    M = acc.unpack_cameras(packed[0], 9)
    sym = np.linalg.eigvalsh(M)[:, ::-1]
    gen = acc.camera_eigenvalues(packed[0], 9)
    scale = np.abs(sym).max(axis=1, keepdims=True)
    assert np.abs(gen - sym).max() <= 1e-12 * scale.max()


def test_oracle_eigvecs_convention(packed):
    for M, V in ((acc.unpack_cameras(packed[0], 9)[:, 3:6, 3:6], acc.center_eigvecs(packed[0], 9)),
                 (acc.unpack_points(packed[1]), acc.point_eigvecs(packed[1]))):
        V = V.reshape(-1, 3, 3)                      # [block, k, component]
        lam = np.einsum("bki,bij,bkj->bk", V, M, V)   # Rayleigh quotients
        assert np.all(np.diff(lam, axis=1) <= 1e-12 * np.abs(lam).max())   # descending
        res = np.einsum("bij,bkj->bki", M, V) - lam[:, :, None] * V
        nrm = np.abs(M).max(axis=(1, 2))[:, None, None]
        assert np.all(np.abs(res) <= 1e-12 * np.maximum(nrm, 1e-300))
        assert np.allclose(np.einsum("bki,bli->bkl", V, V), np.eye(3), atol=1e-12)
        big = np.take_along_axis(V, np.argmax(np.abs(V), axis=2)[:, :, None], axis=2)
        assert np.all(big > 0)


def test_planted_mistake_is_caught(packed, ref):
    """Reading the packed triangle column by column (a plausible index slip)
    must fail the fixture."""
    c = packed[0].reshape(-1, 45)
    wrong = np.zeros((c.shape[0], 9, 9))
    t = 0
    for j in range(9):
        for i in range(j + 1):
            wrong[:, i, j] = wrong[:, j, i] = c[:, t]
            t += 1
    assert np.abs(wrong.reshape(-1) - np.array(ref["compose_camera_covariance_matrices"])).max() > 1e-5
