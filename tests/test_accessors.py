"""cov::Uncertainty accessors (src/uncertaintyTE/uncertainty.cpp:36-185) pinned
by the arrays of the reference's accessor tests
(src/test_eigValues/covarianceEigenValues_test.cpp:81-137, fixture
eig_accessors_daliborka.json), on the golden output.cov, with the reference's
tolerances (absolute 1e-5 for matrices, relative 1e-5 for eigenvalues)."""
import json
import os

import numpy as np
import pytest

from uncertaintyte_amd import Uncertainty, jacob


@pytest.fixture(scope="module")
def unc(golden_dir):
    _, cam, pt = jacob.read_cov(os.path.join(golden_dir, "04_Daliborka_output.cov"))
    return Uncertainty(9, cam, pt)


@pytest.fixture(scope="module")
def ref(golden_dir):
    return json.load(open(os.path.join(golden_dir, "eig_accessors_daliborka.json")))


def test_one_camera_matrix(unc, ref):
    got = unc.getCameraUncMatrix(4).reshape(-1)
    assert np.abs(got - np.array(ref["compose_one_camera_covariance_matrix"])).max() < 1e-5


def test_one_point_matrix(unc, ref):
    got = unc.getPointUncMatrix(142).reshape(-1)
    assert np.abs(got - np.array(ref["compose_one_point_covariance_matrix"])).max() < 1e-5


def test_all_matrices(unc, ref):
    assert np.abs(unc.getCamerasUncMatrices().reshape(-1) -
                  np.array(ref["compose_camera_covariance_matrices"])).max() < 1e-5
    assert np.abs(unc.getPointsUncMatrices().reshape(-1) -
                  np.array(ref["compose_points_covariance_matrices"])).max() < 1e-5


def test_camera_eigenvalues(unc, ref):
    got = unc.getCamerasUncEigenValues().reshape(-1)
    want = np.array(ref["compute_camera_eigenvalues"])
    assert np.all(np.abs(got - want) <= np.abs(got) * 1e-5 + 1e-12)


def test_point_eigenvalues(unc, ref):
    got = unc.getPointsUncEigenValues().reshape(-1)
    want = np.array(ref["compute_points_eigenvalues"])
    tol = np.maximum(np.abs(got) * 1e-5, 1e-10)
    assert np.all(np.abs(got - want) <= tol)
