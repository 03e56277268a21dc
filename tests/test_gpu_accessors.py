"""GPU: the eigen accessors of cov::Uncertainty (src/uncertaintyTE/uncertainty.cpp:
74-115, 146-185) by the batched Jacobi kernels (unc_eig_cameras / unc_eig_points).

Checks, on the golden Daliborka output.cov and on synthetic blocks:
  - eigenvalues against the reference's own test arrays
    (covarianceEigenValues_test.cpp:117-137, eig_accessors_daliborka.json) at
    the reference's tolerance (relative 1e-5), and against the oracle
    (oracle/accessors.py, LAPACK) at 1e-13 of the block's largest eigenvalue;
  - eigenvectors by the eigen equation (residual <= 1e-12 of the block's
    norm), orthonormality (1e-12), the order (descending eigenvalues) and sign
    convention, and equality with the oracle's convention on blocks whose
    eigenvalues are distinct;
  - edge cases: zero blocks (the fixed points), repeated eigenvalues, cp 1..9,
    cp < 6 with centre vectors (BAD_INPUT), empty batches, ragged tiles.
"""
import json
import os

import numpy as np
import pytest

import uncertaintyte_amd as u
from oracle import accessors as acc
from uncertaintyte_amd import jacob

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def packed(golden_dir):
    _, cam, pt = jacob.read_cov(os.path.join(golden_dir, "04_Daliborka_output.cov"))
    return np.asarray(cam, np.float64).reshape(-1, 45), np.asarray(pt, np.float64).reshape(-1, 6)


@pytest.fixture(scope="module")
def ref(golden_dir):
    return json.load(open(os.path.join(golden_dir, "eig_accessors_daliborka.json")))


def _check_vectors(M, V, vals_desc):
    """M (b, 3, 3) symmetric, V (b, 9) vector k at [3k:3k+3]."""
    V = V.reshape(-1, 3, 3)
    nrm = np.maximum(np.abs(M).max(axis=(1, 2)), 1e-300)
    res = np.einsum("bij,bkj->bki", M, V) - vals_desc[:, :, None] * V
    assert np.all(np.abs(res).max(axis=(1, 2)) <= 1e-12 * nrm)
    assert np.abs(np.einsum("bki,bli->bkl", V, V) - np.eye(3)).max() <= 1e-12
    big = np.take_along_axis(V, np.argmax(np.abs(V), axis=2)[:, :, None], axis=2)
    assert np.all(big > 0)


def _distinct(vals, rel=1e-6):
    gaps = np.abs(np.diff(vals, axis=1)).min(axis=1)
    return gaps > rel * np.maximum(np.abs(vals).max(axis=1), 1e-300)


def test_camera_eigenvalues_match_reference_arrays(packed, ref):
    print("eig cameras daliborka", flush=True)
    ev, _ = u.eig_cameras(packed[0], 9, vectors=False)
    want = np.array(ref["compute_camera_eigenvalues"]).reshape(-1, 9)
    assert np.all(np.abs(ev - want) <= np.abs(ev) * 1e-5 + 1e-12)
    sym = np.linalg.eigvalsh(acc.unpack_cameras(packed[0], 9))[:, ::-1]
    scale = np.abs(sym).max(axis=1, keepdims=True)
    assert np.all(np.abs(ev - sym) <= 1e-13 * scale)


def test_point_eigenvalues_match_reference_arrays(packed, ref):
    print("eig points daliborka", flush=True)
    ev, _ = u.eig_points(packed[1], vectors=False)
    want = np.array(ref["compute_points_eigenvalues"]).reshape(-1, 3)
    tol = np.maximum(np.abs(ev) * 1e-5, 1e-10)
    assert np.all(np.abs(ev - want) <= tol)
    sym = np.linalg.eigvalsh(acc.unpack_points(packed[1]))[:, ::-1]
    scale = np.abs(sym).max(axis=1, keepdims=True)
    assert np.all(np.abs(ev - sym) <= 1e-13 * np.maximum(scale, 1e-300))


def test_uncertainty_class_routes_to_device(packed, ref):
    unc = u.Uncertainty(9, packed[0], packed[1])
    ev = unc.getCamerasUncEigenValues()
    want = np.array(ref["compute_camera_eigenvalues"]).reshape(-1, 9)
    assert np.all(np.abs(ev - want) <= np.abs(ev) * 1e-5 + 1e-12)
    pv = unc.getPointsUncEigenValues()
    assert pv.shape == (200, 3)
    assert unc.getCamerasCentersUncEigVec().shape == (64, 9)
    assert unc.getPointsUncEigVectors().shape == (200, 9)


def test_vectors_daliborka(packed):
    print("eig vectors daliborka", flush=True)
    pv, pvec = u.eig_points(packed[1])
    M = acc.unpack_points(packed[1])
    _check_vectors(M, pvec, pv)
    keep = _distinct(pv)
    assert keep.sum() == 197  # the three fixed points are zero blocks
    assert np.abs(pvec[keep] - acc.point_eigvecs(packed[1])[keep]).max() <= 1e-10
    # zero blocks: eigenvalues 0 and the unit vectors
    for i in np.nonzero(~keep)[0]:
        assert np.all(pv[i] == 0) and np.array_equal(pvec[i].reshape(3, 3), np.eye(3))
    cv, cvec = u.eig_cameras(packed[0], 9)
    C = acc.unpack_cameras(packed[0], 9)[:, 3:6, 3:6]
    cval = np.linalg.eigvalsh(C)[:, ::-1]
    _check_vectors(C, cvec, cval)
    keep = _distinct(cval)
    assert np.abs(cvec[keep] - acc.center_eigvecs(packed[0], 9)[keep]).max() <= 1e-10


def _spd(rng, k, n, cond):
    Q, _ = np.linalg.qr(rng.standard_normal((k, n, n)))
    w = np.exp(rng.uniform(0, np.log(cond), (k, n))) * 10.0 ** rng.uniform(-6, 4, (k, 1))
    return np.einsum("kij,kj,klj->kil", Q, w, Q)


@pytest.mark.parametrize("cp", [1, 2, 3, 5, 6, 7, 8, 9])
def test_cameras_all_block_sizes(cp):
    print(f"eig cameras cp={cp}", flush=True)
    rng = np.random.default_rng(cp)
    A = _spd(rng, 333, cp, 1e9)
    iu = np.triu_indices(cp)
    pk = A[:, iu[0], iu[1]]
    ev, vec = u.eig_cameras(pk, cp, vectors=cp >= 6)
    sym = np.linalg.eigvalsh(A)[:, ::-1]
    scale = np.abs(sym).max(axis=1, keepdims=True)
    assert np.all(np.abs(ev - sym) <= 1e-13 * scale)
    if cp >= 6:
        C = A[:, 3:6, 3:6]
        _check_vectors(C, vec, np.linalg.eigvalsh(C)[:, ::-1])
    else:
        with pytest.raises(u.UncError) as e:
            u.eig_cameras(pk, cp, vectors=True)
        assert e.value.code == u.UNC_ERR_BAD_INPUT


def test_indefinite_and_repeated_blocks():
    rng = np.random.default_rng(7)
    Q, _ = np.linalg.qr(rng.standard_normal((6, 3, 3)))
    w = np.array([[1, 1, 2], [3, 3, 3], [-1, 0, 1], [0, 0, 0], [1e-300, 1, 1e300 / 1e290],
                  [-5, -5, 2]], np.float64)
    M = np.einsum("kij,kj,klj->kil", Q, w, Q)
    M[3] = 0
    M = 0.5 * (M + np.transpose(M, (0, 2, 1)))
    pk = M[:, [0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2]]
    ev, vec = u.eig_points(pk)
    sym = np.linalg.eigvalsh(M)[:, ::-1]
    scale = np.maximum(np.abs(sym).max(axis=1, keepdims=True), 1e-300)
    assert np.all(np.abs(ev - sym) <= 1e-13 * scale)
    _check_vectors(M, vec, ev)


@pytest.mark.parametrize("n", [0, 1, 255, 256, 257, 100_003])
def test_points_ragged_batches(n):
    print(f"eig points n={n}", flush=True)
    rng = np.random.default_rng(n)
    if n == 0:
        ev, vec = u.eig_points(np.zeros((0, 6)))
        assert ev.shape == (0, 3) and vec.shape == (0, 9)
        return
    M = _spd(rng, n, 3, 1e8)
    pk = M[:, [0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2]]
    ev, vec = u.eig_points(pk)
    ev_only, _ = u.eig_points(pk, vectors=False)
    _, vec_only = u.eig_points(pk, values=False)
    assert np.array_equal(ev, ev_only) and np.array_equal(vec, vec_only)
    sym = np.linalg.eigvalsh(M)[:, ::-1]
    scale = np.abs(sym).max(axis=1, keepdims=True)
    assert np.all(np.abs(ev - sym) <= 1e-13 * scale)
    _check_vectors(M, vec, ev)
    keep = _distinct(ev)
    assert np.abs(vec[keep] - acc.point_eigvecs(pk)[keep]).max() <= 1e-8


def test_device_entry_points_with_torch():
    import torch
    rng = np.random.default_rng(11)
    n = 5000
    M = _spd(rng, n, 3, 1e6)
    pk = M[:, [0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2]]
    d_in = torch.from_numpy(np.ascontiguousarray(pk)).cuda()
    d_val = torch.empty((n, 3), dtype=torch.float64, device="cuda")
    d_vec = torch.empty((n, 9), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    s = torch.cuda.Stream()  # a real stream: handle 0 would mean the context's stream
    u.eig_points_device(n, d_in.data_ptr(), d_val.data_ptr(), d_vec.data_ptr(),
                        stream=s.cuda_stream)
    s.synchronize()
    ev, vec = u.eig_points(pk)
    assert np.array_equal(d_val.cpu().numpy(), ev) and np.array_equal(d_vec.cpu().numpy(), vec)
    A = _spd(rng, 64, 9, 1e9)
    iu = np.triu_indices(9)
    ck = np.ascontiguousarray(A[:, iu[0], iu[1]])
    c_in = torch.from_numpy(ck).cuda()
    c_val = torch.empty((64, 9), dtype=torch.float64, device="cuda")
    c_vec = torch.empty((64, 9), dtype=torch.float64, device="cuda")
    u.eig_cameras_device(64, 9, c_in.data_ptr(), c_val.data_ptr(), c_vec.data_ptr(),
                         stream=s.cuda_stream)
    s.synchronize()
    cv, cvec = u.eig_cameras(ck, 9)
    assert np.array_equal(c_val.cpu().numpy(), cv) and np.array_equal(c_vec.cpu().numpy(), cvec)


def test_pipeline_output_accessors(daliborka):
    """End to end: the device pipeline's own output through the accessors."""
    r = u.compute_covariances(daliborka)
    unc = u.Uncertainty(9, r.cam, r.pt)
    ev = unc.getCamerasUncEigenValues()
    sym = np.linalg.eigvalsh(acc.unpack_cameras(r.cam, 9))[:, ::-1]
    assert np.all(np.abs(ev - sym) <= 1e-13 * np.abs(sym).max(axis=1, keepdims=True))
    pv = unc.getPointsUncEigenValues()
    sym = np.linalg.eigvalsh(acc.unpack_points(r.pt))[:, ::-1]
    assert np.all(np.abs(pv - sym) <= 1e-13 * np.maximum(np.abs(sym).max(axis=1, keepdims=True),
                                                         1e-300))
