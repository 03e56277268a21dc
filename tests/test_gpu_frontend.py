"""GPU: the scene Jacobian kernel (unc_scene_jacobian; JacobianComposer::
scene2Jacobian, JacobianComposer.cpp:84-194, with analytic derivatives of
snavely_reprojection_error.h:148-347) and the COLMAP and BAL paths end to end.

Checks: for each of the four camera models, the CRS structure exactly (two rows
per observation, cp + 3 columns: camera block then point block) and the values
against the oracle's central differences of the reference's residual functors
(oracle/scene_fd.py, 1e-6 of each row's largest entry; the differencing error
is ~1e-9); for RADIAL also against the generator's analytic blocks at 1e-12.
Out-of-range indices come back as UNC_ERR_BAD_INPUT from the kernel's own
bounds check.  End to end: a COLMAP text model and a BAL problem file through
the CLI give the .cov that unc_compute gives on the same Jacobian and fixed
points."""
import os
import subprocess

import numpy as np
import pytest

import uncertaintyte_amd as u
from oracle import scene_fd
from uncertaintyte_amd import jacob, scene

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("model", [0, 1, 2, 3])
def test_scene_jacobian_matches_finite_differences(model):
    print(f"scene jacobian model {model}", flush=True)
    sc = scene.ring_scene(9, 300, seed=7)
    # a near-identity rotation, placed so that it still looks at the points
    sc.cams[0, 0:3] = [1e-9, -2e-9, 5e-10]
    sc.cams[0, 3:6] = [0.0, 0.0, -3.0]
    J = u.scene_jacobian(sc.cams, sc.points, sc.obs_cam, sc.obs_pt, model)
    cp = u.CAM_PARAMS[model]
    nobs = sc.num_obs
    rl = cp + 3
    assert np.array_equal(J.rows, np.arange(2 * nobs + 1) * rl)
    cols = J.cols.reshape(nobs, 2, rl)
    ecam = sc.obs_cam[:, None] * cp + np.arange(cp)[None, :]
    ept = sc.num_cams * cp + sc.obs_pt[:, None] * 3 + np.arange(3)[None, :]
    for r in range(2):
        assert np.array_equal(cols[:, r, :cp], ecam) and np.array_equal(cols[:, r, cp:], ept)
    got = J.vals.reshape(nobs, 2, rl)
    fd = scene_fd.jacobian_fd(sc.cams, sc.points, sc.obs_cam, sc.obs_pt, model)
    scale = np.abs(fd).max(axis=2, keepdims=True)
    assert np.all(np.abs(got - fd) <= 1e-6 * scale)
    if model == 3:
        Jc, Jx = scene.jacobian_blocks(sc.cams[sc.obs_cam], sc.points[sc.obs_pt])
        ref = np.concatenate([Jc, Jx], -1)
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_scene_jacobian_rejects_bad_indices():
    sc = scene.ring_scene(4, 40, seed=2)
    op = sc.obs_pt.copy()
    op[5] = sc.num_points          # one past the end
    with pytest.raises(u.UncError) as e:
        u.scene_jacobian(sc.cams, sc.points, sc.obs_cam, op, u.CAM_RADIAL)
    assert e.value.code == u.UNC_ERR_BAD_INPUT


def test_scene_jacobian_device_entry():
    import torch
    sc = scene.ring_scene(6, 120, seed=4)
    dev = torch.device("cuda:0")
    cams = torch.from_numpy(np.ascontiguousarray(sc.cams)).to(dev)
    pts = torch.from_numpy(np.ascontiguousarray(sc.points)).to(dev)
    oc = torch.from_numpy(sc.obs_cam.astype(np.int32)).to(dev)
    op = torch.from_numpy(sc.obs_pt.astype(np.int32)).to(dev)
    n = sc.num_obs
    rows = torch.empty(2 * n + 1, dtype=torch.int32, device=dev)
    cols = torch.empty(2 * n * 12, dtype=torch.int32, device=dev)
    vals = torch.empty(2 * n * 12, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    rc = u.lib().unc_scene_jacobian_device(None, u.CAM_RADIAL, sc.num_cams, sc.num_points, n,
                                           cams.data_ptr(), pts.data_ptr(), oc.data_ptr(),
                                           op.data_ptr(), rows.data_ptr(), cols.data_ptr(),
                                           vals.data_ptr(), None)
    assert rc == u.UNC_OK
    J = u.scene_jacobian(sc.cams, sc.points, sc.obs_cam, sc.obs_pt, u.CAM_RADIAL)
    assert np.array_equal(rows.cpu().numpy(), J.rows) and np.array_equal(cols.cpu().numpy(), J.cols)
    assert np.array_equal(vals.cpu().numpy(), J.vals)


def test_colmap_to_cov_through_the_cli(tmp_path):
    sc = scene.ring_scene(12, 600, seed=9)
    d = tmp_path / "model"
    scene.write_colmap(str(d), sc, "RADIAL")
    out = tmp_path / "out.cov"
    print("cli colmap ring12x600", flush=True)
    r = subprocess.run([u.CLI_PATH, f"-in={d}", "-in_form=COLMAP", "-cam=RADIAL", f"-out={out}"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    hdr, cam, pt = jacob.read_cov(str(out))
    cams, pts, oc, op = u.read_colmap(str(d))
    fix = u.find_pts2fix(pts)
    J = u.scene_jacobian(cams, pts, oc, op, u.CAM_RADIAL, fix=fix)
    res = u.compute_covariances(J)
    assert int(hdr["Number of cameras"]) == sc.num_cams
    assert int(hdr["Number of points in 3D"]) == sc.num_points
    want_c, want_p = res.cam.reshape(-1), res.pt.reshape(-1)
    got_c, got_p = np.asarray(cam, np.float64).reshape(-1), np.asarray(pt, np.float64).reshape(-1)
    # the .cov holds 6 significant digits
    assert np.all(np.abs(got_c - want_c) <= 1e-5 * np.abs(want_c) + 1e-300)
    assert np.all(np.abs(got_p - want_p) <= 1e-5 * np.abs(want_p) + 1e-300)


def _block_rel(a, b):
    a = np.asarray(a).reshape(len(a), -1)
    b = np.asarray(b).reshape(len(b), -1)
    nb = np.linalg.norm(b, axis=1)
    d = np.linalg.norm(a - b, axis=1)
    return np.where(nb > 0, d / np.where(nb > 0, nb, 1), d)


def test_bal_file_to_covariances(tmp_path):
    """A BAL problem (observations point by point, as the published BAL files
    list them) read by unc_scene_read_bal, its Jacobian by the device kernel in
    the file's row order against the oracle's central differences of the
    reference's functor, the covariances against the oracle (bit-exact Schur
    complement), the CLI's -in_form=BAL .cov against unc_compute, and the same
    problem with its observations in other orders to 1e-9."""
    import oracle
    sc = scene.ring_scene(12, 600, seed=9)
    path = tmp_path / "problem.txt"
    scene.write_bal(str(path), sc, "point")
    cams, pts, oc, op = u.read_bal(str(path))
    print("bal ring12x600: jacobian", flush=True)
    J = u.scene_jacobian(cams, pts, oc, op, u.CAM_RADIAL)
    fd = scene_fd.jacobian_fd(cams, pts, oc, op, u.CAM_RADIAL)
    got = J.vals.reshape(len(oc), 2, 12)
    assert np.all(np.abs(got - fd) <= 1e-6 * np.abs(fd).max(axis=2, keepdims=True))
    print("bal ring12x600: covariances", flush=True)
    o = oracle.pipeline(J, debug=True)
    r = u.compute_covariances(J, debug=True)
    assert np.array_equal(r.debug["U_s"], o["U_s"]) and np.array_equal(r.debug["Z"], o["Z"])
    assert r.n_terms == o["n_terms"]
    assert _block_rel(r.cam, o["cam"]).max() <= 1e-9
    assert _block_rel(r.pt, o["pt"]).max() <= 1e-9
    print("bal ring12x600: cli", flush=True)
    out = tmp_path / "bal.cov"
    p = subprocess.run([u.CLI_PATH, f"-in={path}", "-in_form=BAL", f"-out={out}"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    hdr, cam, pt = jacob.read_cov(str(out))
    assert int(hdr["Number of cameras"]) == sc.num_cams
    got_c, got_p = np.asarray(cam, np.float64).reshape(-1), np.asarray(pt, np.float64).reshape(-1)
    want_c, want_p = r.cam.reshape(-1), r.pt.reshape(-1)
    assert np.all(np.abs(got_c - want_c) <= 1e-5 * np.abs(want_c) + 1e-300)
    assert np.all(np.abs(got_p - want_p) <= 1e-5 * np.abs(want_p) + 1e-300)
    # the same problem with its observations camera by camera and shuffled:
    # only the row order differs, so only the rounding may
    for order in ("camera", "shuffle"):
        p2 = tmp_path / f"{order}.txt"
        scene.write_bal(str(p2), sc, order)
        r2 = u.compute_covariances(u.scene_jacobian(*u.read_bal(str(p2)), u.CAM_RADIAL, fix=J.fix))
        assert _block_rel(r2.cam, r.cam).max() <= 1e-9, order
        assert _block_rel(r2.pt, r.pt).max() <= 1e-9, order
