"""The newer C++ interface (include/uncertaintyTE/uncertainty.h, the reference's
src/uncertaintyTE/uncertainty.h:24-158): struct layouts, the EAlgorithm names,
the fixed-point choice of getCovariances (JacobianComposer.cpp:63-83), the
unpack accessors against the reference's accessor arrays, and (GPU) a full
getCovariances call against the C ABI on the same inputs.  Driven through
tests/cpp/te_driver.cpp, compiled against the installed headers and library."""
import itertools
import json
import os
import struct
import subprocess

import numpy as np
import pytest

import uncertaintyte_amd as u
from uncertaintyte_amd import jacob, scene

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    exe = tmp_path_factory.mktemp("te") / "te_driver"
    subprocess.run(["g++", "-std=c++14", "-O1", "-I", INC,
                    os.path.join(ROOT, "tests", "cpp", "te_driver.cpp"), "-o", str(exe), "-L",
                    os.path.dirname(u.LIB_PATH), "-luncertainty_amd",
                    "-Wl,-rpath," + os.path.dirname(u.LIB_PATH)], check=True)
    return str(exe)


def _write_input(path, jac, points):
    with open(path, "wb") as f:
        np.array([jac.num_cams, jac.cam_params, jac.num_points, jac.num_obs, jac.num_rows,
                  jac.num_cols], np.int32).tofile(f)
        np.asarray(jac.rows, np.int32).tofile(f)
        np.asarray(jac.cols, np.int32).tofile(f)
        np.asarray(jac.vals, np.float64).tofile(f)
        np.asarray(points, np.float64).reshape(-1).tofile(f)


def _read_arrays(path):
    out = []
    with open(path, "rb") as f:
        while True:
            h = f.read(8)
            if not h:
                return out
            n = struct.unpack("<q", h)[0]
            out.append(np.frombuffer(f.read(8 * n), np.float64))


def test_layout_and_algorithm_names(driver):
    r = subprocess.run([driver, "layout"], capture_output=True, text=True, check=True)
    lines = r.stdout.split("\n")
    # double, double, enum, 7 ints, pointer, bool: the reference's member order
    assert [int(x) for x in lines[0].split()][:4] == [64, 16, 48, 56]
    assert lines[1] == "SVD_DEVIDE_AND_CONQUER 2" and lines[2] == "throws"


def test_fixed_point_choice_is_the_most_distant_triple(driver, tmp_path):
    """40 points: 100000 random triples cover all 9880 triples, so the choice is
    the brute-force maximum of the perimeter (the reference's criterion)."""
    rng = np.random.default_rng(4)
    P = rng.standard_normal((40, 3))
    jac = jacob.Jacobian(-1.0, 1, 9, 40, 0, (0, 1, 2), 0, 9 + 120, np.zeros(1, np.int32),
                         np.zeros(0, np.int32), np.zeros(0))
    p = tmp_path / "in.bin"
    _write_input(p, jac, P)
    r = subprocess.run([driver, "fix", str(p)], capture_output=True, text=True, check=True)
    f = r.stdout.split()
    fix = [int(f[1]), int(f[2]), int(f[3])]
    assert int(f[5]) == u.UNC_ERR_BAD_INPUT  # the empty Jacobian is then refused

    def per(t):
        a, b, c = (P[i] for i in t)
        return np.linalg.norm(a - b) + np.linalg.norm(a - c) + np.linalg.norm(b - c)
    best = max(itertools.combinations(range(40), 3), key=per)
    assert sorted(fix) == sorted(best)
    assert fix[0] < fix[1]  # the reference sorts the first two ids only


def test_unpack_accessors_match_reference_arrays(driver, tmp_path, golden_dir):
    _, cam, pt = jacob.read_cov(os.path.join(golden_dir, "04_Daliborka_output.cov"))
    ref = json.load(open(os.path.join(golden_dir, "eig_accessors_daliborka.json")))
    p = tmp_path / "cov.bin"
    with open(p, "wb") as f:
        np.array([64, 9, 200], np.int32).tofile(f)
        np.asarray(cam, np.float64).reshape(-1).tofile(f)
        np.asarray(pt, np.float64).reshape(-1).tofile(f)
    o = tmp_path / "out.bin"
    r = subprocess.run([driver, "unpack", str(p), str(o)], capture_output=True, text=True,
                       check=True)
    one_cam, one_pt, cams, pts, raw_c, raw_p = _read_arrays(o)
    assert np.abs(one_cam - ref["compose_one_camera_covariance_matrix"]).max() < 1e-5
    assert np.abs(one_pt - ref["compose_one_point_covariance_matrix"]).max() < 1e-5
    assert np.abs(cams - ref["compose_camera_covariance_matrices"]).max() < 1e-5
    assert np.abs(pts - ref["compose_points_covariance_matrices"]).max() < 1e-5
    assert np.array_equal(raw_c, np.asarray(cam).reshape(-1))
    assert np.array_equal(raw_p, np.asarray(pt).reshape(-1))
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if not has_gpu:
        # the eigen accessors run on the device and say so when there is none
        assert "threw 1" in r.stdout and "HIP runtime error" in r.stdout


@pytest.mark.gpu
def test_get_covariances_with_points_matches_c_abi(driver, tmp_path):
    sc = scene.ring_scene(12, 800, seed=9)
    jac = scene.scene_jacobian(sc)
    p = tmp_path / "in.bin"
    _write_input(p, jac, sc.points)
    o = tmp_path / "out.bin"
    print("te_driver run ring12x800", flush=True)
    r = subprocess.run([driver, "run", str(p), str(o)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    f = r.stdout.split()
    fix = (int(f[3]), int(f[4]), int(f[5]))
    assert (int(f[7]), int(f[8]), int(f[9])) == fix
    cam, pt, cev, cvec, pev, pvec = _read_arrays(o)
    want = u.compute_covariances(jac, fix=fix)
    assert np.array_equal(cam, want.cam.reshape(-1)) and np.array_equal(pt, want.pt.reshape(-1))
    assert float(f[11]) == want.lambda_ and int(f[13]) == want.n_terms
    ev, vec = u.eig_cameras(want.cam, 9)
    assert np.array_equal(cev, ev.reshape(-1)) and np.array_equal(cvec, vec.reshape(-1))
    ev, vec = u.eig_points(want.pt)
    assert np.array_equal(pev, ev.reshape(-1)) and np.array_equal(pvec, vec.reshape(-1))
