"""GPU: the reference's canonical run through the `uncertainty` CLI,
`uncertainty -in=04_Daliborka.jacob -out=...cov` (main_uncertainty.cpp:66-93,
saveResults IO.cpp:15-60), against the reference's own output file
test/EigValues/output.cov (tests/golden/04_Daliborka_output.cov): same header
lines (times aside), the printed lambda and all 14 Taylor cycle changes as
text, camera and point values to the 6 printed digits, zero rows at the fixed
points; and -debug's Z.txt / iZ.txt dumps in the reference's Matlab format."""
import os
import subprocess

import numpy as np
import pytest

import oracle
import uncertaintyte_amd as u
from uncertaintyte_amd import jacob

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
TIME_KEYS = {"Loading jacobian time", "Normalization of jacobian time",
             "Compose information matrix time", "Split infromation matrix time",
             "Inverse V time", "Compose Z time", "Inverse Z time", "Taylor expansion time",
             "Point uncertainty time", "Time of the algorithm"}


def _rel(a, b):
    nb = np.linalg.norm(b, axis=1)
    return np.linalg.norm(a - b, axis=1) / np.where(nb > 0, nb, 1.0)


def _run(args, cwd):
    print("cli:", " ".join(args), flush=True)
    r = subprocess.run([u.CLI_PATH] + args, cwd=cwd, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    return r.stdout


def test_canonical_run_matches_output_cov(tmp_path):
    out = tmp_path / "daliborka.cov"
    _run([f"-in={os.path.join(GOLDEN, '04_Daliborka.jacob')}", f"-out={out}",
          "-legacy_time_units"], str(tmp_path))
    h, cam, pt = jacob.read_cov(str(out))
    gh, gcam, gpt = jacob.read_cov(os.path.join(GOLDEN, "04_Daliborka_output.cov"))
    assert set(h) == set(gh)
    for k in gh:
        if k not in TIME_KEYS:
            assert h[k] == gh[k], k
    for k in TIME_KEYS:
        assert h[k].endswith("s") and float(h[k][:-1]) >= 0.0
    assert h["Used lambda"] == "3.67906e-07"
    assert h["TE number of iterations"] == "14"
    assert h["TE cycle change"] == gh["TE cycle change"]
    assert cam.shape == gcam.shape and pt.shape == gpt.shape
    assert _rel(cam, gcam).max() < 1e-5
    nz = np.linalg.norm(gpt, axis=1) > 0
    assert _rel(pt[nz], gpt[nz]).max() < 1e-5
    assert np.array_equal(np.linalg.norm(pt, axis=1) > 0, nz)


def _read_matlab(path, name):
    s = open(path).read().strip()
    assert s.startswith(name + " = [") and s.endswith("];")
    rows = s[len(name) + 4:-2].split(";")
    return np.array([[float(x) for x in r.split()] for r in rows])


def test_debug_writes_z_and_iz(tmp_path):
    out = tmp_path / "cube.cov"
    _run([f"-in={os.path.join(GOLDEN, '01_Cube.jacob')}", f"-out={out}", "-debug"],
         str(tmp_path))
    j = jacob.read_jacob(os.path.join(GOLDEN, "01_Cube.jacob"))
    o = oracle.pipeline(j, debug=True)
    Z = _read_matlab(tmp_path / "Z.txt", "Z")
    X = _read_matlab(tmp_path / "iZ.txt", "iZ")
    N = j.num_cams * j.cam_params
    assert Z.shape == (N, N) and X.shape == (N, N)
    # printed with 6 significant digits
    assert np.abs(Z - o["Z"]).max() <= 1e-5 * np.abs(o["Z"]).max()
    assert np.abs(X - o["X"]).max() <= 1e-5 * np.abs(o["X"]).max()
