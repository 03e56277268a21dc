"""GPU: the gauge-fixed Z^+ modes against the FP64 pseudo-inverse reference.

Reference (oracle.zplus_reference): X = pinv_7(Z_s) from an FP64 eigen-
decomposition of the oracle's Schur complement (bit-identical to the GPU's),
camera output = upper triangles of D X D, points propagated from X's diagonal
blocks.  Contracts, relative Frobenius error per block:
  UNC_PREC_MIXED       <= 1e-5 for cameras and points;
  UNC_PREC_ZPLUS_FP64  <= 1e-8 for cameras, <= 1e-6 for points (short-baseline
                       points amplify the ~1e-11 camera differences by up to
                       1e4 through V^-1; both sides are FP64).
The recursive Cholesky + inverse kernels are checked directly against LAPACK.
"""
import os

import numpy as np
import pytest

import oracle
import uncertaintyte_amd as u
from uncertaintyte_amd import diag, jacob, scene

pytestmark = pytest.mark.gpu


def block_rel(a, b):
    a = np.asarray(a).reshape(len(a), -1)
    b = np.asarray(b).reshape(len(b), -1)
    nb = np.linalg.norm(b, axis=1)
    d = np.linalg.norm(a - b, axis=1)
    return np.where(nb > 0, d / np.where(nb > 0, nb, 1), d)


@pytest.mark.parametrize("n", [1, 37, 64, 65, 100, 128, 129, 130, 200, 777])
def test_chol_inverse_matches_lapack(n):
    rng = np.random.default_rng(n)
    B = rng.standard_normal((n, n))
    A = B @ B.T + n * np.eye(n)
    print(f"chol_inverse n={n}", flush=True)
    _, X = diag.chol_inverse(A)
    Lr = np.linalg.cholesky(A)
    # X = L^-1 is the output (the recursion keeps L only in its diagonal leaf
    # blocks); X Lr = I pins it to LAPACK's factor
    assert np.abs(np.triu(X, 1)).max() == 0.0
    R = X @ Lr - np.eye(n)
    assert np.abs(R).max() < 1e-12


def test_chol_inverse_rejects_indefinite():
    A = np.diag([1.0, -1.0, 2.0])
    with pytest.raises(u.UncError) as e:
        diag.chol_inverse(A)
    assert e.value.code == u.UNC_ERR_SINGULAR


CASES = {
    "daliborka": lambda: jacob.read_jacob(os.path.join(os.path.dirname(__file__), "golden",
                                                       "04_Daliborka.jacob")),
    "cube": lambda: jacob.read_jacob(os.path.join(os.path.dirname(__file__), "golden",
                                                  "01_Cube.jacob")),
    "rings200": lambda: scene.make_workload("rings200"),
    "walk300": lambda: scene.scene_jacobian(scene.walk_scene(300, 60000, 5.0, seed=3)),
}
TOL = {u.PREC_MIXED: (1e-5, 1e-5), u.PREC_ZPLUS_FP64: (1e-8, 1e-6)}


@pytest.fixture(scope="module", params=list(CASES))
def zcase(request):
    j = CASES[request.param]()
    return request.param, j, oracle.zplus_reference(j)


@pytest.mark.parametrize("prec", [u.PREC_MIXED, u.PREC_ZPLUS_FP64])
def test_zplus_modes(zcase, prec):
    name, j, ref = zcase
    print(f"zplus {name} precision={prec}", flush=True)
    r = u.compute_covariances(j, precision=prec)
    rc = block_rel(r.cam, ref["cam"])
    mask = np.linalg.norm(ref["pt"], axis=1) > 0
    rp = block_rel(r.pt[mask], ref["pt"][mask])
    print(f"  {name}: cam max {rc.max():.2e} median {np.median(rc):.2e}  pt max {rp.max():.2e}",
          flush=True)
    tc, tp = TOL[prec]
    assert rc.max() <= tc
    assert rp.max() <= tp
    assert np.all(r.pt[~mask] == 0)


def test_mixed_uses_int8_crt_products():
    """walk300 (N = 2700): the mixed path forms the top node's products on the int8
    CRT kernels (12 moduli by default, or as asked), the FP64 Z+ path does not, and
    the two differ by far less than the mixed contract."""
    j = CASES["walk300"]()
    print("mixed walk300: default moduli, 10 moduli, FP64", flush=True)
    m = u.compute_covariances(j, precision=u.PREC_MIXED)
    m10 = u.compute_covariances(j, precision=u.PREC_MIXED, moduli=10)
    f = u.compute_covariances(j, precision=u.PREC_ZPLUS_FP64)
    assert m.stats.zplus_moduli == 12 and m10.stats.zplus_moduli == 10
    assert f.stats.zplus_moduli == 0
    # the conditioning bound comes from each path's own Cholesky factor
    assert m.stats.zplus_cond_lb > 1.0
    assert abs(m.stats.zplus_cond_lb / f.stats.zplus_cond_lb - 1) < 1e-4
    d12 = block_rel(m.cam, f.cam).max()
    d10 = block_rel(m10.cam, f.cam).max()
    print(f"  12 moduli vs FP64 {d12:.2e}; 10 moduli {d10:.2e}", flush=True)
    assert 0 < d12 < 1e-7 and d12 < d10 < 1e-5


def test_mixed_odd_split_node():
    """441 cameras (N = 3969): the top node's halves are 1984 and 1985 rows, the
    second half the larger; the mixed path sizes its residue areas for it."""
    from uncertaintyte_amd import serve
    j = serve.request_scene(441, 60, 77)
    print("mixed 441 cameras: mixed vs FP64 Z+", flush=True)
    m = u.compute_covariances(j, precision=u.PREC_MIXED)
    f = u.compute_covariances(j, precision=u.PREC_ZPLUS_FP64)
    assert m.stats.zplus_moduli == 12
    d = block_rel(m.cam, f.cam).max()
    print(f"  mixed vs FP64 {d:.2e}", flush=True)
    assert d < 1e-6


def test_mixed_rejects_bad_moduli(daliborka):
    for m in (5, 21, -1):
        with pytest.raises(u.UncError) as e:
            u.compute_covariances(daliborka, precision=u.PREC_MIXED, moduli=m)
        assert e.value.code == u.UNC_ERR_BAD_INPUT


def test_zplus_repeatable_and_context_reuse(daliborka):
    ctx = u.Context()
    a = u.compute_covariances(daliborka, precision=u.PREC_ZPLUS_FP64, ctx=ctx.handle)
    b = u.compute_covariances(scene.make_workload("toy"), precision=u.PREC_ZPLUS_FP64,
                              ctx=ctx.handle)
    c = u.compute_covariances(daliborka, precision=u.PREC_ZPLUS_FP64, ctx=ctx.handle)
    assert np.array_equal(a.cam, c.cam) and np.array_equal(a.pt, c.pt)
    assert b.cam.shape[0] == 10
    ctx.close()


def _sampled_errors(j, r, cams, name):
    """Camera blocks of `cams` and the point blocks of every point whose
    observing cameras all lie in `cams`, against oracle.zplus_sampled (ARPACK +
    LAPACK Cholesky construction of D pinv_7(Z_s) D); also the FP64 noise floor
    of the oracle itself: its two constructions (the sampled one at cams) of the
    same quantity differ by the conditioning of Z, so the floor is stated next
    to the error."""
    print(f"zplus {name}: sampled oracle ({len(cams)} cameras)", flush=True)
    ref, w, blk = oracle.zplus_sampled(j, cams, nthreads=16, blocks=True,
                                       log=lambda m: print("  " + m, flush=True))
    rc = block_rel(r.cam[cams], ref)
    pt, done = oracle.points_from_blocks(j, blk)
    mask = done & (np.linalg.norm(pt, axis=1) > 0)
    rp = block_rel(r.pt[mask], pt[mask])
    print(f"  {name}: {len(cams)} cams max {rc.max():.2e} median {np.median(rc):.2e}; "
          f"{int(mask.sum())} points max {rp.max():.2e} median {np.median(rp):.2e}", flush=True)
    return rc, rp, int(mask.sum())


def test_zplus_at_scale_sampled_cameras():
    """walk1500 (1500 cameras, 1M points, 10% short-baseline; N = 13500): every
    8th camera plus a window of 200 consecutive cameras, and the ~100k points
    seen only from that window, against the sampled FP64 oracle (<= 1e-5)."""
    j = scene.make_workload("walk1500")
    print("zplus walk1500: GPU", flush=True)
    r = u.compute_covariances(j, precision=u.PREC_MIXED)
    cams = sorted(set(range(0, j.num_cams, 8)) | set(range(600, 800)))
    rc, rp, npts = _sampled_errors(j, r, cams, "walk1500")
    assert npts > 10000
    assert rc.max() <= 1e-5
    assert rp.max() <= 1e-5


def test_zplus_config5_large_scene():
    """Config 5: walk4000 (4000 cameras, 3M points, track Poisson(6); N = 36000,
    dense Z 10.4 GB) on one GPU; 200 consecutive cameras and the points seen
    only from them against the sampled FP64 oracle (<= 1e-5)."""
    print("zplus walk4000: generating", flush=True)
    j = scene.make_workload("walk4000")
    assert j.num_cams == 4000
    print(f"zplus walk4000: GPU (N = {j.num_cams * j.cam_params}, {j.num_obs} observations)",
          flush=True)
    r = u.compute_covariances(j, precision=u.PREC_MIXED)
    cams = list(range(2000, 2200))
    rc, rp, npts = _sampled_errors(j, r, cams, "walk4000")
    assert npts > 10000
    assert rc.max() <= 1e-5
    assert rp.max() <= 1e-5
