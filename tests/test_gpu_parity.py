"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Contract (north star): the plain FP64 path matches the reference restatement
to <= 1e-9 relative Frobenius error per 9x9 camera block and per 3x3 point
block.  Stages before the dense inverse (fixPts, scales, blocks, 3x3
inverses, Schur) are required to be bit-identical to the oracle, which runs
the same fma sequence.
"""
import copy
import os
import sys

import numpy as np
import pytest

import oracle
import uncertaintyte_amd as u
from uncertaintyte_amd import diag, jacob, scene

pytestmark = pytest.mark.gpu
TOL_BLOCK = 1e-9


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def block_rel(a, b):
    a = np.asarray(a).reshape(len(a), -1)
    b = np.asarray(b).reshape(len(b), -1)
    nb = np.linalg.norm(b, axis=1)
    d = np.linalg.norm(a - b, axis=1)
    return np.where(nb > 0, d / np.where(nb > 0, nb, 1), d)


# ----------------------------------------------------------------- kernels

@pytest.mark.parametrize("M,N,K", [(64, 64, 64), (130, 77, 45), (300, 257, 129), (576, 576, 576),
                                   (1536, 1280, 640), (1537, 1283, 641), (1408, 1664, 1000)])
def test_gemm_f64(M, N, K):
    say(f"gemm {M}x{N}x{K}")
    rng = np.random.default_rng(M + N + K)
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((K, N))
    C0 = rng.standard_normal((M, N))
    C, m = diag.gemm(A, B, C0, alpha=-0.75, beta=0.5)
    ref = -0.75 * (A @ B) + 0.5 * C0
    assert np.abs(C - ref).max() <= 1e-13 * np.abs(A).max() * np.abs(B).max() * K
    assert m == np.abs(C).max()


@pytest.mark.parametrize("M,N,K", [(1800, 1800, 1800), (1537, 1283, 1031), (1280, 2560, 4099)])
def test_gemm_f64_split_k(M, N, K):
    """The split-K form (two workgroups per 128 x 128 tile, one per half of K,
    then the epilogue kernel adding the halves): the same product as the
    unsplit kernel to FP64 rounding, ragged edges and a K tail that is not a
    multiple of the 16-wide slab included, and max|C| from the epilogue."""
    say(f"gemm split {M}x{N}x{K}")
    rng = np.random.default_rng(M * 3 + N + K)
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((K, N))
    C, m = diag.gemm(A, B, rng.standard_normal((M, N)), alpha=-0.75, split=True)
    ref = -0.75 * (A @ B)
    assert np.abs(C - ref).max() <= 1e-13 * np.abs(A).max() * np.abs(B).max() * K
    assert m == np.abs(C).max()
    C1, _ = diag.gemm(A, B, alpha=-0.75)
    assert np.abs(C - C1).max() <= 1e-14 * np.abs(A).max() * np.abs(B).max() * K


@pytest.mark.parametrize("n", [1800, 1283, 2048])
def test_gemm_f64_symmetric_product(n):
    """The Taylor chain's product T X0 of two powers of one symmetric matrix:
    lower 128 x 128 tiles only, K in three or four chunks, the epilogue adding
    them in chunk order and mirroring the lower triangle.  C must be exactly
    symmetric, equal the full product to FP64 rounding (ragged n included),
    carry max|C| and be the same on a second run."""
    say(f"gemm sym {n}")
    rng = np.random.default_rng(n)
    Q = rng.standard_normal((n, n)) / np.sqrt(n)
    S = Q + Q.T                         # symmetric
    S2 = S @ S                          # a power of S: S2 S is symmetric
    S2 = 0.5 * (S2 + S2.T)
    C, m = diag.gemm_sym(S2, S, alpha=0.5)
    ref = 0.5 * (S2 @ S)
    assert np.array_equal(C, C.T)
    assert np.abs(C - ref).max() <= 1e-13 * np.abs(S2).max() * np.abs(S).max() * n
    assert m == np.abs(C).max()
    # the chunks are added in a fixed order: run to run the result is identical
    C2, m2 = diag.gemm_sym(S2, S, alpha=0.5)
    assert np.array_equal(C, C2) and m == m2


@pytest.mark.parametrize("n,nb", [(1800, 4), (1800, 2), (1283, 3), (2100, 4), (300, 1)])
def test_gemm_f64_symmetric_batch(n, nb):
    """The batched chain's products X0^a X0^(1..nb) in one launch: each C_b exactly
    symmetric (lower tiles formed, mirrored), equal to the full product to FP64
    rounding, max|C_b| from the epilogue, the same on a second launch."""
    say(f"gemm sym batch {n} x {nb}")
    rng = np.random.default_rng(n + nb)
    Q = rng.standard_normal((n, n)) / np.sqrt(n)
    S = Q + Q.T
    P = [S]
    for _ in range(3):
        T = P[-1] @ S
        P.append(0.5 * (T + T.T))
    A = P[3]
    Cs, mx, _ = diag.gemm_sym_batch(A, P[:nb])
    for b in range(nb):
        ref = A @ P[b]
        assert np.array_equal(Cs[b], Cs[b].T), b
        assert np.abs(Cs[b] - ref).max() <= 1e-13 * np.abs(A).max() * np.abs(P[b]).max() * n, b
        assert mx[b] == np.abs(Cs[b]).max(), b
    Cs2, mx2, _ = diag.gemm_sym_batch(A, P[:nb])
    assert all(np.array_equal(x, y) for x, y in zip(Cs, Cs2)) and mx == mx2


def test_gemm_f64_split_k_declines_small_shapes():
    A = np.ones((256, 1024))
    with pytest.raises(u.UncError) as e:
        diag.gemm(A, A.T.copy(), split=True)
    assert e.value.code == u.UNC_ERR_UNSUPPORTED
    with pytest.raises(u.UncError) as e:
        diag.gemm_sym(np.eye(512), np.eye(512))
    assert e.value.code == u.UNC_ERR_UNSUPPORTED


def _residual_case(n, seed=7):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = (Q * np.logspace(0, -8, n)) @ Q.T
    A = 0.5 * (A + A.T)
    return A, np.linalg.inv(A)


@pytest.mark.parametrize("n,tile_cols", [(1, 0), (70, 0), (150, 0), (150, 16), (150, 32),
                                         (333, 16), (333, 32)])
def test_residual_kernel_extraction(n, tile_cols):
    """Error-free extraction form on both tile shapes (64 x 16 and 64 x 32, n
    not a multiple of either): the extracted parts sum exactly, the rest is
    summed in FP64, so |R - R_exact| <= 16 n^3 eps^2 rowmax|A|_i colmax|X|_j
    (+ 2 eps |R_ij| for the last two roundings; 17 below for slack).  R_exact
    is the oracle's compensated (Dot2) residual.  Dropping the rest would
    leave errors near eps n rowmax colmax, eleven orders above this bound at
    n = 150."""
    say(f"residual (extraction) n={n} tile_cols={tile_cols}")
    A, X = _residual_case(n)
    R = diag.residual(A, X, tile_cols=tile_cols)
    Ro = oracle.residual_dd(A, X)
    eps = 2.0 ** -53
    bound = (17.0 * n ** 3 * eps ** 2 * np.outer(np.abs(A).max(axis=1), np.abs(X).max(axis=0))
             + 4 * eps * np.abs(Ro))
    assert np.all(np.abs(R - Ro) <= bound)
    # and in practice far inside it: 1e-10 of the largest entry (measured
    # 1.2e-11 at n = 333, whose residual is itself only 5.5e-10)
    assert np.abs(R - Ro).max() <= 1e-10 * np.abs(Ro).max() + 1e-300


@pytest.mark.parametrize("n,blocks", [(1, 1), (100, 1), (300, 1), (300, 3), (1100, 1),
                                      (1100, 4)])
def test_residual_int8_crt(n, blocks):
    """The Newton residual on the int8 matrix cores (residual_crt): rows of A and
    columns of X scaled to q = 77-bit integers, their exact product formed modulo
    20 moduli and rebuilt by the CRT, the identity subtracted in the residues.
    Bound against the compensated residual: 2^(3 - q) sqrt(n) |A_i.| |X_.j| (the
    cut) + 4 eps |R_ij| (the final roundings), plus the oracle's own
    17 n^3 eps^2 rowmax colmax.  A must be symmetric (its rows are cut as
    columns), X need not be.  The row blocks change no value."""
    say(f"residual (int8 crt) n={n} blocks={blocks}")
    A, X = _residual_case(n, seed=11)
    R, _ = diag.residual_crt(A, X, blocks=blocks)
    Ro = oracle.residual_dd(A, X)
    eps = 2.0 ** -53
    ra, cx = np.linalg.norm(A, axis=1), np.linalg.norm(X, axis=0)
    rmax, cmax = np.abs(A).max(axis=1), np.abs(X).max(axis=0)
    bound = (2.0 ** (3 - 77) * np.sqrt(n) * np.outer(ra, cx)
             + 17.0 * n ** 3 * eps ** 2 * np.outer(rmax, cmax) + 4 * eps * np.abs(Ro))
    assert np.all(np.abs(R - Ro) <= bound)
    if blocks > 1:
        R1, _ = diag.residual_crt(A, X, blocks=1)
        assert np.array_equal(R, R1)


def test_inverse_kernel():
    say("gj inverse")
    rng = np.random.default_rng(3)
    n = 200
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = (Q * np.logspace(0, -6, n)) @ Q.T
    A = 0.5 * (A + A.T)
    X = diag.inverse(A, refine=True)
    Xl = oracle.inv_lu(A)
    Xl = Xl + oracle.gemm(Xl, oracle.residual_dd(A, Xl))
    assert np.linalg.norm(X - Xl) / np.linalg.norm(Xl) < 1e-12


# ----------------------------------------------------------------- pipeline

def _cases():
    return {
        "cube": lambda: jacob.read_jacob(os.path.join(os.path.dirname(__file__), "golden", "01_Cube.jacob")),
        "daliborka": lambda: jacob.read_jacob(os.path.join(os.path.dirname(__file__), "golden", "04_Daliborka.jacob")),
        "toy10x1000": lambda: scene.scene_jacobian(scene.ring_scene(10, 1000, seed=1)),
        "rings40x4000": lambda: scene.scene_jacobian(scene.rings_scene(40, 4000, 4.0, seed=11)),
        "rings40x4000_shuffled": lambda: _shuffled(scene.scene_jacobian(scene.rings_scene(40, 4000, 4.0, seed=11))),
        # ~8000 observations per camera: the Schur kernel's global-table form
        "ring10x20000": lambda: scene.scene_jacobian(scene.ring_scene(10, 20000, seed=4)),
        # N = 1080: the Taylor chain in batches of four products (from N = 1024)
        "rings120x12000": lambda: scene.scene_jacobian(scene.rings_scene(120, 12000, 4.0, seed=13)),
    }


def _shuffled(j):
    """Same Jacobian with its observations (row pairs) in a random order."""
    rng = np.random.default_rng(5)
    rl = j.cam_params + 3
    perm = rng.permutation(j.num_rows // 2)
    j = copy.deepcopy(j)
    j.cols = j.cols.reshape(-1, 2 * rl)[perm].reshape(-1).copy()
    j.vals = j.vals.reshape(-1, 2 * rl)[perm].reshape(-1).copy()
    return j


@pytest.fixture(scope="module", params=list(_cases().keys()))
def case(request):
    j = _cases()[request.param]()
    o = oracle.pipeline(j, debug=True)
    return request.param, j, o


def test_pipeline_parity(case):
    name, j, o = case
    say(f"pipeline parity {name}: cams={j.num_cams} pts={j.num_points} obs={j.num_obs}")
    r = u.compute_covariances(j, debug=True)
    g = r.debug
    # bit-exact stages
    assert np.array_equal(g["d"], o["d"]), "column scales"
    assert np.array_equal(g["U_s"], o["U_s"]), "camera blocks"
    assert np.array_equal(g["Vinv_s"], o["Vinv_s"]), "3x3 inverses"
    assert np.array_equal(g["Z"], o["Z"]), "Schur complement"
    assert r.lambda_ == o["lambda_"]
    assert r.fixed == o["fixed"]
    rl = j.cam_params + 3
    obs_pt = (j.cols.reshape(-1, 2 * rl)[:, j.cam_params] - j.num_cams * j.cam_params) // 3
    assert r.stats.num_obs_kept == int(np.count_nonzero(~np.isin(obs_pt, list(j.fix))))
    # tolerance stages: X0 after one compensated Newton step is ~cond * err0^2
    ex0 = np.linalg.norm(g["X0"] - o["X0"]) / np.linalg.norm(o["X0"])
    rc = block_rel(r.cam, o["cam"])
    rp = block_rel(r.pt, o["pt"])
    say(f"  {name}: X0 rel {ex0:.2e}  cam max {rc.max():.2e}  pt max {rp.max():.2e}")
    assert ex0 < 1e-10, ex0
    assert r.n_terms == o["n_terms"]
    np.testing.assert_allclose(r.cycle_change, o["cycle_change"], rtol=1e-9)
    assert rc.max() <= TOL_BLOCK, rc.max()
    assert rp.max() <= TOL_BLOCK, rp.max()
    assert np.array_equal(r.pt == 0, o["pt"] == 0)


def test_config2_parity():
    """200 cameras on 4 rings, 100k points (the bench workload), vs the oracle."""
    say("config2 parity: generating")
    j = scene.make_workload("rings200")
    o = oracle.pipeline(j)
    say("config2 parity: gpu")
    r = u.compute_covariances(j)
    rc = block_rel(r.cam, o["cam"])
    rp = block_rel(r.pt, o["pt"])
    say(f"  config2: cam max {rc.max():.2e} median {np.median(rc):.2e}  pt max {rp.max():.2e}")
    assert r.n_terms == o["n_terms"]
    assert rc.max() <= TOL_BLOCK
    assert rp.max() <= TOL_BLOCK


def test_daliborka_golden_on_gpu(daliborka, golden_dir):
    say("daliborka vs output.cov")
    hdr, cam, pt = jacob.read_cov(os.path.join(golden_dir, "04_Daliborka_output.cov"))
    r = u.compute_covariances(daliborka)
    assert block_rel(r.cam, cam).max() < 1e-5
    mask = np.linalg.norm(pt, axis=1) > 0
    assert block_rel(r.pt[mask], pt[mask]).max() < 1e-5
    assert ["%g" % c for c in r.cycle_change] == hdr["TE cycle change"].split()


def test_unrefined_path_within_noise_floor(daliborka):
    say("unrefined inverse")
    o = oracle.pipeline(daliborka)
    r = u.compute_covariances(daliborka, refine=False)
    # plain FP64 inverse: cond(Z + lambda I) * eps ~ 1e-9 noise, no compensated step
    assert block_rel(r.cam, o["cam"]).max() < 1e-7


# ----------------------------------------------------------------- errors

def test_error_no_fixed_points(golden_dir):
    say("error: no fixed points")
    j = jacob.read_jacob(os.path.join(golden_dir, "03_Toy.jacob"))
    with pytest.raises(u.UncError) as e:
        u.compute_covariances(j)
    assert e.value.code == u.UNC_ERR_NO_FIXED_POINTS


def test_error_zero_column(cube):
    say("error: zero column")
    j = copy.deepcopy(cube)
    j.vals = j.vals.copy()
    j.vals[j.cols == 2 * 9 + 6] = 0.0
    with pytest.raises(u.UncError) as e:
        u.compute_covariances(j)
    assert e.value.code == u.UNC_ERR_SINGULAR


def test_error_duplicate_observation(cube):
    """A camera observing one point twice (two equal row pairs): refused, as the
    oracle refuses it (the Schur order takes one observation per camera and
    point)."""
    say("error: duplicate observation")
    j = copy.deepcopy(cube)
    rl = j.cam_params + 3
    cols = j.cols.reshape(-1, 2 * rl)
    vals = j.vals.reshape(-1, 2 * rl)
    pts = (cols[:, j.cam_params] - j.num_cams * j.cam_params) // 3
    o = int(np.flatnonzero(~np.isin(pts, list(j.fix)))[0])
    j.cols = np.concatenate([cols, cols[o:o + 1]]).reshape(-1).copy()
    j.vals = np.concatenate([vals, vals[o:o + 1]]).reshape(-1).copy()
    j.num_obs += 1
    j.num_rows += 2
    j.rows = (np.arange(j.num_rows + 1) * rl).astype(np.int32)
    with pytest.raises(RuntimeError):
        oracle.pipeline(j)
    with pytest.raises(u.UncError) as e:
        u.compute_covariances(j)
    assert e.value.code == u.UNC_ERR_UNSUPPORTED


def test_error_bad_layout(cube):
    say("error: bad layout")
    j = copy.deepcopy(cube)
    j.cols = j.cols.copy()
    j.cols[1], j.cols[2] = j.cols[2], j.cols[1]
    with pytest.raises(u.UncError) as e:
        u.compute_covariances(j)
    assert e.value.code == u.UNC_ERR_BAD_INPUT


def test_error_camera_out_of_range_regular_layout(cube):
    """An observation whose two rows keep the regular layout (cp consecutive camera
    columns, then 3 point columns) but name a camera past the last: the host-side
    layout check passes it on as ids, and the device's range check refuses it."""
    say("error: camera out of range, regular layout")
    j = copy.deepcopy(cube)
    cp, rl = j.cam_params, j.cam_params + 3
    cols = j.cols.reshape(-1, rl).copy()
    # the camera block of the first observation's two rows one camera past the end
    cols[0:2, :cp] = j.num_cams * cp + np.arange(cp)
    j.cols = cols.reshape(-1)
    with pytest.raises(u.UncError) as e:
        u.compute_covariances(j)
    assert e.value.code == u.UNC_ERR_BAD_INPUT


# ----------------------------------------------------------------- device API

def test_device_api_matches_host_api(daliborka):
    say("device API")
    import torch
    j = daliborka
    dev = torch.device("cuda:0")
    rows = torch.from_numpy(j.rows).to(dev)
    cols = torch.from_numpy(j.cols).to(dev)
    vals = torch.from_numpy(j.vals).to(dev)
    cam = torch.zeros(j.num_cams * 45, dtype=torch.float64, device=dev)
    pt = torch.zeros(j.num_points * 6, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    o = u.make_options(j)
    st = u.compute_device(o, j.num_rows, j.num_cols, rows.data_ptr(), cols.data_ptr(),
                          vals.data_ptr(), cam.data_ptr(), pt.data_ptr())
    r = u.compute_covariances(j)
    assert np.array_equal(cam.cpu().numpy().reshape(r.cam.shape), r.cam)
    assert np.array_equal(pt.cpu().numpy().reshape(r.pt.shape), r.pt)
    # the chain's products: T1 and one per accepted term
    assert st.num_terms == r.n_terms and st.gemm_count == 1 + r.n_terms
