"""Benchmark: covariance solves/s of the plain FP64 pipeline on one MI355X.

Workload: 200 cameras on 4 rings, 100k points, track length max(2, Poisson(4))
(seeded synthetic scene, analytic RADIAL Jacobian: ~410k observations, CSR
J ~8.2e5 x 3.0e5 with 9.8e6 nonzeros, dense Z 1800 x 1800).  A step is one
complete solve (all cameras + all points) with the Jacobian already resident
in HBM; outputs stay on the device.

    python bench.py [--steps K] [--warmup W] [--no-cpu] [--no-zplus] [--no-serving]

Prints one JSON line.  `roofline` describes the dominant kernel (the FP64
MFMA product T X0 of the Taylor chain; it is symmetric, so only its lower
128 x 128 tiles are computed and N^2 (N + 1) flops are counted per launch)
from HIP events recorded by the library around each launch inside the timed
region; `cpu_baseline` times the CPU oracle (C, OpenMP) on one solve of the
same scene.

Two further measurements ride on the same line:
  zplus    the gauge-fixed Z+ path (UNC_PREC_MIXED) against the plain FP64 TE
           path on walk1500 (1500 cameras on a random walk, 1M points, track
           Poisson(5) min 2, 10% short-baseline points; N = 13491), both with
           the Jacobian resident in HBM, plus the camera-block error of the Z+
           output against D pinv_7(Z_s) D on 188 sampled cameras (oracle);
  accessors the eigen accessors on the device: 8M point blocks (3x3 values
           + vectors, HBM-bound) and 4000 camera blocks (9x9 values);
  serving  512 requests (cameras log-uniform 20..800, points per camera
           25..400, 64 distinct seeded scenes) arriving as a Poisson process
           into the request server (4 slots), p50/p99 latency and solves/s.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cov solves/s, mixed vs FP64 Z⁺ speedup at ≤1e-5 err; HBM GB/s; serving p50/p99"
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
WORKLOADS = {
    "rings200": "rings200: 200 cameras on 4 rings, 100k points, track Poisson(4) min 2",
    "walk1500": "walk1500: 1500 cameras on a random walk, 1M points, track Poisson(5) min 2, "
                "10% short-baseline points",
    "walk4000": "walk4000: 4000 cameras on a random walk, 3M points, track Poisson(6) min 2",
    "toy": "toy: 10 cameras on a ring, 1000 points",
}
PATHS = {"fp64": "FP64 plain TE path", "mixed": "gauge-fixed Z+ path (UNC_PREC_MIXED)",
         "zplus_fp64": "gauge-fixed Z+ path (UNC_PREC_ZPLUS_FP64)"}


def _device_solver(u, torch, dev, J, precision):
    rows = torch.from_numpy(J.rows).to(dev)
    cols = torch.from_numpy(J.cols).to(dev)
    vals = torch.from_numpy(J.vals).to(dev)
    per = J.cam_params * (J.cam_params + 1) // 2
    cam = torch.zeros(J.num_cams * per, dtype=torch.float64, device=dev)
    pt = torch.zeros(J.num_points * 6, dtype=torch.float64, device=dev)
    ctx = u.Context(0)
    opts = u.make_options(J, precision)

    def call():
        return u.compute_device(opts, J.num_rows, J.num_cols, rows.data_ptr(), cols.data_ptr(),
                                vals.data_ptr(), cam.data_ptr(), pt.data_ptr(), ctx=ctx)
    return call, ctx, cam, pt


def _timed(call, torch, reps):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        st = call()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps * 1e3, st


def _zplus_errors(J, cams, zcam_h, zpt_h, with_points, log=None):
    """Camera blocks of `cams` (and, with_points, the point blocks of every point
    seen only from `cams`) against the sampled FP64 oracle D pinv_7(Z_s) D."""
    import numpy as np
    import oracle
    nth = min(os.cpu_count() or 1, 16)
    out = {}
    if with_points:
        ref, _, blk = oracle.zplus_sampled(J, cams, nthreads=nth, blocks=True, log=log)
        pref, done = oracle.points_from_blocks(J, blk)
        mask = done & (np.linalg.norm(pref, axis=1) > 0)
    else:
        ref, _ = oracle.zplus_sampled(J, cams, nthreads=nth, log=log)
    for name, cam_h, pt_h in zip(("mixed", "zplus_fp64"), zcam_h, zpt_h):
        rel = np.linalg.norm(cam_h[cams] - ref, axis=1) / np.linalg.norm(ref, axis=1)
        o = {"cam_err_max": float(rel.max()), "cam_err_median": float(np.median(rel))}
        if with_points:
            pr = np.linalg.norm(pt_h[mask] - pref[mask], axis=1) / np.linalg.norm(pref[mask], axis=1)
            o.update(pt_err_max=float(pr.max()), pt_err_median=float(np.median(pr)),
                     points_checked=int(mask.sum()))
        out[name] = o
    return out


def _zplus_pair(u, torch, dev, J, reps):
    """Time the mixed (int8 CRT products) and the FP64 Z+ paths on device-resident
    inputs; returns ms, stats and host copies of both outputs."""
    res = {}
    for name, prec in (("mixed", u.PREC_MIXED), ("zplus_fp64", u.PREC_ZPLUS_FP64)):
        call, ctx, cam, pt = _device_solver(u, torch, dev, J, prec)
        call()
        ms, st = _timed(call, torch, reps)
        res[name] = (ms, st, cam.cpu().numpy().reshape(J.num_cams, -1),
                     pt.cpu().numpy().reshape(J.num_points, -1))
        ctx.close()
    return res


def measure_zplus(u, scene, torch, dev, check=True):
    """Config 3, walk1500: the mixed-precision Z+ path (int8 CRT products in the
    Cholesky recursion) against the FP64 Z+ path and the plain FP64 TE path,
    inputs resident in HBM; camera and point-block errors against the sampled
    FP64 oracle."""
    J = scene.make_workload("walk1500")
    out = {"workload": "walk1500: 1500 cameras on a random walk, 1M points, track Poisson(5) "
                       "min 2, 10% short-baseline points", "N": J.num_cams * J.cam_params,
           "observations": J.num_obs}
    res = _zplus_pair(u, torch, dev, J, 3)
    (z_ms, zst, zc, zp), (g_ms, gst, gc, gp) = res["mixed"], res["zplus_fp64"]
    fcall, fctx, _, _ = _device_solver(u, torch, dev, J, u.PREC_FP64)
    fcall()
    f_ms, fst = _timed(fcall, torch, 1)
    fctx.close()
    out.update(zplus_ms=round(z_ms, 2), zplus_fp64_ms=round(g_ms, 2),
               speedup_vs_fp64_zplus=round(g_ms / z_ms, 3),
               fp64_te_ms=round(f_ms, 2), speedup=round(f_ms / z_ms, 2),
               zplus_inverse_ms=round(zst.stage_ms[6], 2),
               zplus_fp64_inverse_ms=round(gst.stage_ms[6], 2),
               fp64_inverse_plus_taylor_ms=round(fst.stage_ms[6] + fst.stage_ms[7], 2),
               moduli=zst.zplus_moduli, cond_lb=float(zst.zplus_cond_lb),
               arithmetic="mixed: Cholesky selected inversion, products of the recursion nodes "
                          f">= 2048 rows exact on int8 MFMA modulo {zst.zplus_moduli} moduli "
                          "(CRT, 46-bit scaled operands), the rest FP64 MFMA; zplus_fp64: all FP64")
    if check:
        cams = sorted(set(range(0, J.num_cams, 8)) | set(range(600, 800)))
        err = _zplus_errors(J, cams, (zc, gc), (zp, gp), True)
        out.update(errors=err, cam_err_max=err["mixed"]["cam_err_max"],
                   pt_err_max=err["mixed"]["pt_err_max"],
                   err_sample=f"{len(cams)} cameras (every 8th + 600..799) and the points seen "
                              "only from them vs D pinv_7(Z_s) D (ARPACK + LAPACK oracle)",
                   contract=1e-5)
    return out


def measure_large(u, scene, torch, dev, fp64=False, check=True):
    """Config 5, walk4000 (4000 cameras, 3M points, track Poisson(6) min 2;
    N = 36000): the mixed and FP64 Z+ paths, optionally the plain FP64 TE path,
    and the camera errors on 200 sampled cameras against the oracle."""
    def log(msg):
        print(f"[large] {msg}", file=sys.stderr, flush=True)
    t0 = time.time()
    J = scene.make_workload("walk4000")
    log(f"scene generated ({time.time() - t0:.0f} s)")
    out = {"workload": "walk4000: 4000 cameras on a random walk, 3M points, track Poisson(6) min 2",
           "N": J.num_cams * J.cam_params, "observations": J.num_obs,
           "scene_gen_s": round(time.time() - t0, 1),
           "workspace_GiB": round(u.workspace_bytes(u.make_options(J, u.PREC_MIXED), J.num_rows)
                                  / 2 ** 30, 2)}
    res = _zplus_pair(u, torch, dev, J, 2)
    (z_ms, zst, zc, zp), (g_ms, gst, gc, gp) = res["mixed"], res["zplus_fp64"]
    log(f"Z+ mixed {z_ms:.0f} ms, FP64 {g_ms:.0f} ms per solve")
    out.update(zplus_ms=round(z_ms, 1), zplus_fp64_ms=round(g_ms, 1),
               speedup_vs_fp64_zplus=round(g_ms / z_ms, 3), moduli=zst.zplus_moduli,
               cond_lb=float(zst.zplus_cond_lb),
               zplus_stages_ms={k: round(float(zst.stage_ms[i]), 2)
                                for i, k in enumerate(["upload", "decode", "point_lists", "assemble",
                                                       "r", "schur", "inverse", "taylor", "cam_out",
                                                       "point_out", "d", "total"])
                                if zst.stage_ms[i] > 0})
    if fp64:
        fcall, fctx, _, _ = _device_solver(u, torch, dev, J, u.PREC_FP64)
        f_ms, fst = _timed(fcall, torch, 1)
        fctx.close()
        log(f"plain FP64 path {f_ms:.0f} ms")
        out.update(fp64_te_ms=round(f_ms, 1), speedup=round(f_ms / z_ms, 2),
                   fp64_taylor_terms=fst.num_terms)
    if check:
        cams = list(range(0, J.num_cams, 20))
        err = _zplus_errors(J, cams, (zc, gc), (zp, gp), False, log=log)
        out.update(errors=err, cam_err_max=err["mixed"]["cam_err_max"],
                   err_sample=f"{len(cams)} cameras vs D pinv_7(Z_s) D (ARPACK + LAPACK oracle)")
    return out


def measure_accessors(u, torch, dev, n_points=8_000_000, n_cams=4000, reps=20):
    """Eigen accessors (cov::Uncertainty, uncertainty.cpp:74-185) on device-
    resident packed outputs: 3x3 eigenvalues + eigenvectors of n_points point
    blocks (48 B in, 24 + 72 B out per point) and 9x9 eigenvalues + centre
    eigenvectors of n_cams camera blocks.  HIP events on the launch stream."""
    g = torch.Generator(device=dev)
    g.manual_seed(5)
    # SPD blocks B B^T + 1e-3 I, packed (xx, xy, xz, yy, yz, zz); elementwise
    # products (no batched GEMM) keep the set-up simple
    B = torch.randn((n_points, 3, 3), dtype=torch.float64, device=dev, generator=g)
    pk = torch.stack([(B[:, i, :] * B[:, j, :]).sum(-1) + (1e-3 if i == j else 0.0)
                      for i in range(3) for j in range(i, 3)], 1).contiguous()
    del B
    vals = torch.empty((n_points, 3), dtype=torch.float64, device=dev)
    vecs = torch.empty((n_points, 9), dtype=torch.float64, device=dev)
    C = torch.randn((n_cams, 9, 12), dtype=torch.float64, device=dev, generator=g)
    ck = torch.stack([(C[:, i, :] * C[:, j, :]).sum(-1) for i in range(9) for j in range(i, 9)],
                     1).contiguous()
    cvals = torch.empty((n_cams, 9), dtype=torch.float64, device=dev)
    cvecs = torch.empty((n_cams, 9), dtype=torch.float64, device=dev)
    # a stream of its own (torch's default stream is the null stream, handle 0,
    # which the library reads as "the context's stream"); events on it see the
    # kernels
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    pcall = lambda: u.eig_points_device(n_points, pk.data_ptr(), vals.data_ptr(), vecs.data_ptr(),
                                        stream=s.cuda_stream)
    ccall = lambda: u.eig_cameras_device(n_cams, 9, ck.data_ptr(), cvals.data_ptr(),
                                         cvecs.data_ptr(), stream=s.cuda_stream)
    out = {}
    for name, call in (("points", pcall), ("cameras", ccall)):
        call()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        for _ in range(reps):
            call()
        e1.record(s)
        torch.cuda.synchronize()
        out[name + "_ms"] = e0.elapsed_time(e1) / reps
    bytes_pt = 144 * n_points
    gbs = bytes_pt / (out["points_ms"] * 1e-3) / 1e9
    res = {"kernel": "k_eig_points<true,true> (3x3 cyclic Jacobi, one lane per point)",
           "points": n_points, "points_ms": round(out["points_ms"], 4),
           "points_per_s": round(n_points / (out["points_ms"] * 1e-3), 0),
           "algorithmic_bytes_per_point": 144, "achieved_GBps": round(gbs, 1),
           "peak_GBps": PEAK_HBM_GBS, "frac": round(gbs / PEAK_HBM_GBS, 4),
           "cameras": n_cams, "cameras_ms": round(out["cameras_ms"], 4),
           "camera_kernel": "k_eig_cameras<9> (9x9 Jacobi eigenvalues + 3x3 centre vectors)"}
    del pk, vals, vecs, C, ck, cvals, cvecs
    torch.cuda.empty_cache()
    return res


def measure_serving(u, rate=150.0, n_requests=512):
    from uncertaintyte_amd import serve
    scenes, order = serve.make_requests(n_requests, 64, workers=min(os.cpu_count() or 1, 16),
                                        pinned=True)
    # two slots: at 150 req/s the GPU is ~70 % busy, and a third or fourth
    # request in flight mostly stretches the largest ones (measured p50 / p99
    # 3.2 / 24.8 ms with 2 slots, 4.7 / 29-33 ms with 4)
    srv = u.Server(slots=2)
    # warm state: every distinct scene once, so the slots' cached buffers and the
    # kernels' code objects are in place before the timed load
    for tk in [srv.submit(s, precision=u.PREC_MIXED) for s in scenes]:
        try:
            srv.wait(tk)
        except u.UncError:
            pass
    res = serve.run_load(srv, scenes, order, rate=rate, precision=u.PREC_MIXED)
    st = srv.stats()
    srv.close()
    out = {k: (round(v, 3) if isinstance(v, float) else v) for k, v in res.summary().items()}
    out.update(rate_per_s=rate, slots=2, precision="Z+ (UNC_PREC_MIXED)",
               peak_held_GiB=round(st.peak_held_bytes / 2 ** 30, 2),
               max_concurrent=st.max_concurrent,
               workload="512 requests, cameras log-uniform 20..800, points/camera 25..400, "
                        "64 distinct seeded walk scenes, Poisson arrivals; request payloads in "
                        "page-locked host memory (unc_host_alloc), uploaded per request")
    return out


def copy_calibration() -> int:
    """A plain copy of known size: 1 GiB read + 1 GiB written per launch."""
    import torch
    n = (1 << 30) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read_per_launch": 1 << 30, "copy_bytes_written_per_launch": 1 << 30,
                      "launches": 5}), flush=True)
    return 0


def gemm_traffic():
    """HBM bytes per Taylor-chain product (GEMM + epilogue) measured offline with
    rocprofv3 --pmc (FETCH_SIZE and WRITE_SIZE, each scaled by the copy
    calibration); recorded in profiles/gemm_batch_rings200_traffic.md as
    'traffic_bytes_per_launch: <int>'."""
    path = os.path.join(ROOT, "profiles", "gemm_batch_rings200_traffic.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def cpu_reference_arithmetic(J, threads):
    """The reference's TE arithmetic on the host, timed: fixPts, J^T J blocks,
    3x3 inverses and the Schur complement (the C oracle, OpenMP, `threads`),
    then SDM::inv -- row and column 2-norm equilibration, LAPACK dgetrf +
    dgetri, unscaling (ScaledDenseMatrix.cpp:150-279) -- and the Taylor chain
    by dgemm with the reference's coefficients and stopping rule
    (compute.cpp:121-160), both through the host BLAS (scipy's OpenBLAS,
    `threads` threads), then the camera output and point step (oracle).  No
    compensated refinement.  Returns (seconds, n_terms, BLAS name)."""
    import numpy as np
    import scipy.linalg.blas as blas
    import scipy.linalg.lapack as lapack
    from threadpoolctl import threadpool_info, threadpool_limits

    import oracle
    t0 = time.perf_counter()
    Z, d = oracle.front(J, nthreads=threads)
    N = Z.shape[0]
    lam = oracle.auto_lambda(N) if J.lam == -1 else J.lam
    with threadpool_limits(limits=threads):
        A = np.asfortranarray(Z + lam * np.eye(N))
        sL = np.sqrt((A * A).sum(axis=1))
        A /= sL[:, None]
        sR = np.sqrt((A * A).sum(axis=0))
        A /= sR[None, :]
        lu, piv, info = lapack.dgetrf(A, overwrite_a=True)
        Xi, info2 = lapack.dgetri(lu, piv, overwrite_lu=True)
        if info or info2:
            raise RuntimeError("LAPACK inverse failed")
        X0 = np.asfortranarray(Xi / sR[:, None] / sL[None, :])
        X = X0.copy(order="F")
        T = blas.dgemm(1.0, X0, X0)
        old, terms = np.finfo(np.float64).max, 0
        for i in range(1, 20):
            k = lam ** i / float(oracle.factorial32(i - 1))
            change = abs(k * max(np.abs(T).max(), np.finfo(np.float64).tiny))
            if change < 1e-5 or change > old:
                break
            old = change
            terms += 1
            X += k * T
            T = blas.dgemm(1.0, T, X0)
    cp = J.cam_params
    iu = np.triu_indices(cp)
    blocks = {}
    cam = np.empty((J.num_cams, cp * (cp + 1) // 2))
    for a in range(J.num_cams):
        sl = slice(a * cp, (a + 1) * cp)
        B = X[sl, sl]
        cam[a] = ((B + B.T / 2) * d[sl][:, None] * d[sl][None, :])[iu]
        blocks[a] = B
    oracle.points_from_blocks(J, blocks)
    secs = time.perf_counter() - t0
    blas_name = ", ".join(sorted({f"{x.get('internal_api')} {x.get('version')}"
                                  for x in threadpool_info() if x.get("user_api") == "blas"}))
    return secs, terms, blas_name


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="rings200")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed", "zplus_fp64"],
                    help="path of the main measurement (the headline value is fp64)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-zplus", action="store_true")
    ap.add_argument("--no-check", action="store_true", help="skip the Z+ accuracy check")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-accessors", action="store_true")
    ap.add_argument("--rate", type=float, default=150.0, help="serving arrival rate (1/s)")
    ap.add_argument("--large", action="store_true",
                    help="also run the 4000-camera / 3M-point scene (N = 36000), minutes")
    ap.add_argument("--large-fp64", action="store_true",
                    help="with --large, also time the plain FP64 TE path on it (about a minute)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy 5 times (the reference kernel for "
                         "scaling rocprofv3 FETCH_SIZE / WRITE_SIZE counts to bytes)")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the pipeline is single-GPU",
              file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import uncertaintyte_amd as u
    from uncertaintyte_amd import scene

    t0 = time.time()
    J = scene.make_workload(args.workload)
    gen_s = time.time() - t0
    dev = torch.device("cuda:0")
    rows = torch.from_numpy(J.rows).to(dev)
    cols = torch.from_numpy(J.cols).to(dev)
    vals = torch.from_numpy(J.vals).to(dev)
    per = J.cam_params * (J.cam_params + 1) // 2
    cam = torch.zeros(J.num_cams * per, dtype=torch.float64, device=dev)
    pt = torch.zeros(J.num_points * 6, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    ctx = u.Context(0)
    prec = {"fp64": u.PREC_FP64, "mixed": u.PREC_MIXED, "zplus_fp64": u.PREC_ZPLUS_FP64}[args.precision]
    opts = u.make_options(J, prec)
    call = lambda: u.compute_device(opts, J.num_rows, J.num_cols, rows.data_ptr(), cols.data_ptr(),
                                    vals.data_ptr(), cam.data_ptr(), pt.data_ptr(), ctx=ctx)
    for _ in range(args.warmup):
        call()
    torch.cuda.synchronize()
    gemm_ms = 0.0
    gemm_flops = 0.0
    gemm_n = 0
    batch_ms = batch_flops = 0.0
    batch_n = 0
    stage = np.zeros(12)
    terms = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = call()
        gemm_ms += st.gemm_ms
        gemm_flops += st.gemm_flops
        gemm_n += st.gemm_count
        batch_ms += st.batch_ms
        batch_flops += st.batch_flops
        batch_n += st.batch_count
        stage += np.array(st.stage_ms[:])
        terms = st.num_terms
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    value = args.steps / el
    N = J.num_cams * J.cam_params
    # The dominant kernel: the chain's launches of four products each
    # (k_gemm_sym_batch, where the chain is batched), algorithmic flops
    # 4 N^2 (N + 1) per launch (the symmetric form forms the lower tiles only).
    # Where no such launch ran, all of the chain's product launches.
    if batch_n > 0:
        flops = batch_flops / batch_n
        avg_gemm_s = batch_ms / batch_n * 1e-3
        launches = batch_n
        kernel = ("Taylor-chain powers, four per launch: k_gemm_sym_batch (X0^4j times X0^1..4, "
                  "lower 128x128 tiles, whole K per workgroup, mirrored in the epilogue)")
    else:
        flops = gemm_flops / max(gemm_n, 1)
        avg_gemm_s = gemm_ms / max(gemm_n, 1) * 1e-3
        launches = gemm_n
        kernel = "Taylor-chain product launches (k_gemm_f64w8 + epilogue)"
    achieved = flops / avg_gemm_s / 1e12 if avg_gemm_s > 0 else 0.0
    chain_tflops = gemm_flops / (gemm_ms * 1e-3) / 1e12 if gemm_ms > 0 else 0.0
    stage /= args.steps
    names = [u.lib().unc_stage_name(i).decode() for i in range(12)]
    nC, nO, nP = J.num_cams, J.num_obs, J.num_points
    nnz = J.nnz
    asm_bytes = 12 * nnz + 4 * (2 * nO + 1) + 8 * (81 * nC + 27 * nO + 9 * nP)
    schur_bytes = 8 * (81 * nC + 27 * nO + 9 * nP) + 8 * N * N

    cpu = None
    parity = None
    if not args.no_cpu:
        import oracle
        threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        # the reference's arithmetic with the host BLAS (the baseline)
        cs, cterms, blas_name = cpu_reference_arithmetic(J, threads)
        # the reference hard-codes 8 OpenMP / Eigen threads (compute.cpp:460-463)
        cs8, _, _ = cpu_reference_arithmetic(J, 8)
        # the parity oracle (C restatement with a compensated Newton step and its
        # own blocked dgemm), whose output is the GPU's parity reference
        t1 = time.perf_counter()
        o_ref = oracle.pipeline(J, nthreads=threads)
        co = time.perf_counter() - t1
        model = ""
        try:
            for line in open("/proc/cpuinfo"):
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
        cpu = {"value": round(1.0 / cs, 4), "unit": "solves/s", "cores": threads, "kind": "port",
               "value_8_threads": round(1.0 / cs8, 4), "cpu_model": model,
               "nproc": os.cpu_count(), "blas": blas_name, "taylor_terms": cterms,
               "parity_oracle_solves_per_s": round(1.0 / co, 4),
               "sample": f"1 full solve of the same {args.workload} scene with the reference's "
                         f"arithmetic ({cs:.2f} s on {threads} threads, {cs8:.2f} s on 8): C "
                         f"oracle front end (fixPts, blocks, Schur; OpenMP), equilibrated LAPACK "
                         f"dgetrf/dgetri and the dgemm Taylor chain through {blas_name}, outputs"}
        # parity of the GPU result of the last timed solve against the oracle,
        # beside the FP64 noise floor (the reference-shaped unrefined LU inverse
        # against the refined oracle on the same Z)
        if args.precision == "fp64":
            o_lu = oracle.pipeline(J, inverse_mode=oracle.INV_REFERENCE, nthreads=threads)
            gcam = cam.cpu().numpy().reshape(nC, per)
            gpt = pt.cpu().numpy().reshape(nP, 6)

            def brel(a, b):
                nb = np.linalg.norm(b, axis=1)
                m = nb > 0
                return (np.linalg.norm(a - b, axis=1)[m] / nb[m])

            parity = {"cam_max": float(brel(gcam, o_ref["cam"]).max()),
                      "pt_max": float(brel(gpt, o_ref["pt"]).max()),
                      "noise_floor_cam_max": float(brel(o_lu["cam"], o_ref["cam"]).max()),
                      "noise_floor_pt_max": float(brel(o_lu["pt"], o_ref["pt"]).max()),
                      "contract": 1e-9, "terms_equal": terms == o_ref["n_terms"],
                      "against": "C oracle (refined FP64 inverse); noise floor = unrefined "
                                 "equilibrated-LU oracle vs refined oracle"}

    # the main measurement's context and tensors are released first: the
    # request server's streams then have the device's hardware queues to
    # themselves
    ctx.close()
    del rows, cols, vals, cam, pt
    torch.cuda.empty_cache()
    # serving before the Z+ accuracy check: the CPU oracle's BLAS threads would
    # otherwise compete with the request threads
    accessors = None if args.no_accessors else measure_accessors(u, torch, dev)
    serving = None if args.no_serving else measure_serving(u, rate=args.rate)
    zplus = None if args.no_zplus else measure_zplus(u, scene, torch, dev, check=not args.no_check)
    large = measure_large(u, scene, torch, dev, fp64=args.large_fp64,
                          check=not args.no_check) if args.large else None

    line = {
        "metric": METRIC, "value": round(value, 3), "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * el / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded RADIAL-camera scene, analytic Jacobian)",
        "config": {"workload": WORKLOADS.get(args.workload, args.workload) + ", " + PATHS[args.precision],
                   "cameras": nC, "points": nP, "observations": nO,
                   "jacobian_nnz": nnz, "N": N, "taylor_terms": terms, "parallelism": "1 GPU"},
        "roofline": {"bound": "mfma", "kernel": kernel, "achieved": round(achieved, 3),
                     "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP64_TFLOPS, 4), "traffic": gemm_traffic(),
                     "per_launch_flops": flops, "avg_launch_ms": round(avg_gemm_s * 1e3, 4),
                     "launches": launches, "chain_launches": gemm_n,
                     "chain_TFLOPs": round(chain_tflops, 3)},
        "cpu_baseline": cpu,
        "parity": parity,
        "stages_ms": {names[i]: round(float(stage[i]), 4) for i in range(12) if stage[i] > 0},
        "hbm": {"assemble_GBps": round(asm_bytes / (stage[3] * 1e-3) / 1e9, 1) if stage[3] > 0 else None,
                "schur_GBps": round(schur_bytes / (stage[5] * 1e-3) / 1e9, 1) if stage[5] > 0 else None,
                "peak_GBps": PEAK_HBM_GBS},
        "scene_gen_s": round(gen_s, 2),
        "zplus": zplus,
        "accessors": accessors,
        "serving": serving,
        "large": large,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
