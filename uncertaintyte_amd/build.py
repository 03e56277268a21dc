"""Build recipe for the gfx950 library and CLI (hipcc, in-tree output).

    python -m uncertaintyte_amd.build          # -> uncertaintyte_amd/_build/

Produces
  _build/libuncertainty_amd.so       C ABI (include/uncertainty_amd.h) + the C++
                                     getCovariances (include/uncertainty.h)
  _build/libuncertainty_amd_diag.so  the same kernels plus the kernel-level test
                                     hooks of csrc/diag.h (tests only)
  _build/uncertainty                 CLI with the reference's flags
"""
from __future__ import annotations

import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libuncertainty_amd.so")
CLI = os.path.join(OUT, "uncertainty")

HIP_SOURCES = ["k_assemble.hip", "k_sort.hip", "k_schur.hip", "k_dense.hip", "k_output.hip", "k_zplus.hip", "k_eig.hip", "k_scene.hip", "k_svd.hip", "k_crt_gemm.hip", "pipeline.hip"]
# MFMA accumulators in VGPR form: the loop-carried accumulators of the GEMM
# then stay in place instead of being copied between AGPRs and VGPRs each slab.
PER_FILE_FLAGS = {"k_dense.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
                  "k_zplus.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}
CPP_SOURCES = ["uncertainty_api.cpp", "uncertainty_te.cpp", "scene_io.cpp", "server.cpp"]
DIAG_SOURCES = ["diag.hip"]
DIAG_LIB = os.path.join(OUT, "libuncertainty_amd_diag.so")

ARCH = os.environ.get("UNC_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: the bit-exact stages spell out every fused multiply-add.
# -packed-fp32-ops off: packed FP32 instructions are avoided on this platform.
HIPCC_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-Wno-unused-command-line-argument", "-I", os.path.join(ROOT, "include")]


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.isabs(c) and os.path.exists(c) or not os.path.isabs(c)):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build failed: " + " ".join(cmd[:4]) + " ...")
    return r.stdout


def build(verbose: bool = False, jobs: int = 8) -> str:
    os.makedirs(OUT, exist_ok=True)
    hipcc = _hipcc()
    t0 = time.time()
    objs = []
    diag_objs = []
    procs = []
    for src in HIP_SOURCES + CPP_SOURCES + DIAG_SOURCES:
        obj = os.path.join(OUT, os.path.splitext(src)[0] + ".o")
        lang = [] if src.endswith(".hip") else ["-x", "hip"]
        cmd = [hipcc] + HIPCC_FLAGS + PER_FILE_FLAGS.get(src, []) + lang + \
            ["-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                            text=True)))
        (diag_objs if src in DIAG_SOURCES else objs).append(obj)
        while len([p for _, p in procs if p.poll() is None]) >= jobs:
            time.sleep(0.05)
    failed = False
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed = True
            sys.stderr.write(out)
        elif verbose and out.strip():
            sys.stderr.write(out)
    if failed:
        raise RuntimeError("HIP compilation failed")
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    tmp = DIAG_LIB + ".tmp"
    _run([hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + diag_objs)
    os.replace(tmp, DIAG_LIB)
    _run([hipcc, "-O2", "-std=c++17", "-x", "hip", f"--offload-arch={ARCH}",
          "-I", os.path.join(ROOT, "include"), os.path.join(CSRC, "cli.cpp"),
          "-o", CLI, "-L", OUT, "-luncertainty_amd", "-Wl,-rpath,$ORIGIN"])
    if verbose:
        print(f"built {LIB} in {time.time() - t0:.1f}s")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
