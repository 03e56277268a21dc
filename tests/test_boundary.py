"""CPU-side checks of the drop-in boundary: the C-ABI library loads and exports
every symbol the headers declare, struct layouts agree between C and the
Python binding, the formats round-trip, and the C++ header compiles.  No
compute call is made on a GPU here."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import uncertaintyte_amd as u
from uncertaintyte_amd import jacob

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")


def _declared_c_functions():
    txt = open(os.path.join(INC, "uncertainty_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(unc_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = u.lib()
    names = _declared_c_functions()
    assert "unc_compute" in names and "unc_compute_device" in names
    for n in names:
        assert hasattr(L, n), n
    out = subprocess.run(["nm", "-DC", "--defined-only", u.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    # the reference's C++ entry point (include/uncertainty.h) with its exact signature
    assert "getCovariances(cov::Options&, cov::Statistic&, ceres::CRSMatrix&, double*, double*)" in out
    assert "getCovariancesLastError()" in out


def test_pointer_taking_entry_points_have_ctypes_signatures():
    """Every declared function with pointer parameters must have argtypes in the
    binding (without them ctypes passes 64-bit addresses as 32-bit ints)."""
    txt = open(os.path.join(INC, "uncertainty_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    L = u.lib()
    for name, params in re.findall(r"\b(unc_[a-z0-9_]+)\s*\(([^)]*)\)", txt):
        if "*" in params:
            assert getattr(L, name).argtypes is not None, name


def test_struct_layouts_match_c(tmp_path):
    src = tmp_path / "lay.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "uncertainty_amd.h"\n'
                   'int main(){printf("%zu %zu %zu %zu %zu %zu\\n", sizeof(unc_options),'
                   ' sizeof(unc_stats), sizeof(unc_debug), offsetof(unc_options, pts2fix),'
                   ' offsetof(unc_stats, cycle_change), offsetof(unc_stats, stage_ms));return 0;}\n')
    exe = tmp_path / "lay"
    subprocess.run(["gcc", "-I", INC, str(src), "-o", str(exe)], check=True)
    vals = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                           check=True).stdout.split()]
    assert vals == [ctypes.sizeof(u.Options), ctypes.sizeof(u.Stats), ctypes.sizeof(u.Debug),
                    u.Options.pts2fix.offset, u.Stats.cycle_change.offset, u.Stats.stage_ms.offset]


def test_default_options_and_strings():
    o = u.Options()
    u.lib().unc_default_options(ctypes.byref(o))
    assert o.algorithm == 2 and o.lambda_ == -1 and o.precision == 0 and o.refine == -1
    assert list(o.pts2fix) == [-1, -1, -1]
    assert u.lib().unc_abi_version() == 3
    for c in range(8):
        assert u.error_string(c)
    assert u.lib().unc_stage_name(5).decode() == "schur"


def test_workspace_bytes_scales_with_n2():
    j = jacob.read_jacob(os.path.join(ROOT, "tests", "golden", "04_Daliborka.jacob"))
    o = u.make_options(j)
    b = u.workspace_bytes(o, j.num_rows)
    N = j.num_cams * j.cam_params
    assert 5 * 8 * N * N <= b < 5 * 8 * N * N + 64 * 1024 * 1024


def test_workspace_chain_slots_only_on_the_fp64_path():
    """From N = 1024 the FP64 TE path forms the Taylor chain's powers in batches
    and holds five more N x N buffers for them; the Z+ paths never run the chain
    and must not reserve them (their footprint drives serving admission)."""
    from uncertaintyte_amd import scene
    j = scene.scene_jacobian(scene.ring_scene(120, 600, seed=3))
    N = j.num_cams * j.cam_params
    assert N >= 1024
    te = u.workspace_bytes(u.make_options(j, u.PREC_FP64), j.num_rows)
    zp = u.workspace_bytes(u.make_options(j, u.PREC_MIXED), j.num_rows)
    assert 5 * 8 * N * N <= te - zp < 5 * 8 * N * N + (1 << 20)


def test_no_gpu_is_a_clean_error():
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    j = jacob.read_jacob(os.path.join(ROOT, "tests", "golden", "01_Cube.jacob"))
    with pytest.raises(u.UncError) as e:
        u.compute_covariances(j)
    assert e.value.code == u.UNC_ERR_DEVICE


def test_server_without_gpu_is_a_clean_error():
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    with pytest.raises(u.UncError) as e:
        u.Server(slots=2)
    assert e.value.code == u.UNC_ERR_DEVICE


def test_jacob_roundtrip(tmp_path, cube):
    p = tmp_path / "c.jacob"
    jacob.write_jacob(str(p), cube)
    j2 = jacob.read_jacob(str(p))
    assert (j2.num_cams, j2.cam_params, j2.num_points, j2.num_obs, j2.fix) == \
        (cube.num_cams, cube.cam_params, cube.num_points, cube.num_obs, cube.fix)
    assert np.array_equal(j2.rows, cube.rows) and np.array_equal(j2.cols, cube.cols)
    assert np.array_equal(j2.vals, cube.vals)
    q = tmp_path / "c.jacob.npz"
    jacob.write_jacob_npz(str(q), cube)
    j3 = jacob.read_jacob(str(q))
    assert np.array_equal(j3.vals, cube.vals) and j3.fix == cube.fix


def test_cov_writer_is_byte_compatible(tmp_path, golden_dir):
    src = os.path.join(golden_dir, "04_Daliborka_output.cov")
    hdr, cam, pt = jacob.read_cov(src)
    h = jacob.CovHeader(64, 9, 200, 5205, fixed=(13, 40, 113), lam=float(hdr["Used lambda"]),
                        time_create_j=0.01, time_norm_j=0.19, time_multiply_jj=2.82,
                        time_split_jj=0.41, time_inv_v=0, time_compose_z=10.81, time_inv_z=9.7,
                        time_te=147.54, time_pts_unc=78.48, time_all=268.63,
                        cycle_change=[float(x) for x in hdr["TE cycle change"].split()])
    out = tmp_path / "o.cov"
    jacob.write_cov(str(out), h, cam, pt)
    assert open(out).read() == open(src).read()


def test_cpp_header_compiles_and_links(tmp_path):
    src = tmp_path / "use.cpp"
    src.write_text('#include "uncertainty.h"\n#include "uncertainty_amd.h"\n'
                   'int main(){cov::Options o(1,9,4,2); cov::Statistic s = cov::Statistic();'
                   ' ceres::CRSMatrix J; (void)o; (void)s; (void)J; return UNC_OK;}\n')
    exe = tmp_path / "use"
    subprocess.run(["g++", "-std=c++11", "-I", INC, str(src), "-o", str(exe), "-L",
                    os.path.dirname(u.LIB_PATH), "-luncertainty_amd",
                    "-Wl,-rpath," + os.path.dirname(u.LIB_PATH)], check=True)
    assert subprocess.run([str(exe)]).returncode == 0


def test_cli_rejects_unknown_flag():
    r = subprocess.run([u.CLI_PATH, "-bogus=1"], capture_output=True, text=True)
    assert r.returncode == 2 and "unknown flag" in r.stderr
