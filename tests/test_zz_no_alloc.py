"""unc_compute_device_noalloc: the solve with every device buffer in the
library's fixed per-device scratch (no allocation), against unc_compute_device
on the same device inputs.  Same kernels, so the outputs must be identical
bit for bit; two calls in a row, the second on a smaller scene (the scratch
keeps the first call's contents), both precision paths, and refusal of a
scene whose workspace exceeds the scratch."""
import ctypes
import os
import sys

import numpy as np
import pytest

import uncertaintyte_amd as u
from uncertaintyte_amd import jacob, scene

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _device_inputs(torch, j):
    dev = torch.device("cuda:0")
    per = j.cam_params * (j.cam_params + 1) // 2
    return dict(rows=torch.from_numpy(np.ascontiguousarray(j.rows, np.int32)).to(dev),
                cols=torch.from_numpy(np.ascontiguousarray(j.cols, np.int32)).to(dev),
                vals=torch.from_numpy(np.ascontiguousarray(j.vals, np.float64)).to(dev),
                cam=torch.zeros(j.num_cams * per, dtype=torch.float64, device=dev),
                pt=torch.zeros(j.num_points * 6, dtype=torch.float64, device=dev))


def _solve(torch, fn, ctx, j, precision):
    o = u.make_options(j, precision)
    b = _device_inputs(torch, j)
    b["cam"].fill_(-1.0)
    b["pt"].fill_(-1.0)
    torch.cuda.synchronize()
    st = u.Stats()
    rc = fn(ctx.handle, ctypes.byref(o), j.num_rows, j.num_cols, b["rows"].data_ptr(),
            b["cols"].data_ptr(), b["vals"].data_ptr(), b["cam"].data_ptr(), b["pt"].data_ptr(),
            ctypes.byref(st), None)
    torch.cuda.synchronize()
    return rc, b["cam"].cpu().numpy(), b["pt"].cpu().numpy(), st


@pytest.mark.parametrize("precision", [u.PREC_FP64, u.PREC_MIXED])
def test_noalloc_matches_device_entry(precision):
    import torch
    L = u.lib()
    ctx = u.Context(0)
    big = jacob.read_jacob(os.path.join(GOLDEN, "04_Daliborka.jacob"))
    small = jacob.read_jacob(os.path.join(GOLDEN, "01_Cube.jacob"))
    try:
        for name, j in (("daliborka", big), ("cube", small)):
            need = u.workspace_bytes(u.make_options(j, precision), j.num_rows)
            assert 0 < need <= u.NOALLOC_SCRATCH_BYTES
            say(f"noalloc precision={precision} {name}: workspace {need} bytes")
            rc0, cam0, pt0, st0 = _solve(torch, L.unc_compute_device, ctx, j, precision)
            rc1, cam1, pt1, st1 = _solve(torch, L.unc_compute_device_noalloc, ctx, j, precision)
            assert rc0 == u.UNC_OK and rc1 == u.UNC_OK, (rc0, rc1)
            assert np.array_equal(cam0, cam1), name
            assert np.array_equal(pt0, pt1), name
            assert st0.num_terms == st1.num_terms and st0.num_obs_kept == st1.num_obs_kept
            assert list(st0.cycle_change[:st0.num_terms]) == list(st1.cycle_change[:st1.num_terms])
    finally:
        ctx.close()


def test_noalloc_refuses_oversize_and_null_context():
    import torch
    L = u.lib()
    ctx = u.Context(0)
    try:
        j = scene.make_workload("rings200")
        o = u.make_options(j, u.PREC_FP64)
        assert u.workspace_bytes(o, j.num_rows) > u.NOALLOC_SCRATCH_BYTES
        say("noalloc: oversize scene")
        rc, _, _, _ = _solve(torch, L.unc_compute_device_noalloc, ctx, j, u.PREC_FP64)
        assert rc == u.UNC_ERR_TOO_LARGE
        small = jacob.read_jacob(os.path.join(GOLDEN, "01_Cube.jacob"))
        o = u.make_options(small, u.PREC_FP64)
        b = _device_inputs(torch, small)
        rc = L.unc_compute_device_noalloc(None, ctypes.byref(o), small.num_rows, small.num_cols,
                                          b["rows"].data_ptr(), b["cols"].data_ptr(),
                                          b["vals"].data_ptr(), b["cam"].data_ptr(),
                                          b["pt"].data_ptr(), None, None)
        assert rc == u.UNC_ERR_BAD_INPUT
        # after the refusals a solve still works
        rc0, cam0, pt0, _ = _solve(torch, L.unc_compute_device, ctx, small, u.PREC_FP64)
        rc1, cam1, pt1, _ = _solve(torch, L.unc_compute_device_noalloc, ctx, small, u.PREC_FP64)
        assert rc0 == rc1 == u.UNC_OK
        assert np.array_equal(cam0, cam1) and np.array_equal(pt0, pt1)
    finally:
        ctx.close()
