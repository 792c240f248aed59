"""mdsf_rfft3_noalloc (the plan-free transform whose working set lives in a
fixed device array) against mdsf_rfft3 on a plan, through the C ABI.

Both run the same kernels with the same twiddles, so the spectra must agree
bit for bit.  Two calls in a row share the one device array: the second, on a
smaller grid, must not see anything the first left behind.  The transform is
also checked against numpy's FP64 rfftn (library oracle) at the tolerance of
its precision.
"""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _noalloc(lib, density, precision):
    from mdsf import _lib
    n = (ctypes.c_int32 * 3)(*density.shape)
    out = torch.empty(tuple(density.shape[:2]) + (density.shape[2] // 2 + 1, 2),
                      dtype=torch.float64, device=density.device)
    stream = ctypes.c_void_p(torch.cuda.current_stream(density.device).cuda_stream)
    _lib.check(lib.mdsf_rfft3_noalloc(n, precision, ctypes.c_void_p(density.data_ptr()),
                                      ctypes.c_void_p(out.data_ptr()), stream))
    return torch.view_as_complex(out)


def _plan_rfft3(density, fft64):
    from mdsf import engine
    n = list(density.shape)
    plan = engine.Plan(n, 0, [1.0, 1.0, 1.0], np.eye(3), [], [], [], [], density.device,
                       fft64=fft64)
    out = plan.rfft3(density)
    torch.cuda.synchronize()
    plan.close()
    return out


@pytest.mark.parametrize("fft64", [False, True])
def test_noalloc_matches_plan_rfft3_twice(fft64, gpu):
    from mdsf import _lib
    lib = _lib.load()
    dtype = torch.float64 if fft64 else torch.float32
    precision = _lib.F64 if fft64 else _lib.F32
    rng = np.random.default_rng(7)
    # the second call is on a smaller grid (radix 3 and 5 factors on the first)
    for shape in ((96, 80, 120), (32, 48, 16)):
        host = rng.random(shape)
        dens = torch.from_numpy(host).to(device=gpu, dtype=dtype)
        got = _noalloc(lib, dens, precision)
        want = _plan_rfft3(dens, fft64)
        assert got.shape == want.shape == (shape[0], shape[1], shape[2] // 2 + 1)
        g, w = got.cpu().numpy(), want.cpu().numpy()
        assert np.array_equal(g.view(np.float64), w.view(np.float64)), (
            shape, np.max(np.abs(g - w)))
        ref = np.fft.rfftn(dens.double().cpu().numpy())
        tol = 1e-12 if fft64 else 2e-6
        assert np.max(np.abs(g - ref)) <= tol * np.max(np.abs(ref)), shape


def test_noalloc_rejects_oversized_grids(gpu):
    from mdsf import _lib
    lib = _lib.load()
    n = (ctypes.c_int32 * 3)(512, 512, 512)
    dummy = torch.empty(1, device=gpu)
    out = torch.empty(1, dtype=torch.float64, device=gpu)
    rc = lib.mdsf_rfft3_noalloc(n, _lib.F32, ctypes.c_void_p(dummy.data_ptr()),
                                ctypes.c_void_p(out.data_ptr()), None)
    assert rc == _lib.ERR_ARG
    assert b"too large" in lib.mdsf_last_error()
