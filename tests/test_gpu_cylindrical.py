"""GPU cylindrical (r, z) average and point interpolation against the oracle
(PLOT_RAD_NEW, plot2d.py:559-655) and scipy's RegularGridInterpolator
(plot2d.py:588).  Tolerance: FP64, 1e-12 relative (the sums run in the same
order; cos / sin and the 3x3 product may differ in the last bit)."""
import math

import numpy as np
import pytest

from oracle import plot2d_oracle as P2

pytestmark = pytest.mark.gpu


def monoclinic(theta_deg):
    th = math.radians(theta_deg)
    return np.array([[1, 0, 0], [np.cos(th), np.sin(th), 0], [0, 0, 1]])


def _kgridplt(n, L, seed):
    import torch
    from mdsf import engine
    rng = np.random.default_rng(seed)
    sf = rng.random((n[0], n[1], n[2] // 2 + 1)) + 1.0
    plan = engine.Plan(n, 0, [1.0] * 3, np.eye(3), [], [], [], [])
    _, _, kp = plan.outputs(torch.from_numpy(sf).to(plan.device), L, False)
    return kp


@pytest.mark.parametrize("theta", [90.0, 120.0])
@pytest.mark.parametrize("fmt", [False, True])
def test_rz_average_matches_oracle(theta, fmt, gpu):
    from mdsf import engine
    print(f"rz_average theta={theta} format={fmt}", flush=True)
    kp = _kgridplt((24, 20, 26), [9.0, 8.0, 10.0], 21)
    ucell = monoclinic(theta)
    got = engine.rz_average(kp, ucell, format=fmt)
    want = P2.rz_average(kp.cpu().numpy(), ucell, format=fmt)
    for key in ("rarr", "zar", "rfin", "zfin"):
        np.testing.assert_array_equal(got[key], want[key])
    for key in ("oa", "final"):
        g = got[key].cpu().numpy()
        assert g.shape == want[key].shape
        np.testing.assert_allclose(g, want[key], rtol=1e-12, atol=0)


def test_rz_average_nan_bins(gpu):
    """Monoclinic cell at 60 degrees: the outer r bins leave the grid (NaN
    before the fill); the fill value and the finite bins match the oracle."""
    from mdsf import engine
    kp = _kgridplt((16, 18, 14), [7.0, 9.0, 8.0], 22)
    ucell = monoclinic(60.0)
    got = engine.rz_average(kp, ucell, format=True)
    want = P2.rz_average(kp.cpu().numpy(), ucell, format=True)
    np.testing.assert_allclose(got["final"].cpu().numpy(), want["final"], rtol=1e-12, atol=0)


def test_interp_points_match_scipy(gpu):
    import torch
    from scipy.interpolate import RegularGridInterpolator
    from mdsf import engine
    kp = _kgridplt((12, 10, 14), [6.0, 5.0, 7.0], 23)
    D = kp.cpu().numpy()
    X, Y, Z = D[:, 0, 0, 0], D[0, :, 0, 1], D[0, 0, :, 2]
    ES = RegularGridInterpolator((X, Y, Z), D[..., 3], bounds_error=False)
    rng = np.random.default_rng(5)
    lo = np.array([X[0], Y[0], Z[0]]) - 0.3
    hi = np.array([X[-1], Y[-1], Z[-1]]) + 0.3
    pts = rng.uniform(lo, hi, size=(5000, 3))
    pts[:20] = np.stack([X[rng.integers(len(X), size=20)], Y[rng.integers(len(Y), size=20)],
                         Z[rng.integers(len(Z), size=20)]], -1)     # exact grid points
    want = ES(pts)
    got = engine.interp_points(kp, torch.from_numpy(pts).to(kp.device)).cpu().numpy()
    assert np.array_equal(np.isnan(got), np.isnan(want))
    ok = ~np.isnan(want)
    np.testing.assert_allclose(got[ok], want[ok], rtol=1e-12, atol=0)
