"""Pins of the cylindrical-average oracle (``oracle/plot2d_oracle.py``) by
what the mathematics fixes, on the CPU (``PLOT_RAD_NEW``, plot2d.py:559-655;
``normalize_alkanes``, plot2d.py:874-918)."""
import math

import numpy as np
import pytest

from oracle import plot2d_oracle as P2


def monoclinic(theta_deg):
    th = math.radians(theta_deg)
    return np.array([[1, 0, 0], [np.cos(th), np.sin(th), 0], [0, 0, 1]])


def grid(field, n=(9, 8, 7), lo=(-3.0, -2.5, -2.0), hi=(3.0, 2.5, 2.0)):
    """A kgridplt-shaped array (nx, ny, nz, 4) with ascending axes and channel 3 = field(X, Y, Z)."""
    X, Y, Z = (np.linspace(lo[d], hi[d], n[d]) for d in range(3))
    D = np.zeros(n + (4,))
    D[..., 0] = X[:, None, None]
    D[..., 1] = Y[None, :, None]
    D[..., 2] = Z[None, None, :]
    D[..., 3] = field(D[..., 0], D[..., 1], D[..., 2])
    return D


@pytest.mark.parametrize("theta", [90.0, 120.0, 75.0])
def test_reciprocal_inverse_is_the_transposed_cell(theta):
    # b_i . a_j = delta_ij makes (b1; b2; b3) = (A^T)^-1, so b_inv = A^T
    u = monoclinic(theta)
    np.testing.assert_allclose(P2.reciprocal_inverse(u), u.T, atol=1e-14)


@pytest.mark.parametrize("theta", [90.0, 120.0])
def test_constant_field(theta):
    D = grid(lambda x, y, z: np.full_like(x, 2.5))
    out = P2.rz_average(D, monoclinic(theta), rbins=60)
    assert np.allclose(out["final"], 1.0, rtol=0, atol=1e-14)
    raw = P2.rz_average(D, monoclinic(theta), format=False, rbins=60)
    assert raw["final"].shape == (2 * 60 - 1, 7)
    assert np.array_equal(raw["rfin"], np.append(-raw["rarr"][::-1], raw["rarr"][1:]))


def test_linear_field_closed_form():
    """S = a + b X + c Y + d Z is reproduced exactly by the trilinear
    interpolant; the theta average of equally spaced samples removes the
    x / y terms, so every bin whose samples stay inside the grid holds
    a + d z (orthogonal cell, Rmax inside the grid: no NaN bins)."""
    a, b, c, d = 1.5, 0.7, -0.4, 0.3
    D = grid(lambda x, y, z: a + b * x + c * y + d * z)
    rb = 50
    X, Y, Z = D[:, 0, 0, 0], D[0, :, 0, 1], D[0, 0, :, 2]
    out = P2.rz_average(D, np.eye(3), format=True, rbins=rb)   # format: no mean normalisation
    oa = out["oa"]
    rmax = 0.95 * min(X[-1] - X[0], Y[-1] - Y[0]) / 2.0
    np.testing.assert_allclose(out["rarr"], np.linspace(0.0, rmax, rb), rtol=0, atol=0)
    np.testing.assert_allclose(out["zar"], Z, rtol=0, atol=1e-15)
    want = a + d * Z[None, :] + 0.0 * out["rarr"][:, None]
    np.testing.assert_allclose(oa, want, rtol=0, atol=1e-12)
    # and the mean normalisation of the unformatted output
    un = P2.rz_average(D, np.eye(3), format=False, rbins=rb)
    np.testing.assert_allclose(un["oa"], want / np.average(want), rtol=0, atol=1e-12)


@pytest.mark.parametrize("theta", [120.0, 60.0])
def test_bilinear_field_monoclinic_closed_form(theta):
    """S = X Y + Z on a monoclinic cell.  The trilinear interpolant reproduces
    X Y exactly; a sample at (r cos t, r sin t, z) is taken at p = (r cos t,
    r sin t, z) @ b_inv (plot2d.py:619-627), and with b_inv = A^T,
    X = r cos t, Y = r (cos t cos(theta) + sin t sin(theta)), so the average
    over n >= 3 equally spaced t is r^2 cos(theta) / 2 + z.  The transposed
    product b_inv @ p gives r^2 cos(theta) sin(theta) / 2 instead, and a wrong
    r or z bin grid shifts the values."""
    u = monoclinic(theta)
    D = grid(lambda x, y, z: x * y + z)
    rb = 50
    out = P2.rz_average(D, u, format=True, rbins=rb)   # format: no mean normalisation
    r = out["rarr"][:, None]
    zar = out["zar"][None, :]
    X, Y = D[:, 0, 0, 0], D[0, :, 0, 1]
    rmax = 0.95 * min((X[-1] - X[0]) * u[0][0], (Y[-1] - Y[0]) * u[1][1]) / 2.0
    np.testing.assert_allclose(out["rarr"], np.linspace(0.0, rmax, rb), rtol=0, atol=0)
    nth = np.maximum((P2.THETA_BINS_PER_INV_ANG * 2 * np.pi * out["rarr"]).astype(int), 1)
    ok = (nth >= 3) | (out["rarr"] == 0.0)
    want = r ** 2 * np.cos(np.radians(theta)) / 2.0 + zar
    np.testing.assert_allclose(out["oa"][ok], (want + 0 * zar)[ok], rtol=0, atol=1e-12)
    wrong = r ** 2 * np.cos(np.radians(theta)) * np.sin(np.radians(theta)) / 2.0 + zar
    assert np.max(np.abs(out["oa"][ok] - wrong[ok])) > 1e-2


def test_out_of_grid_bins_take_the_minimum():
    """A cell whose r range leaves the grid on the y axis: those bins are NaN
    before the fill and take the smallest finite bin (plot2d.py:629-630)."""
    D = grid(lambda x, y, z: 3.0 + z, lo=(-3.0, -1.0, -2.0), hi=(3.0, 5.0, 2.0))
    out = P2.rz_average(D, np.eye(3), format=True, rbins=40)
    oa = out["oa"]
    r = out["rarr"]
    inside = r <= 1.0   # the y axis reaches down to -1 only
    np.testing.assert_allclose(oa[inside], 3.0 + out["zar"][None, :] + 0 * r[inside, None],
                               atol=1e-12)
    np.testing.assert_allclose(oa[~inside], np.min(3.0 + out["zar"]), atol=1e-12)


def test_trilinear_value_by_hand():
    from scipy.interpolate import RegularGridInterpolator
    rng = np.random.default_rng(0)
    v = rng.random((2, 2, 2))
    ES = RegularGridInterpolator(([0.0, 2.0], [0.0, 1.0], [0.0, 4.0]), v, bounds_error=False)
    x, y, z = 0.5, 0.25, 3.0
    tx, ty, tz = 0.25, 0.25, 0.75
    want = sum(v[i, j, k] * (tx if i else 1 - tx) * (ty if j else 1 - ty) * (tz if k else 1 - tz)
               for i in (0, 1) for j in (0, 1) for k in (0, 1))
    assert abs(ES([x, y, z])[0] - want) < 1e-15
    assert np.isnan(ES([2.5, 0.5, 1.0])[0])


def test_normalize_alkanes_ring_mean():
    # points (r, z): (1.5, 0) and (1.5, 0.3) lie in the ring at 0 and 11.3 degrees
    # (inside the kept bins); (1.1, 1.1) lies in the ring at 45 degrees (outside
    # them); the rest are off the ring
    R = np.array([1.1, 1.5])
    Z = np.array([0.0, 0.3, 1.1])
    I = np.array([[7.0, 9.0, 100.0], [2.0, 4.0, 50.0]])
    assert P2.normalize_alkanes(R, Z, I, 1.4, 1.57, 120) == pytest.approx(3.0, abs=1e-15)
