"""Seeded feature generator (mdsf.gentraj, after gentraj.py:101-333), on the CPU."""
import numpy as np
import pytest

from mdsf import gentraj as G
from mdsf import trajio


def frac(points, cell):
    return points @ np.linalg.inv(cell)


def test_box_vectors_closed_forms():
    A, B, C = G.box_vectors((8.0, 7.0, 6.0), (90.0, 90.0, 60.0))
    np.testing.assert_allclose(A, [8.0, 0, 0])
    np.testing.assert_allclose(B, [3.5, 7.0 * np.sqrt(3) / 2, 0], atol=1e-12)
    np.testing.assert_allclose(C, [0, 0, 6.0], atol=1e-12)
    ang = lambda u, v: np.degrees(np.arccos(u @ v / np.linalg.norm(u) / np.linalg.norm(v)))
    # alpha = beta = 90 (the cells the features are meant for): |C| = c and
    # the angles as given, for any gamma
    A, B, C = G.box_vectors((5.0, 6.0, 7.0), (90.0, 90.0, 75.0))
    assert np.isclose(np.linalg.norm(B), 6.0) and np.isclose(np.linalg.norm(C), 7.0)
    assert np.isclose(ang(A, B), 75.0) and np.isclose(ang(A, C), 90.0)
    # general cells keep the reference's volume expression (gentraj.py:238),
    # whose last term is 2 cos(alpha) cos(beta) sin(gamma), not ... cos(gamma):
    # C's direction is right, its z component follows that volume
    al, be, ga = np.radians([80.0, 70.0, 60.0])
    A, B, C = G.box_vectors((5.0, 6.0, 7.0), (80.0, 70.0, 60.0))
    V = 5 * 6 * 7 * np.sqrt(1 - np.cos(al) ** 2 - np.cos(be) ** 2 - np.cos(ga) ** 2
                            + 2 * np.cos(al) * np.cos(be) * np.sin(ga))
    assert np.isclose(C[2], V / (5 * 6 * np.sin(ga)))
    assert np.isclose(C[0], 7 * np.cos(be)) and np.isclose(ang(A, B), 60.0)


def test_points_fill_the_cell_and_are_seeded():
    p, cell = G.generate(seed=3, npoints=4000, nframes=2)
    f = frac(p.reshape(-1, 3), cell)
    assert f.min() >= 0.0 and f.max() < 1.0
    # uniform in fractional coordinates: each half holds about half the points
    assert abs(np.mean(f[:, 0] < 0.5) - 0.5) < 0.03
    q, _ = G.generate(seed=3, npoints=4000, nframes=2)
    assert np.array_equal(p, q)


def test_pores_excluded_and_inverted():
    r = 0.9
    pl = G.pore_locations(8.0, 60.0, 2)
    p, _ = G.generate(seed=4, npoints=3000, nframes=1, pores=True, radius=r)
    d = np.linalg.norm(p[0][:, None, :2] - pl[None], axis=-1).min(-1)
    assert d.min() > r
    p, _ = G.generate(seed=4, npoints=1000, nframes=1, pores=True, radius=r, invert=True)
    d = np.linalg.norm(p[0][:, None, :2] - pl[None], axis=-1).min(-1)
    assert d.max() <= r


def test_gaussian_pores_thin_out_towards_the_axis():
    r = 1.5
    pl = G.pore_locations(8.0, 60.0, 2)
    p, _ = G.generate(seed=5, npoints=20000, nframes=1, pores=True, radius=r, gaussian=True)
    d = np.linalg.norm(p[0][:, None, :2] - pl[None], axis=-1).min(-1)
    inner = np.mean(d < 0.5 * r)
    shell = np.mean((d > 0.75 * r) & (d <= r))
    # acceptance 2 Phi((d - r) / (r / 2)): 0.05 at the axis, 0.62 at 3r/4; the
    # shell holds 7/9 of the annulus area of the inner disk's 1/3, so without
    # the weight shell / inner would be ~1.75
    assert 0 < inner and shell / inner > 4.0


def test_layers_and_disks_excluded():
    p, _ = G.generate(seed=6, npoints=3000, nframes=1, layers=True, layer_width=0.12)
    lz = G.layer_locations(8.0, 0.37)
    assert np.abs(p[0][:, 2:3] - lz[None]).min() > 0.06
    p, _ = G.generate(seed=6, npoints=2000, nframes=1, disks=True, disk_radius=0.4,
                      layer_width=0.2, offset=True)
    pl = G.pore_locations(8.0, 60.0, 2)
    dl = G.disk_locations(pl, G.layer_locations(8.0, 0.37), 8.0, 0.37, 0.5, 5, True)
    assert not G.in_disks(p[0], dl, 0.2, 0.4).any()
    # disk centres sit at the pore radius from their pore axis
    d = np.linalg.norm(dl[:, None, :2] - pl[None], axis=-1).min(-1)
    np.testing.assert_allclose(d, 0.5, atol=1e-12)


def test_files_round_trip(tmp_path):
    p, cell = G.generate(seed=7, npoints=300, nframes=3, pores=True)
    G.write(str(tmp_path / "g"), p, cell, name="C")
    xyz, box, t, step = trajio.read_trr(str(tmp_path / "g.trr"))
    assert np.array_equal(xyz, p.astype(np.float32))
    np.testing.assert_allclose(box[0], cell, atol=1e-6)
    np.testing.assert_allclose(t, np.linspace(0, 1000, 3), rtol=1e-6)
    gx, gbox, names, _ = trajio.read_gro(str(tmp_path / "g.gro"))
    np.testing.assert_allclose(gx[0], p[-1], atol=6e-4)
    assert names == ["C"] * 300
