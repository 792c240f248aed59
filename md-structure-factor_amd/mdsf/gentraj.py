"""Seeded synthetic configurations with the features of the reference's
``gentraj.py``: a triclinic box filled uniformly with particles, optionally
keeping them out of (or, inverted, only inside) pores, layers and stacked
disks, written as ``.gro`` (last frame) + ``.trr`` (all frames) in nm.

Follows ``gentraj.py``:
* box vectors A, B, C from lengths (a, b, c) and angles (alpha, beta, gamma)
  (``:236-246``), with the reference's volume expression (``:238``), whose
  last term is 2 cos(alpha) cos(beta) sin(gamma) (the textbook term has
  cos(gamma); the two agree when alpha = beta = 90, the cells the features
  are meant for);
* pore axes on a rows x rows lattice (``:248-256``), layer centres
  ``linspace(0, c, nlayers)`` with nlayers = c / dbwl (``:229-232``), disk
  centres rotated about each pore axis per layer (``:258-274``);
* the membership tests ``check_pores`` (``:101-125``, with the optional
  Gaussian acceptance 2 Phi((d - r) / (r / 2))), ``check_layers``
  (``:128-153``) and ``check_disks`` (``:156-182``), each taking the first
  feature that contains the point;
* rejection sampling in the reference's order (``:278-318``): a point is
  redrawn until it passes the pore test, then until it passes the layer test
  (without re-testing pores), then the disk test.

Differences: the random stream is ``numpy.random.default_rng(seed)`` (the
reference draws from the unseeded global generator), sampling is vectorised
over the points of a frame, and the layer count passed to ``linspace`` is
``int(c / dbwl)`` (the reference passes the float, which numpy >= 1.18
rejects with a TypeError).
"""
import numpy as np
from scipy.special import ndtr

from . import trajio


def box_vectors(lengths, angles_deg):
    """A, B, C of a triclinic cell (gentraj.py:236-246); angles (yz, xz, xy) in degrees."""
    a, b, c = (float(x) for x in lengths)
    alpha, beta, gamma = (x * (np.pi / 180) for x in angles_deg)
    V = a * b * c * np.sqrt(1 - np.cos(alpha) ** 2 - np.cos(beta) ** 2 - np.cos(gamma) ** 2
                            + 2 * np.cos(alpha) * np.cos(beta) * np.sin(gamma))
    A = np.array([a, 0, 0])
    B = np.array([b * np.cos(gamma), b * np.sin(gamma), 0])
    C = np.array([c * np.cos(beta), c * ((np.cos(alpha) - np.cos(gamma) * np.cos(beta)) / np.sin(gamma)),
                  V / (a * b * np.sin(gamma))])
    return A, B, C


def pore_locations(a, gamma_deg, rows):
    """Pore axes (x, y) on a rows x rows lattice (gentraj.py:248-256)."""
    gamma = gamma_deg * np.pi / 180
    p2p = a / rows
    loc = np.zeros([rows * rows, 2])
    for i in range(rows):
        for j in range(rows):
            loc[i * rows + j, :] = [p2p * (0.5 + j) + p2p * (0.5 + i) * np.cos(gamma),
                                    p2p / 2 * np.sin(gamma) + i * p2p * np.sin(gamma)]
    return loc


def layer_locations(c, dbwl):
    """Layer centres (gentraj.py:229-232)."""
    nlayers = c / dbwl
    c = nlayers * dbwl
    return np.linspace(0, c, int(nlayers))


def disk_locations(pores, layers_z, c, dbwl, pore_radius, disks_per_layer, offset):
    """Disk centres around each pore axis, per layer (gentraj.py:258-274)."""
    nlayers = int(c / dbwl)
    out = np.zeros([pores.shape[0] * nlayers * disks_per_layer, 3])
    for i in range(pores.shape[0]):
        for j in range(nlayers):
            for k in range(disks_per_layer):
                theta = 2 * np.pi * k / disks_per_layer
                if offset:
                    theta += (j % 2) * (np.pi / disks_per_layer)
                pt = np.array([pore_radius * np.cos(theta), pore_radius * np.sin(theta), 0.0])
                out[i * nlayers * disks_per_layer + j * disks_per_layer + k, :] = \
                    pt + np.append(pores[i, :], layers_z[j])
    return out


def _first_hit(dist, limit):
    """Index of the first feature with dist <= limit per point, -1 if none."""
    hit = dist <= limit
    first = np.argmax(hit, axis=1)
    return np.where(hit.any(axis=1), first, -1), np.take_along_axis(dist, first[:, None], 1)[:, 0]


def in_pores(pts, pores, r, gaussian, rng):
    """check_pores (gentraj.py:101-125) for each point."""
    d = np.linalg.norm(pores[None, :, :] - pts[:, None, :2], axis=2)
    idx, dist = _first_hit(d, r)
    inside = idx >= 0
    if gaussian:
        p = 2 * ndtr((dist - r) / (r / 2))
        inside &= rng.random(len(pts)) >= p
    return inside


def in_layers(pts, layers_z, width, gaussian, rng):
    """check_layers (gentraj.py:128-153) for each point."""
    r = 0.5 * width
    d = np.abs(pts[:, 2:3] - layers_z[None, :])
    idx, dist = _first_hit(d, r)
    inside = idx >= 0
    if gaussian:
        p = 2 * ndtr((dist - r) / (r / 2))
        inside &= rng.random(len(pts)) >= p
    return inside


def in_disks(pts, disks, layer_width, disk_radius):
    """check_disks (gentraj.py:156-182) for each point."""
    rr = np.linalg.norm(disks[None, :, :2] - pts[:, None, :2], axis=2)
    zz = np.abs(disks[None, :, 2] - pts[:, None, 2])
    return np.any((rr <= disk_radius) & (zz <= 0.5 * layer_width), axis=1)


def generate(seed=0, npoints=10000, nframes=100, box=(8.0, 8.0, 8.0), angles=(90.0, 90.0, 60.0),
             pores=False, radius=0.5, rows=2, layers=False, dbwl=0.37, layer_width=0.1,
             disks=False, disks_per_layer=5, disk_radius=0.1, offset=False, gaussian=False,
             invert=False, max_rounds=100000):
    """Positions (nframes, npoints, 3) in nm and the cell vectors (3, 3) (rows A, B, C)."""
    rng = np.random.default_rng(seed)
    a, b, c = (float(x) for x in box)
    if layers or disks:
        lz = layer_locations(c, dbwl)
        c = (c / dbwl) * dbwl
    A, B, C = box_vectors((a, b, c), angles)
    pl = pore_locations(a, angles[2], rows) if (pores or disks) else None
    dl = disk_locations(pl, lz, c, dbwl, radius, int(disks_per_layer), offset) if disks else None

    def draw(n):
        u = rng.random((n, 3))
        return u[:, 0:1] * A + u[:, 1:2] * B + u[:, 2:3] * C

    tests = []
    if pores:
        tests.append(lambda p: in_pores(p, pl, radius, gaussian, rng))
    if layers:
        tests.append(lambda p: in_layers(p, lz, layer_width, gaussian, rng))
    if disks:
        tests.append(lambda p: in_disks(p, dl, layer_width, float(disk_radius)))
    points = np.zeros([nframes, npoints, 3])
    for t in range(nframes):
        pts = draw(npoints)
        for test in tests:
            # redraw until the point is outside (inside, with invert) this feature
            bad = test(pts) != invert
            rounds = 0
            while bad.any():
                rounds += 1
                if rounds > max_rounds:
                    raise RuntimeError("feature regions leave no room for the points")
                pts[bad] = draw(int(bad.sum()))
                bad[bad] = test(pts[bad]) != invert
        points[t] = pts
    return points, np.stack([A, B, C])


def write(prefix, points, cell, name="NA"):
    """``prefix.gro`` (last frame) and ``prefix.trr`` (all frames), as gentraj.py:320-333."""
    n = points.shape[1]
    trajio.write_gro(prefix + ".gro", points[-1], cell, [name] * n, [name] * n,
                     title="This is a .gro file")
    times = np.linspace(0, 1000, points.shape[0])
    trajio.write_trr(prefix + ".trr", points.astype(np.float32), cell, times=times)
