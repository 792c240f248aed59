"""The torch autograd reference of the S(q) forward path, pinned against the oracle.

``sf_torch_ref.structure_factor`` must reproduce the oracle's ``compute_sf``
(itself pinned by the reference's golden outputs) before its autograd
gradients may serve as the reference for the HIP adjoint.  Cases cover all 8
cell corners (the reference's corner z slot), a monoclinic 120 degree cell and
per-frame boxes (rescale).
"""
import math

import numpy as np
import pytest
import torch

import sf_torch_ref as T
from oracle import dens_oracle as O


def corner_system(seed, box=(7.0, 8.0, 9.0), frames=2, n_each=2, names=("C", "O", "H", "N")):
    rng = np.random.default_rng(seed)
    box = np.asarray(box)
    pts = []
    for cx in (0.05, 0.9):
        for cy in (0.1, 0.95):
            for cz in (0.02, 0.97):
                for _ in range(n_each):
                    pts.append(np.array([cx, cy, cz]) * box + rng.normal(0, 0.3, 3))
    pts += list(rng.uniform(0, 1, (5, 3)) * box)
    base = np.mod(np.array(pts), box)
    coords = np.stack([np.mod(base + rng.normal(0, 0.1, base.shape), box) for _ in range(frames)])
    dims = np.stack([box * (1.0 + 0.01 * t) for t in range(frames)])
    typ = [names[i % len(names)] for i in range(base.shape[0])]
    return coords, dims, typ


def spill_system(seed, frames=1, box=(40.0, 44.0, 42.0), names=("H", "C", "O", "Na"), per_corner=3):
    """Atoms within about an Angstrom of all 8 cell corners in a ~40 A box, with
    Na setting Nborder = 17 at Sres = 1: at N ~ 40, the reference's corner z
    slot (u_z + Nb for y high / z low, u_z - Nb for y low / z high,
    dens.py:107) sends contributions to z tiles of 16 cells that the atom's
    own window does not reach."""
    rng = np.random.default_rng(seed)
    box = np.asarray(box)
    pts = []
    for cx in (0, 1):
        for cy in (0, 1):
            for cz in (0, 1):
                for _ in range(per_corner):
                    pts.append(np.array([cx, cy, cz]) * box + rng.normal(0, 0.8, 3))
    pts += list(rng.uniform(0, 1, (12, 3)) * box)
    base = np.mod(np.array(pts), box)
    coords = np.stack([np.mod(base + rng.normal(0, 0.05, base.shape), box) for _ in range(frames)])
    dims = np.stack([box * (1.0 + 0.002 * t) for t in range(frames)])
    typ = [names[i % len(names)] for i in range(base.shape[0])]
    return coords, dims, typ


def spill_images(rframe, typ, rad, sres, dims, tile=16):
    """Quirk images (corner, y and z on opposite sides) whose destination z tile
    the atom's z window [ir - A, ir + A) does not reach, on the oracle's set-up."""
    r = np.array(rframe)[None]
    o = O.setup(r, np.array(dims), typ, rad, sres, False)
    n, nb, count = [int(v) for v in o["N"]], o["Nb"], 0
    for a in range(r.shape[1]):
        ir = (r[0, a] / o["dr"]).astype(int)
        A = o["bdict"][typ[a]].astype(int)
        win = [range(ir[d] - A[d], ir[d] + A[d]) for d in range(3)]
        xs = [u for u in win[0] if u < 0 or u >= n[0]]
        for ylo in (False, True):
            ys = [u for u in win[1] if (u < 0 if ylo else u >= n[1])]
            zs = [u for u in win[2] if (u >= n[2] if ylo else u < 0)]
            for uz in zs:
                zd = uz - nb if ylo else uz + nb
                t0 = (zd // tile) * tile
                t1 = min(t0 + tile, n[2])
                reached = any(max(ir[2] - A[2], t0 + j * n[2]) < min(ir[2] + A[2], t1 + j * n[2])
                              for j in (-1, 0, 1))
                if not reached:
                    count += len(xs) * len(ys)
    return count


@pytest.mark.parametrize("theta", [None, 120.0])
def test_torch_reference_matches_oracle_spill_system(rad, theta):
    coords, dims, typ = spill_system(11)
    assert spill_images(coords[0], typ, rad, 1.0, dims) > 100
    ucell = monoclinic(theta)
    want = O.compute_sf(np.array(coords), np.array(dims), typ, None, rad, ucell, 1.0,
                        better=False)["sf"]
    got = T.structure_factor(torch.tensor(coords), dims, typ, rad, 1.0, ucell).numpy()
    mask = want > 1e-12 * want.max()
    err = np.abs(got[mask] - want[mask]) / want[mask]
    assert err.max() < 1e-9, err.max()


def test_spill_count_is_zero_without_corner_atoms(rad):
    """The counter itself: atoms away from the corners give no spill images."""
    rng = np.random.default_rng(4)
    box = np.array([40.0, 44.0, 42.0])
    r = 8.0 + rng.uniform(0, 1, (30, 3)) * (box - 16.0)
    typ = [("H", "C", "O", "Na")[i % 4] for i in range(30)]
    assert spill_images(r, typ, rad, 1.0, box[None]) == 0


def monoclinic(theta_deg):
    if theta_deg is None:
        return np.eye(3)
    th = math.radians(theta_deg)
    return np.array([[1, 0, 0], [np.cos(th), np.sin(th), 0], [0, 0, 1]])


@pytest.mark.parametrize("theta", [None, 120.0])
def test_torch_reference_matches_oracle(rad, theta):
    coords, dims, typ = corner_system(5)
    ucell = monoclinic(theta)
    want = O.compute_sf(np.array(coords), np.array(dims), typ, None, rad, ucell, 0.5,
                        better=False)["sf"]
    got = T.structure_factor(torch.tensor(coords), dims, typ, rad, 0.5, ucell).numpy()
    mask = want > 1e-12 * want.max()
    err = np.abs(got[mask] - want[mask]) / want[mask]
    assert err.max() < 1e-9, err.max()


def test_torch_reference_matches_clipped_windows_golden(rad):
    """Windows that leave the padded grid (clipped as the reference's numpy
    slices clip them): the torch reference against the reference's own
    output for the clipped_windows fixture."""
    from conftest import load_golden
    c = load_golden("clipped_windows")
    S = T.structure_factor(torch.tensor(c["in_coords"]), c["in_dims"], c["typ"], rad, c["sres"],
                           c["in_ucell"], better=bool(c["in_better"]))
    want = c["out_sf"]
    np.testing.assert_allclose(S.numpy(), want, rtol=1e-9, atol=1e-12 * want.max())


def test_torch_reference_gradient_matches_fd(rad):
    """Autograd of the torch reference against central differences of the oracle."""
    coords, dims, typ = corner_system(6, frames=1)
    rng = np.random.default_rng(0)
    probe = None

    def oracle_loss(c):
        sf = O.compute_sf(np.array(c), np.array(dims), typ, None, rad, np.eye(3), 0.5,
                          better=False)["sf"]
        return float(np.sum(w * sf))

    sf0 = O.compute_sf(np.array(coords), np.array(dims), typ, None, rad, np.eye(3), 0.5,
                       better=False)["sf"]
    w = rng.normal(size=sf0.shape) / sf0.max()
    x = torch.tensor(coords, requires_grad=True)
    loss = torch.sum(torch.tensor(w) * T.structure_factor(x, dims, typ, rad, 0.5))
    loss.backward()
    g = x.grad.numpy()
    h = 1e-4
    for (a, d) in [(0, 0), (3, 1), (7, 2), (12, 0), (17, 2), (20, 1)]:
        cp, cm = np.array(coords), np.array(coords)
        cp[0, a, d] += h
        cm[0, a, d] -= h
        fd = (oracle_loss(cp) - oracle_loss(cm)) / (2 * h)
        assert abs(fd - g[0, a, d]) <= 1e-5 * np.abs(g).max() + 1e-6 * abs(fd), (a, d, fd, g[0, a, d])
