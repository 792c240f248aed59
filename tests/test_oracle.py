"""Pins the CPU oracle (oracle/dens_oracle.py) before the HIP path is tested against it.

* golden vectors written by the reference's own compute_sf (tests/golden);
* brute force / closed forms for the pieces the goldens only cover as a
  composition: the border fold, the single-atom deposition, get_dplot;
* known entries of radii.txt.
"""
import math

import numpy as np
import pytest

from conftest import RADII, golden_cases, load_golden
from oracle import dens_oracle as O


@pytest.mark.parametrize("name", golden_cases())
def test_oracle_matches_reference_goldens(name, rad):
    c = load_golden(name)
    r = np.array(c["in_coords"])
    dims = np.array(c["in_dims"])
    if int(c["in_seed"]) >= 0:
        np.random.seed(int(c["in_seed"]))
    res = O.compute_sf(r, dims, c["typ"], None, rad, c["in_ucell"], c["sres"],
                       better=bool(c["in_better"]), random_noise=int(c["in_noise"]))
    # in-place mutation of r (rescale + wrap) is bit-exact, dtype kept
    assert r.dtype == c["out_r"].dtype
    assert np.array_equal(r, c["out_r"])
    assert np.array_equal(res["N"], c["out_N"])
    assert res["L"].dtype == c["out_L"].dtype and np.array_equal(res["L"], c["out_L"])
    for k in ("sf", "sfplt", "kgrid", "kgridplt"):
        assert res[k].shape == c["out_" + k].shape
        np.testing.assert_allclose(res[k], c["out_" + k], rtol=1e-12, atol=0)


def test_load_radii_known_entries():
    rad = O.load_radii(RADII)
    assert rad["H"] == (1.0, 0.53)
    assert rad["C"] == (6.0, 0.70)
    assert rad["N"] == (7.0, 0.56)
    assert rad["O"] == (8.0, 0.60)
    assert rad["Na"] == (11.0, 2.27) and rad["NA"] == (11.0, 2.27)
    assert rad["R3"] == (1.0, 0.90)
    assert len(rad) == 226


def _fold_bruteforce(d0, n, nb):
    out = np.zeros(tuple(n))
    for p in np.ndindex(d0.shape):
        out[O.fold_destination(p, n, nb)] += d0[p]
    return out


@pytest.mark.parametrize("n,nb", [((8, 10, 12), 2), ((8, 8, 8), 3), ((6, 8, 10), 4),
                                  ((8, 6, 6), 5), ((10, 10, 10), 0), ((6, 6, 6), 6)])
def test_fold_closed_form_matches_slicing(n, nb):
    rng = np.random.default_rng(1)
    d0 = rng.random(tuple(x + 2 * nb for x in n))
    des, ori = O.fold_tables(np.array(n), nb)
    d1 = O.fold(d0, des, ori)
    np.testing.assert_allclose(d1, _fold_bruteforce(d0, n, nb), rtol=1e-13, atol=1e-13)
    assert abs(d1.sum() - d0.sum()) < 1e-9 * d0.sum()


def test_fold_corner_bug_is_y_side_dependent():
    # a unit mass at padded corner (L, L, H) lands shifted in z by Nz - Nb
    n, nb = np.array([8, 8, 8]), 2
    d0 = np.zeros(tuple(n + 2 * nb))
    p = (0, 1, nb + 8 + 1)          # x on L, y on L, z on H
    d0[p] = 1.0
    d1 = O.fold(d0, *O.fold_tables(n, nb))
    std = ((p[0] - nb) % 8, (p[1] - nb) % 8, (p[2] - nb) % 8)
    assert d1[std] == 0.0
    assert d1[std[0], std[1], std[2] + 8 - nb] == 1.0


def test_single_atom_deposit_closed_form(rad):
    # orthorhombic grid: value = Nel/s^3 * prod_d exp(-(r_d - x_d)^2 / 2 s^2) on the window
    dr = np.array([0.5, 0.4, 0.45])
    nb = 12
    shape = (30 + 2 * nb, 30 + 2 * nb, 30 + 2 * nb)
    r = np.array([[3.3, 4.1, 5.05]])
    bd = O.borders(rad, dr, {"C"})
    d0, ir = O.deposit_frame(r, ["C"], rad, bd, dr, nb, np.eye(3), shape)
    nel, sig = rad["C"]
    A = bd["C"].astype(int)
    assert list(ir[0]) == [int(3.3 / 0.5), int(4.1 / 0.4), int(5.05 / 0.45)]
    ax = [np.arange(ir[0][d] - A[d], ir[0][d] + A[d]) for d in range(3)]
    g = [np.exp(-(r[0, d] - ax[d] * dr[d]) ** 2 / (2 * sig ** 2)) for d in range(3)]
    expect = nel / sig ** 3 * g[0][:, None, None] * g[1][None, :, None] * g[2][None, None, :]
    win = d0[ax[0][0] + nb:ax[0][-1] + nb + 1, ax[1][0] + nb:ax[1][-1] + nb + 1,
             ax[2][0] + nb:ax[2][-1] + nb + 1]
    np.testing.assert_allclose(win, expect, rtol=1e-13)
    assert np.count_nonzero(d0) == win.size
    # truncation at 1e-24 relative: the window edge value is close to PRECISION * amp
    mass = d0.sum() * np.prod(dr)
    assert abs(mass - nel * (2 * math.pi) ** 1.5) < 1e-6 * mass


def test_deposit_translation_by_whole_cells(rad):
    dr = np.array([0.5, 0.5, 0.5])
    nb = 12
    shape = (40 + 2 * nb,) * 3
    bd = O.borders(rad, dr, {"O"})
    a, _ = O.deposit_frame(np.array([[5.2, 6.3, 7.4]]), ["O"], rad, bd, dr, nb, np.eye(3), shape)
    b, _ = O.deposit_frame(np.array([[5.2 + 1.0, 6.3, 7.4 + 1.5]]), ["O"], rad, bd, dr, nb,
                           np.eye(3), shape)
    np.testing.assert_allclose(np.roll(np.roll(a, 2, axis=0), 3, axis=2), b, rtol=1e-9,
                               atol=1e-12 * a.max())


@pytest.mark.parametrize("n", [(8, 8, 8), (6, 10, 12), (16, 8, 4), (12, 14, 10)])
def test_dplot_closed_form(n):
    rng = np.random.default_rng(2)
    sf = rng.random((n[0], n[1], n[2] // 2 + 1))
    out = O.get_dplot(sf)
    assert out.shape == (n[0] - 2, n[1] - 2, n[2] - 1)
    for a, b, c in np.ndindex(out.shape):
        src = O.dplot_index(a, b, c, n)
        if src is None:
            want = 1.0 / 3.0 * (sf[1, 0, 0] + sf[0, 1, 0] + sf[0, 0, 1])
        else:
            want = sf[src]
        assert out[a, b, c] == want


def test_frame_power_matches_bruteforce_dft():
    rng = np.random.default_rng(3)
    d = rng.random((6, 4, 8))
    p = O.frame_power(d)
    idx = np.indices(d.shape).reshape(3, -1).T
    for k in [(0, 0, 0), (1, 2, 3), (5, 3, 4), (2, 1, 1)]:
        ph = np.exp(-2j * np.pi * (idx @ (np.array(k) / np.array(d.shape))))
        f = np.sum(d.reshape(-1) * ph)
        assert abs(p[k] - abs(f) ** 2) < 1e-9 * p[0, 0, 0]
    assert abs(p[0, 0, 0] - d.sum() ** 2) < 1e-12 * p[0, 0, 0]


def test_wrap_range_buffered_branch():
    assert O.wrap_atom_range(10, 5) == (0, 5)
    assert O.wrap_atom_range(100, 1000000) == (0, 100)
    assert O.wrap_atom_range(3000000, 1000000) == (0, 0)
    assert O.wrap_atom_range(2500000, 1200000) == (2000000, 2500000)
