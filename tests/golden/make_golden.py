"""Write golden input/output vectors for ``dens.compute_sf`` from the real reference.

The reference (joeyelk/MD-Structure-Factor, ``dens.py``) pins no result of its
hot path in its own tests (``test/test_load_traj.py:16-25`` only checks atom
names), so this script runs the reference's ``compute_sf`` itself on tiny
inputs and stores inputs and every ``.npz`` key it writes.  Only the numbers are
kept; no reference source or bytecode enters this repository.

``dens.py:4`` imports ``past.utils.old_div`` (PyPI ``future``), which is not
installed here.  For the float / ndarray operands it receives
(``dens.py:42,53``) ``old_div`` is plain true division, so a stub module
providing ``old_div = a / b`` is inserted into ``sys.modules`` first.

Run from the repository root (needs ``/root/reference``, this container only)::

    python tests/golden/make_golden.py [case ...]

With case names, only those fixtures are (re)written.
"""
import math
import os
import sys
import tempfile
import types

import numpy as np

REF = "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))


def _import_reference():
    past = types.ModuleType("past")
    utils = types.ModuleType("past.utils")
    utils.old_div = lambda a, b: a / b
    past.utils = utils
    sys.modules.setdefault("past", past)
    sys.modules.setdefault("past.utils", utils)
    sys.path.insert(0, REF)
    import dens  # noqa: E402  (the reference module)
    return dens


def _water_like(rng, n, L, names):
    r = rng.uniform(0.0, 1.0, size=(n, 3)) * np.asarray(L)
    typ = np.array([names[i % len(names)] for i in range(n)])
    return r, typ


def _run(dens, name, coords, dims, typ, ucell, sres, *, better=True, noise=0,
         seed=None, note=""):
    """Run the reference compute_sf on copies of the inputs and save everything."""
    dens.USE_BETTER_RESOLUTION = better
    dens.RANDOM_NOISE = noise
    dens.PRINT_DETAILS = False
    rad = dens.load_radii(os.path.join(REF, "radii.txt"))
    r_in = np.array(coords, copy=True)
    dims_in = np.array(dims, copy=True)
    r_work = np.array(coords, copy=True)
    dims_work = np.array(dims, copy=True)
    if seed is not None:
        np.random.seed(seed)
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "sf")
        dens.compute_sf(r_work, dims_work, typ, out, rad, ucell, sres)
        res = dict(np.load(out + ".npz"))
    payload = dict(
        in_coords=r_in, in_dims=dims_in, in_typ=np.asarray(typ).astype("U8"),
        in_ucell=np.asarray(ucell, dtype=np.float64),
        in_sres=np.asarray(sres, dtype=np.float64),
        in_better=np.asarray(better), in_noise=np.asarray(noise),
        in_seed=np.asarray(-1 if seed is None else seed),
        out_r=r_work,
        note=np.asarray(note),
    )
    for k, v in res.items():
        payload["out_" + k] = v
    np.savez_compressed(os.path.join(OUT, name + ".npz"), **payload)
    print(f"{name}: N={res['N']} sf{res['sf'].shape} kgridplt{res['kgridplt'].shape}")


def _clipped_windows(dens, eye):
    """G: windows that leave the padded grid.  The reference slices the same
    padded index range out of `grid` and `d0` (dens.py:292-308), so numpy
    clips both alike instead of raising: a window crossing the low edge
    becomes empty, one crossing the high edge is cut there, one entirely at
    negative indices counts from the far end.  Coordinates outside (-L, 2L)
    keep such windows after the single wrap of dens.py:206-212."""
    rng = np.random.default_rng(20261203)
    L = np.array([12.0, 12.0, 12.0])                 # Sres 0.75: N = 16, Nb = 7 (C)
    r0, typ = _water_like(rng, 10, L, ["C", "O", "H"])
    out = np.array([
        [-14.25, 6.0, 6.0],      # -> -2.25: x window crosses the low edge (empty)
        [-24.0, 5.0, 7.0],       # -> -12: x window wholly at negative indices
        [27.2, 6.5, 5.5],        # -> 15.2: x window cut at the high edge
        [34.5, 6.0, 6.0],        # -> 22.5: x window starts past the high edge (empty)
        [5.0, 26.9, 4.0],        # -> 14.9: y window cut at the high edge
        [27.1, 27.3, 27.0],      # -> 15.1 .. 15.3: cut on all three axes (corner slot)
        [11.9, 27.4, -0.05],     # cut on y, z wraps to 11.95
    ])
    otyp = np.array(["C", "C", "C", "O", "H", "C", "O"])
    r = np.vstack([r0, out])
    t = np.concatenate([typ, otyp])
    frames = np.stack([r, r + rng.normal(0, 0.05, r.shape)])
    dims = np.tile(L, (2, 1))
    _run(dens, "clipped_windows", frames, dims, t, eye, 0.75,
         note="windows off the padded grid: clipped as numpy slices, N=16, Nb=7")


def main():
    dens = _import_reference()
    eye = np.eye(3)
    only = set(sys.argv[1:])
    if only:
        if "clipped_windows" in only:
            _clipped_windows(dens, eye)
        if only - {"clipped_windows"}:
            raise SystemExit("cases A-F are written together (one random stream); run "
                             "without arguments to rewrite them")
        return
    rng = np.random.default_rng(20261202)

    # A: orthorhombic float64, atoms in all 8 corners and one at exactly 0.0
    #    (wraps to L, home cell ir = N), Nborder > N/2.
    L = np.array([8.0, 8.0, 8.0])
    base = []
    for cx in (0.15, 7.9):
        for cy in (0.3, 7.75):
            for cz in (0.05, 7.6):
                base.append((cx, cy, cz))
    base.append((0.0, 0.0, 0.0))
    extra, _ = _water_like(rng, 7, L, ["O"])
    r0 = np.vstack([np.array(base), extra])
    typ = np.array((["H", "C", "O"] * 8)[: r0.shape[0]])
    frames = np.stack([r0, np.mod(r0 + rng.normal(0, 0.2, r0.shape), L)])
    frames[1, 8] = (0.0, 0.0, 0.0)
    dims = np.tile(L, (2, 1))
    _run(dens, "orth_corners_f64", frames, dims, typ, eye, 0.5,
         note="orthorhombic f64, 8 corner atoms + exact zero, N=16, Nb=10")

    # B: monoclinic theta=120 deg, float32 trajectory through the caller's
    #    coordinate transform (main_gromacs.py:204-207), varying dims -> rescale.
    th = 120.0 * math.pi / 180.0
    ucell = np.array([[1, 0, 0], [np.cos(th), np.sin(th), 0], [0, 0, 1]])
    L = np.array([9.0, 8.0, 10.0])
    nat = 40
    r0, typ = _water_like(rng, nat, L, ["O", "H", "H", "C", "N"])
    fr = []
    dd = []
    for t in range(3):
        s = 1.0 + 0.01 * (t - 1)
        fr.append((r0 + rng.normal(0, 0.1, r0.shape)) * s)
        dd.append(L * s)
    T = np.asarray(fr, dtype=np.float32)
    T[..., 1] = T[..., 1] / np.sin(th)
    T[..., 0] = T[..., 0] - T[..., 1] * np.cos(th)
    dims = np.asarray(dd, dtype=np.float32)
    _run(dens, "mono120_f32", T, dims, typ, ucell, 0.6,
         note="monoclinic 120 deg float32 coords+dims, 3 frames, rescale")

    # C: per-axis Sres array with USE_BETTER_RESOLUTION=False (explicit grid).
    L = np.array([6.0, 8.0, 10.0])
    Nw = np.array([12, 16, 20])
    r0, typ = _water_like(rng, 30, L, ["C", "H", "O"])
    frames = np.stack([r0, np.mod(r0 + rng.normal(0, 0.3, r0.shape), L)])
    dims = np.tile(L, (2, 1))
    _run(dens, "aniso_sres_array", frames, dims, typ, eye, L / Nw, better=False,
         note="Sres=L/N array (12,16,20), USE_BETTER_RESOLUTION=False")

    # D: a Na atom whose window (2A=44) exceeds the grid (N=24), Nb=22<=N.
    L = np.array([18.0, 18.0, 18.0])
    r0 = np.array([[2.0, 17.0, 9.0], [9.0, 9.0, 9.0], [0.5, 0.5, 17.5],
                   [16.0, 3.0, 1.0], [12.0, 1.0, 15.0]])
    typ = np.array(["Na", "O", "H", "C", "H"])
    frames = np.stack([r0, np.mod(r0 + rng.normal(0, 0.3, r0.shape), L)])
    dims = np.tile(L, (2, 1))
    _run(dens, "na_wide", frames, dims, typ, eye, 0.75,
         note="Na window 44 > N=24 (wraps twice); Nb=22")

    # E: RANDOM_NOISE branch with a fixed legacy-numpy seed.
    L = np.array([4.0, 4.0, 5.0])
    r0, typ = _water_like(rng, 6, L, ["H"])
    frames = np.stack([r0, r0])
    dims = np.tile(L, (2, 1))
    _run(dens, "random_noise", frames, dims, typ, eye, 0.5, noise=2, seed=7,
         note="RANDOM_NOISE=2, np.random.seed(7), N=(8,8,10)")

    # F: 24^3, float32, dims varying per frame, grid bumped by the resolution
    #    rule (23 -> 24), 60 atoms of H/C/O/N, 3 frames.
    L = np.array([11.7, 11.7, 11.7])
    r0, typ = _water_like(rng, 60, L, ["O", "H", "H", "C", "N", "H"])
    fr, dd = [], []
    for t in range(3):
        s = np.array([1.0 + 0.004 * t, 1.0 - 0.003 * t, 1.0 + 0.002 * t])
        fr.append(np.mod(r0 + rng.normal(0, 0.15, r0.shape), L) * s)
        dd.append(L * s)
    _run(dens, "mixed_f32_24", np.asarray(fr, np.float32),
         np.asarray(dd, np.float32), typ, eye, 0.5,
         note="24^3 float32, per-frame dims, N bumped 23->24")

    _clipped_windows(dens, eye)


if __name__ == "__main__":
    main()
