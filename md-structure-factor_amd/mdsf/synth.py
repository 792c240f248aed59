"""Seeded synthetic systems for the measured workloads (all lengths in Angstrom).

* ``water``  : SPC-geometry waters, uniform O, random orientations, fresh
               frames, written/read through nm float32 like a trajectory
               (load_traj.py:95-96: coords = xyz * 10 in float32).
* ``hii``    : inverted-hexagonal phase: 4 pore columns (radius 8) in a
               104 x 90 x 80 box lined with C/O/Na headgroups, C/H tails
               elsewhere, 50,000 atoms; frames = base + N(0, 0.2) noise.
* ``uniform``: uniform-random atoms of given types in a cubic box.
* ``adjoint``: 20,000 C/O/H atoms uniform in a 60 box, 16 frames of
               base + N(0, 0.2) noise, plus a target copy of the base
               perturbed by N(0, 0.5) (grid 128^3 at Sres = 60/128).
* ``large``  : 1,000,000 H/C/N/O atoms uniform in a 200 box, fresh frames
               (grid 512^3 at Sres = 200/512).
* ``batch``  : 32 independent cubic systems, grids {64..192}^3, boxes 30-80,
               5k-60k H/C/O/Na atoms, fresh frames.
"""
import numpy as np

SPC_OH = 1.0
SPC_HOH_DEG = 109.47


def _random_rotations(rng, n):
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1),
    ], 1)


def water_files(workdir, seed=1001, nwater=900, box=30.0, nframes=10):
    """Write the water workload as GROMACS files (nm): ``water.gro`` (topology,
    first frame) and ``water.trr`` (all frames, single precision).  Returns
    the two paths."""
    import os
    from . import trajio
    coords, dims, typ = water(seed, nwater, box, nframes)
    nm = coords / np.float32(10)
    cell = np.diag([box / 10.0] * 3)
    gro = os.path.join(workdir, "water.gro")
    trr = os.path.join(workdir, "water.trr")
    resids = [k // 3 + 1 for k in range(len(typ))]
    trajio.write_gro(gro, nm[:1], cell, typ, resnames=["SOL"] * len(typ), resids=resids,
                     title="water workload")
    trajio.write_trr(trr, nm, cell, times=np.arange(nframes, dtype=float))
    return gro, trr


def water(seed=1001, nwater=900, box=30.0, nframes=10):
    """Returns (coords float32 (F, 3*nwater, 3), dims float32 (F, 3), typ list)."""
    rng = np.random.default_rng(seed)
    half = np.radians(SPC_HOH_DEG) / 2
    geom = np.array([[0.0, 0.0, 0.0],
                     [SPC_OH * np.sin(half), SPC_OH * np.cos(half), 0.0],
                     [-SPC_OH * np.sin(half), SPC_OH * np.cos(half), 0.0]])
    frames = []
    for _ in range(nframes):
        o = rng.uniform(0.0, box, size=(nwater, 3))
        rot = _random_rotations(rng, nwater)
        mol = o[:, None, :] + np.einsum("nij,aj->nai", rot, geom)
        nm = (mol.reshape(-1, 3) / 10.0).astype(np.float32)   # as written to .gro/.trr
        frames.append(nm * np.float32(10))                     # as read back (float32)
    coords = np.stack(frames)
    dims = np.tile((np.array([box] * 3) / 10.0).astype(np.float32) * np.float32(10), (nframes, 1))
    typ = ["O", "H", "H"] * nwater
    return coords, dims, typ


HII_BOX = (104.0, 90.0, 80.0)
HII_AXES = ((13.0, 11.25), (65.0, 11.25), (39.0, 56.25), (91.0, 56.25))
HII_PORE = 8.0


def hii(seed=2002, nframes=200, natoms=50000, layers=20, per_layer=6, dtype=np.float32):
    """Returns (coords (F, natoms, 3), dims (F, 3) float64, typ list)."""
    rng = np.random.default_rng(seed)
    bx, by, bz = HII_BOX
    pos, typ = [], []
    for cx, cy in HII_AXES:
        for k in range(layers):
            z = (k + 0.5) * bz / layers
            for j in range(per_layer):
                phi = 2 * np.pi * j / per_layer + (np.pi / per_layer) * (k % 2)
                for name, rad, dphi in (("C", 8.0, 0.0), ("O", 7.5, 0.12), ("O", 7.5, -0.12),
                                        ("Na", 5.5, 0.0)):
                    rr = rad + rng.normal(0.0, 0.5)
                    pos.append(((cx + rr * np.cos(phi + dphi)) % bx,
                                (cy + rr * np.sin(phi + dphi)) % by, z))
                    typ.append(name)
    ntail = natoms - len(pos)
    tails = []
    while len(tails) < ntail:
        cand = rng.uniform(0.0, 1.0, size=(2 * ntail, 3)) * np.array(HII_BOX)
        ok = np.ones(len(cand), bool)
        for cx, cy in HII_AXES:
            dx = (cand[:, 0] - cx + bx / 2) % bx - bx / 2
            dy = (cand[:, 1] - cy + by / 2) % by - by / 2
            ok &= dx * dx + dy * dy > HII_PORE ** 2
        tails.extend(cand[ok][: ntail - len(tails)])
    pos.extend(tails)
    typ.extend(["C" if i % 3 == 0 else "H" for i in range(ntail)])
    base = np.asarray(pos, dtype=np.float64)
    coords = np.empty((nframes, natoms, 3), dtype=dtype)
    for t in range(nframes):
        coords[t] = np.mod(base + rng.normal(0.0, 0.2, size=base.shape), np.array(HII_BOX))
    dims = np.tile(np.array(HII_BOX), (nframes, 1))
    return coords, dims, typ


def hii_sres():
    """Per-axis resolution giving exactly 256^3 (USE_BETTER_RESOLUTION = False)."""
    return np.array(HII_BOX) / 256.0


def uniform(seed, natoms, box, names, nframes, dtype=np.float32):
    rng = np.random.default_rng(seed)
    typ = [names[i] for i in rng.integers(0, len(names), size=natoms)]
    coords = rng.uniform(0.0, box, size=(nframes, natoms, 3)).astype(dtype)
    dims = np.full((nframes, 3), float(box))
    return coords, dims, typ


ADJ_BOX = 60.0


def adjoint(seed=4004, target_seed=4005, natoms=20000, nframes=16, dtype=np.float64):
    """Returns (coords (F, natoms, 3), target coords (F, natoms, 3), dims (F, 3), typ)."""
    rng = np.random.default_rng(seed)
    names = ["C", "O", "H"]
    typ = [names[i] for i in rng.integers(0, 3, size=natoms)]
    base = rng.uniform(0.0, ADJ_BOX, size=(natoms, 3))
    coords = np.stack([np.mod(base + rng.normal(0.0, 0.2, size=base.shape), ADJ_BOX)
                       for _ in range(nframes)]).astype(dtype)
    trng = np.random.default_rng(target_seed)
    tbase = np.mod(base + trng.normal(0.0, 0.5, size=base.shape), ADJ_BOX)
    target = np.stack([np.mod(tbase + trng.normal(0.0, 0.2, size=base.shape), ADJ_BOX)
                       for _ in range(nframes)]).astype(dtype)
    dims = np.full((nframes, 3), ADJ_BOX)
    return coords, target, dims, typ


def adjoint_sres():
    return ADJ_BOX / 128.0


LARGE_BOX = 200.0


def large(seed=5005, natoms=1_000_000, nframes=1, dtype=np.float64):
    """1M uniform-random H/C/N/O atoms in a 200 box, fresh frames (grid 512^3 at Sres = 200/512)."""
    rng = np.random.default_rng(seed)
    names = ["H", "C", "N", "O"]
    typ = [names[i] for i in rng.integers(0, 4, size=natoms)]
    coords = rng.uniform(0.0, LARGE_BOX, size=(nframes, natoms, 3)).astype(dtype)
    dims = np.full((nframes, 3), LARGE_BOX)
    return coords, dims, typ


def large_sres():
    return LARGE_BOX / 512.0


BATCH_GRIDS = (64, 96, 128, 160, 192)


def batch(seed=3003, nsys=32, nframes=50, grids=BATCH_GRIDS, box=(30.0, 80.0),
          natoms=(5000, 60000), names=("H", "C", "O", "Na"), dtype=np.float32):
    """Independent cubic systems of uneven size: per system a grid n from
    ``grids``, a box L ~ U[box], natoms ~ U{natoms}, types uniform over
    ``names``, fresh uniform frames.  Returns a list of (coords, dims, typ, sres)
    with sres = L / (n + 1/2), so that int(L / sres) = n exactly."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nsys):
        n = int(rng.choice(grids))
        L = float(rng.uniform(*box))
        na = int(rng.integers(natoms[0], natoms[1] + 1))
        typ = [names[i] for i in rng.integers(0, len(names), size=na)]
        coords = rng.uniform(0.0, L, size=(nframes, na, 3)).astype(dtype)
        dims = np.full((nframes, 3), L)
        out.append((coords, dims, typ, L / (n + 0.5)))
    return out
