"""GPU parity of the paths the headline tests do not reach, against the oracle.

* wide borders on a small grid: N = 64, Nborder 58 (Na at Sres 0.3), so the
  corner-slot bands (dens.py:107) span 4 z tiles and the fold wraps windows
  nearly as wide as the cell;
* systems with 17-40 parameter sets (the per-type LDS layout sized for 64
  types, in the MFMA deposition, the xy-coupled deposition and the tiled
  adjoint);
* more than 4,096 spill tiles (N = 272, Nborder 136): the corner-slot spill
  is added to the density after the tiled pass instead of being staged;
* 1,000,000 atoms, the wrap branch of dens.py:213-221 that wraps only atom
  indices below the frame count: in-place coordinates, home cells, the
  density of the atoms left outside the cell and the range error;
* cells whose metric couples z with x or y (any 3x3 ucell, dens.py:301);
* the drop-in's rescale, remap_grid and remap_grid_tcl (dens.py:46-108);
* RANDOM_NOISE at FP64 (dens.py:279-280).

Tolerances as in test_gpu_parity.py: S(q) max relative error 1e-5 over voxels
above 1e-6 max(S); densities 1e-5 relative + 2e-6 of the maximum (FP32) or
1e-12 relative + 1e-14 of the maximum (FP64); home cells, in-place
coordinates and folds bit-exact.
"""
import math

import numpy as np
import pytest
import torch

from conftest import load_golden
from oracle import dens_oracle as O
from test_gpu_parity import _gpu, _oracle_frames, assert_density_close, assert_sq_close

pytestmark = pytest.mark.gpu


def _monoclinic(theta_deg):
    if theta_deg is None:
        return np.eye(3)
    th = math.radians(theta_deg)
    return np.array([[1, 0, 0], [np.cos(th), np.sin(th), 0], [0, 0, 1]])


def _triclinic():
    """Rows A, B, C of a triclinic cell with alpha = 80, beta = 105, gamma = 70
    degrees (unit edge lengths): the metric U U^T couples z with x and y."""
    from mdsf import gentraj
    return np.vstack(gentraj.box_vectors((1.0, 1.0, 1.0), (80.0, 105.0, 70.0)))


def _sq_dropin(coords, dims, typ, rad, sres, ucell, precision, tmp_path, better=False):
    import dens
    saved = (dens.PRINT_DETAILS, dens.FFT_PRECISION, dens.USE_BETTER_RESOLUTION)
    dens.PRINT_DETAILS, dens.USE_BETTER_RESOLUTION = False, better
    dens.FFT_PRECISION = precision
    try:
        r = np.array(coords)
        dens.compute_sf(r, dims, typ, str(tmp_path / "s"), rad, ucell, sres)
    finally:
        dens.PRINT_DETAILS, dens.FFT_PRECISION, dens.USE_BETTER_RESOLUTION = saved
    return r, dict(np.load(str(tmp_path / "s.npz")))


def _density_check(coords, dims, typ, rad, sres, ucell, fp64, frames=(0,)):
    args = (coords, dims, typ, rad, sres, False, ucell)
    s, plan, c, types, scale = _gpu(*args, fp64=fp64)
    _, o, ref = _oracle_frames(*args)
    for t in frames:
        home = torch.empty((s.natoms, 3), dtype=torch.int32, device=plan.device)
        d = plan.deposit_fold(c[t], types, scale[t], [float(x) for x in s.L], s.arith, s.wrap,
                              home=home)
        plan.check_errors()
        assert np.array_equal(home.cpu().numpy(), ref[t][1]), "home cells differ"
        assert_density_close(d.cpu().numpy().astype(np.float64), ref[t][0], fp64)
    plan.close()
    return s, o


# --------------------------------------------------------------- wide borders
def _wide_border_system(seed, frames=2):
    """64^3 at Sres 0.3 (L = 19.2 A): Na (half-width 58) at all 8 corners and
    inside, C / O / H elsewhere."""
    rng = np.random.default_rng(seed)
    L = 19.2
    pts, typ = [], []
    for k in range(8):
        corner = np.array([(k >> 2) & 1, (k >> 1) & 1, k & 1], dtype=np.float64) * L
        pts.append(corner + rng.uniform(-1.0, 1.0, 3))
        typ.append("Na")
    for name in ("C", "O", "H", "Na", "C", "H"):
        for _ in range(4):
            pts.append(rng.uniform(0, L, 3))
            typ.append(name)
    base = np.mod(np.array(pts), L)
    coords = np.stack([np.mod(base + rng.normal(0, 0.1, base.shape), L)
                       for _ in range(frames)]).astype(np.float32)
    dims = np.full((frames, 3), L)
    return coords, dims, typ


@pytest.mark.parametrize("fp64", [False, True])
def test_wide_border_small_grid(fp64, rad, gpu, tmp_path):
    coords, dims, typ = _wide_border_system(21)
    s, o = _density_check(coords, dims, typ, rad, 0.3, np.eye(3), fp64, frames=(0, 1))
    assert list(s.N) == [64, 64, 64] and s.nborder >= 56
    want = O.compute_sf(np.array(coords), np.array(dims), typ, None, rad, np.eye(3), 0.3,
                        better=False)["sf"]
    _, got = _sq_dropin(coords, dims, typ, rad, 0.3, np.eye(3), "fp64" if fp64 else "fp32",
                        tmp_path)
    err = assert_sq_close(got["sf"], want)
    print(f"wide border ({'fp64' if fp64 else 'fp32'}): S(q) max rel {err.max():.2e}")


# ------------------------------------------------------------ many types
_MANY = ["H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Mg", "Al", "Si", "P", "S", "Cl",
         "Ar", "Ca", "Sc", "Ti", "V", "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As",
         "Se", "Br", "Kr", "Sr", "Zr", "Mo", "Ag", "Sn", "Xe"]


def _many_types_system(seed, frames=2, box=30.0):
    rng = np.random.default_rng(seed)
    typ = [_MANY[i % len(_MANY)] for i in range(3 * len(_MANY))]
    base = rng.uniform(0, box, (len(typ), 3))
    base[:8] = np.mod(np.array([[(k >> 2) & 1, (k >> 1) & 1, k & 1] for k in range(8)]) * box
                      + rng.uniform(-0.8, 0.8, (8, 3)), box)       # corner atoms
    coords = np.stack([np.mod(base + rng.normal(0, 0.05, base.shape), box)
                       for _ in range(frames)])
    dims = np.full((frames, 3), box)
    return coords, dims, typ


@pytest.mark.parametrize("fp64,theta", [(False, None), (False, 120.0), (True, None)])
def test_many_parameter_sets(fp64, theta, rad, gpu, tmp_path):
    from mdsf import engine
    coords, dims, typ = _many_types_system(31)
    ucell = _monoclinic(theta)
    s = engine.Setup(coords, dims, typ, rad, 0.6, False, ucell)
    assert len(s.amp) == 40 and len(s.amp) > 16
    _density_check(coords, dims, typ, rad, 0.6, ucell, fp64)
    want = O.compute_sf(np.array(coords), np.array(dims), typ, None, rad, ucell, 0.6,
                        better=False)["sf"]
    _, got = _sq_dropin(coords, dims, typ, rad, 0.6, ucell, "fp64" if fp64 else "fp32",
                        tmp_path)
    assert_sq_close(got["sf"], want)


def test_many_parameter_sets_tiled_adjoint(rad, gpu):
    """40 parameter sets through the tiled MFMA adjoint (FP32, orthogonal cell)
    against autograd of the torch reference."""
    import sf_torch_ref as T
    from mdsf.autograd import SFModel
    coords, dims, typ = _many_types_system(32, frames=2)
    model = SFModel(coords, dims, typ, rad, 0.6, fp64=False)
    assert model.plan.natoms == len(typ)
    rng = np.random.default_rng(9)
    x = torch.tensor(coords, device=model.plan.device, requires_grad=True)
    S = model(x)
    w = torch.tensor(rng.normal(size=tuple(S.shape)), device=S.device) / S.detach().max()
    torch.sum(w * S).backward()
    got = x.grad.cpu().numpy()
    xr = torch.tensor(coords, requires_grad=True)
    Sr = T.structure_factor(xr, dims, typ, rad, 0.6, np.eye(3))
    torch.sum(w.cpu() * Sr).backward()
    want = xr.grad.numpy()
    err = float(np.max(np.abs(got - want)) / np.max(np.abs(want)))
    assert err <= 2e-4, err


# --------------------------------------------------- in-place spill (> 4096 tiles)
@pytest.mark.parametrize("fp64", [False, True])
def test_inplace_spill_many_spill_tiles(fp64, rad, gpu):
    """N = 272, Nborder 134 (one Na, half-width 134): the spill tiles of an
    axis, [0, ceil(Nb / 16)) and [floor((N - Nb) / 16), 17), cover all 17
    tiles, so there are 17^3 = 4,913 > 4,096 of them and the corner-slot spill
    is added in place after the tiled deposition.  H / C atoms at the 8
    corners spill."""
    rng = np.random.default_rng(41)
    L = 272 * 0.128
    pts, typ = [np.array([L / 2, L / 2, L / 2])], ["Na"]
    for k in range(8):
        corner = np.array([(k >> 2) & 1, (k >> 1) & 1, k & 1], dtype=np.float64) * L
        for name in ("H", "C"):
            pts.append(corner + rng.uniform(-1.0, 1.0, 3))
            typ.append(name)
    coords = np.mod(np.array(pts), L)[None].astype(np.float32)
    dims = np.full((1, 3), L)
    s, o = _density_check(coords, dims, typ, rad, L / 272, np.eye(3), fp64)
    assert list(s.N) == [272, 272, 272] and s.nborder == 134


# ------------------------------------------------------------ 1e6 atoms
def _million(rng, nframes=2, L=100.0):
    natoms = 1_000_000
    typ = np.array(["H"] * natoms, dtype="<U2")
    typ[::100000] = "Na"                                # Nborder 22 at 128^3
    coords = rng.uniform(0.5, L - 0.5, (nframes, natoms, 3)).astype(np.float32)
    return coords, np.full((nframes, 3), L), typ


def test_million_atoms_wrap_branch(rad, gpu):
    """At 1e6 atoms the reference wraps only atom indices [0, frames)
    (dens.py:213-221).  Atoms 0 and 1 (wrapped) and, at indices >= frames (not
    wrapped), atoms outside the cell: within the border (deposited from the
    padding), across the padded grid's low edge (numpy's slice comes out
    empty), wholly at negative padded indices (counted from the far end,
    where the Gaussian underflows), cut at the high edge (the kept part
    deposits), past the high edge (empty), and a NaN coordinate (empty
    slices).  The reference raises for none of them.  Checked: the in-place
    coordinates and the last frame's home cells bit for bit against the
    oracle (home cells of the NaN atom excepted: numpy gives INT64_MIN, the
    int32 output INT32_MIN); the density those atoms add (the FP64 frame
    density minus that of the same frame without them) against the oracle's
    deposition + fold of them alone."""
    from mdsf import engine
    rng = np.random.default_rng(51)
    coords, dims, typ = _million(rng)
    L = 100.0
    dr = L / 128
    out = [(0, [-0.5, 3.0, 50.0]), (1, [100.25, 99.0, 100.5])]
    ext = [(500_001 + k, [-5.0 + 0.3 * k, 104.0 - 0.2 * k, 50.0 + k]) for k in range(10)]
    ext += [(600_001, [-20.5 * dr, 50.0, 50.0]),      # x0 < 0 < x1: empty slice
            (600_002, [-40.5 * dr, 40.0, 60.0]),      # x1 < 0: the far end of the grid
            (600_003, [50.0, 147.5 * dr, 30.0]),      # y cut at the high edge
            (600_004, [160.5 * dr, 50.0, 50.0]),      # past the high edge: empty
            (600_005, [150.2 * dr, 148.9 * dr, -23.7 * dr]),  # cut on x, y; z low side
            (600_006, [float("nan"), 50.0, 50.0])]
    for i, p in out + ext:
        coords[:, i] = np.array(p, dtype=np.float32)
        typ[i] = "H"
    sres = L / 128
    s = engine.Setup(coords, dims, typ, rad, sres, False, np.eye(3))
    assert s.wrap == (0, 2) and s.nborder == 22
    ro = np.array(coords)
    O.rescale(ro, np.array(dims))
    O.wrap(ro, s.L)
    plan = engine.Plan.from_setup(s, fft64=True)
    dev = plan.device
    c = torch.from_numpy(coords).to(dev)
    types = torch.from_numpy(s.type_index.astype(np.int32)).to(dev)
    scale = torch.from_numpy(np.ascontiguousarray(s.scale, dtype=np.float64)).to(dev)
    home = torch.empty((s.natoms, 3), dtype=torch.int32, device=dev)
    plan.accumulate_frames(c, types, scale, [float(x) for x in s.L], s.arith, s.wrap, home=home)
    plan.check_errors()
    assert np.array_equal(c.cpu().numpy(), ro, equal_nan=True), "in-place coordinates differ"
    h = home.cpu().numpy()
    with np.errstate(invalid="ignore"):
        want_h = O.home_cells(ro[-1], s.dr)
    ok = np.isfinite(ro[-1]).all(axis=1)
    assert np.array_equal(h[ok], want_h[ok]), "home cells differ"
    assert np.all(h[~ok][:, 0] == np.iinfo(np.int32).min)
    for i, _ in ext[:10]:
        assert ro[0, i, 0] < 0 or ro[0, i, 0] >= L        # left outside by the reference
    # the outside atoms' density: full frame minus the frame without them
    c0 = torch.from_numpy(np.ascontiguousarray(ro[:1])).to(dev)
    full = plan.deposit_fold(c0[0], types, torch.ones_like(scale[0]), [float(x) for x in s.L],
                             s.arith, (0, 0))
    plan.check_errors()
    idx = [i for i, _ in ext]
    keep = np.ones(s.natoms, bool)
    keep[idx] = False
    sub = engine.Setup(ro[:1, keep], dims[:1], list(typ[keep]), rad, sres, False, np.eye(3))
    assert sub.nborder == s.nborder and list(sub.N) == list(s.N)
    psub = engine.Plan.from_setup(sub, fft64=True)
    csub = torch.from_numpy(np.ascontiguousarray(ro[:1, keep])).to(dev)
    tsub = torch.from_numpy(sub.type_index.astype(np.int32)).to(dev)
    dsub = psub.deposit_fold(csub[0], tsub, torch.ones_like(scale[0]), [float(x) for x in s.L],
                             sub.arith, (0, 0))
    psub.check_errors()
    diff = (full - dsub).cpu().numpy()
    o = O.setup(np.array(ro[:1, :3]), np.array(dims[:1]), ["H", "H", "Na"], rad, sres, False)
    assert o["Nb"] == s.nborder
    padded = tuple(int(o["N"][d]) + 2 * o["Nb"] for d in range(3))
    with np.errstate(invalid="ignore"):
        d0, _ = O.deposit_frame(ro[0, idx], [str(typ[i]) for i in idx], rad, o["bdict"],
                                o["dr"], o["Nb"], np.eye(3), padded)
    want = O.fold(d0, *O.fold_tables(o["N"], o["Nb"]))
    scale_max = float(np.max(np.abs(full.cpu().numpy())))
    assert np.max(np.abs(diff - want)) <= 1e-10 * scale_max
    assert np.max(want) > 1e-3 * scale_max           # the outside atoms are visible
    plan.close()
    psub.close()


# ------------------------------------------------------------ general cell
@pytest.mark.parametrize("fp64", [False, True])
def test_general_cell_density_and_sq(fp64, rad, gpu, tmp_path):
    """A triclinic ucell (alpha 80, beta 105, gamma 70 degrees) with atoms at
    all 8 corners and Nborder 17 at N ~ 40: the per-voxel gather and the
    corner-slot spill with the z couplings, against the oracle."""
    from test_adjoint_ref import spill_system
    coords, dims, typ = spill_system(61, frames=2)
    ucell = _triclinic()
    g = ucell @ ucell.T
    assert abs(g[0, 2]) > 0.1 and abs(g[1, 2]) > 0.1
    _density_check(coords, dims, typ, rad, 1.0, ucell, fp64, frames=(0, 1))
    want = O.compute_sf(np.array(coords), np.array(dims), typ, None, rad, ucell, 1.0,
                        better=False)["sf"]
    _, got = _sq_dropin(coords, dims, typ, rad, 1.0, ucell, "fp64" if fp64 else "fp32", tmp_path)
    assert_sq_close(got["sf"], want)


def test_general_cell_adjoint_rejected(rad, gpu):
    from mdsf.autograd import SFModel
    from test_adjoint_ref import spill_system
    coords, dims, typ = spill_system(62, frames=1)
    model = SFModel(coords, dims, typ, rad, 1.0, ucell=_triclinic(), fp64=True)
    x = torch.tensor(coords, device=model.plan.device, requires_grad=True)
    S = model(x)
    with pytest.raises(NotImplementedError):
        torch.sum(S).backward()


# ------------------------------------------------------------ drop-in helpers
@pytest.mark.parametrize("cdtype,ddtype", [(np.float32, np.float32), (np.float32, np.float64),
                                           (np.float64, np.float64)])
def test_dropin_rescale(cdtype, ddtype, gpu):
    import dens
    rng = np.random.default_rng(71)
    coords = rng.uniform(0, 30, (4, 500, 3)).astype(cdtype)
    dims = (30.0 + rng.normal(0, 0.3, (4, 3))).astype(ddtype)
    want = np.array(coords)
    _, avg_want = O.rescale(want, dims)
    got = np.array(coords)
    out, avg = dens.rescale(got, dims)
    assert out is got and got.dtype == cdtype
    assert np.array_equal(got, want)
    assert np.array_equal(avg, avg_want) and avg.dtype == avg_want.dtype


@pytest.mark.parametrize("n,nb", [((8, 10, 12), 2), ((8, 8, 8), 3), ((6, 8, 10), 4),
                                  ((8, 6, 6), 5), ((10, 10, 10), 0), ((16, 12, 14), 12)])
def test_dropin_remap_grid_tcl_bitwise(n, nb, gpu):
    import dens
    rng = np.random.default_rng(sum(n) + nb)
    d0 = rng.normal(size=tuple(x + 2 * nb for x in n))
    des, ori = O.fold_tables(np.array(n), nb)
    got = dens.remap_grid_tcl(d0, des, ori)
    want = O.fold(d0, des, ori)
    assert got.dtype == np.float64 and np.array_equal(got, want)
    assert abs(got.sum() - d0.sum()) <= 1e-9 * np.abs(d0).sum()


def test_dropin_remap_grid_cubic_and_bad_tables(gpu):
    import dens
    rng = np.random.default_rng(3)
    n, nb = 10, 3
    d0 = rng.normal(size=(n + 2 * nb,) * 3)
    des = [0, nb, n - nb, n]
    ori = [0, nb, nb + n, 2 * nb + n]
    got = dens.remap_grid(d0, des, ori)
    assert np.array_equal(got, O.fold(d0, [des] * 3, [ori] * 3))
    with pytest.raises(ValueError):
        dens.remap_grid(d0, [0, 2, 7, 10], ori)


def test_random_noise_fp64_plan(rad, gpu, tmp_path):
    """RANDOM_NOISE (dens.py:267-280) on the FP64 plan: numpy's FP64 noise
    enters unrounded, so S(q) matches the reference's golden to FP64 rounding."""
    import dens
    c = load_golden("random_noise")
    saved = (dens.RANDOM_NOISE, dens.PRINT_DETAILS, dens.FFT_PRECISION,
             dens.USE_BETTER_RESOLUTION)
    dens.RANDOM_NOISE, dens.PRINT_DETAILS, dens.FFT_PRECISION = int(c["in_noise"]), False, "fp64"
    dens.USE_BETTER_RESOLUTION = bool(c["in_better"])
    try:
        np.random.seed(int(c["in_seed"]))
        r = np.array(c["in_coords"])
        dens.compute_sf(r, np.array(c["in_dims"]), c["typ"], str(tmp_path / "n"), rad,
                        c["in_ucell"], c["sres"])
    finally:
        (dens.RANDOM_NOISE, dens.PRINT_DETAILS, dens.FFT_PRECISION,
         dens.USE_BETTER_RESOLUTION) = saved
    got = np.load(str(tmp_path / "n.npz"))["sf"]
    assert_sq_close(got, c["out_sf"], rel=1e-11)


# ----------------------------------------------------------------- tile lists
def test_tile_list_deposition_reproducible(rad, gpu):
    """The FP32 orthogonal-cell deposition reads per-tile atom lists built in
    the prepare stage (regular atoms upward, corner atoms downward in each
    tile's region).  A mixed system with small windows (bucket edge 16) and
    Na windows wider than 3 tiles (bucket edge 32, up to 8 slices per axis):
    the folded density of a frame matches the oracle, and depositing it again
    gives the same bits (the list order, hence the sum order, is fixed)."""
    from mdsf import synth
    coords, dims, typ = synth.uniform(707, 1000, 38.4, ["H", "C", "O"] * 10 + ["Na"], 1)
    args = (coords, dims, typ, rad, 0.3, False, np.eye(3))
    s, plan, c, types, scale = _gpu(*args, fp64=False)
    assert max(int(h) for h in np.max(s.half, axis=1)) >= 24  # a type with 32-cell buckets
    _, o, ref = _oracle_frames(*args)
    box = [float(x) for x in s.L]
    c0 = c[0].clone()
    d1 = plan.deposit_fold(c[0], types, scale[0], box, s.arith, s.wrap).cpu().numpy()
    d2 = plan.deposit_fold(c0, types, scale[0], box, s.arith, s.wrap).cpu().numpy()
    plan.check_errors()
    assert np.array_equal(d1, d2)
    assert_density_close(d1.astype(np.float64), ref[0][0], False)
    plan.close()
