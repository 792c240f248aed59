"""GPU parity: the HIP path (through the C ABI) against the oracle and the
reference's own golden outputs.

Tolerances (FP32 deposition and FFT, FP64 frame sum):
  * S(q): max relative error <= 1e-5 on voxels above 1e-6 * max(S);
  * folded density: |d - d_oracle| <= 1e-5 * |d_oracle| + 2e-6 * max(d_oracle);
  * home cells, occupancy, grid sizes, k-grid axes, in-place coordinates: bit-exact.
"""
import math
import os

import numpy as np
import pytest

from conftest import golden_cases, load_golden
from oracle import dens_oracle as O

pytestmark = pytest.mark.gpu

SQ_REL = 1e-5
SQ_FLOOR = 1e-6


def assert_sq_close(got, want, rel=SQ_REL, floor=SQ_FLOOR):
    assert got.shape == want.shape
    m = np.max(want)
    mask = want > floor * m
    err = np.abs(got[mask] - want[mask]) / want[mask]
    assert err.max() <= rel, f"max rel err {err.max():.3e} (median {np.median(err):.3e})"
    return err


def assert_density_close(got, want, fp64=False):
    if fp64:
        tol = 1e-12 * np.abs(want) + 1e-14 * np.max(np.abs(want))
    else:
        tol = 1e-5 * np.abs(want) + 2e-6 * np.max(np.abs(want))
    bad = np.abs(got - want) > tol
    assert not bad.any(), (f"{bad.sum()} voxels off; max abs err "
                           f"{np.max(np.abs(got - want)):.3e}, max {np.max(want):.3e}")


def _gpu(r, dims, typ, rad, sres, better, ucell, fp64=False):
    import torch
    from mdsf import engine
    s = engine.Setup(r, dims, typ, rad, sres, better, ucell)
    plan = engine.Plan.from_setup(s, fft64=fp64)
    coords = torch.from_numpy(np.ascontiguousarray(r)).to(plan.device)
    types = torch.from_numpy(s.type_index.astype(np.int32).reshape(-1)).to(plan.device)
    scale = torch.from_numpy(np.ascontiguousarray(s.scale, dtype=np.float64)).to(plan.device)
    return s, plan, coords, types, scale


def _oracle_frames(r, dims, typ, rad, sres, better, ucell):
    rr = np.array(r)
    o = O.setup(rr, np.array(dims), typ, rad, sres, better)
    padded = tuple(int(o["N"][i]) + 2 * o["Nb"] for i in range(3))
    des, ori = O.fold_tables(o["N"], o["Nb"])
    out = []
    for t in range(rr.shape[0]):
        d0, ir = O.deposit_frame(rr[t], typ, rad, o["bdict"], o["dr"], o["Nb"], ucell, padded)
        out.append((O.fold(d0, des, ori), ir))
    return rr, o, out


def _run_dropin(c, rad, tmp_path, precision="auto"):
    import dens
    saved = (dens.USE_BETTER_RESOLUTION, dens.RANDOM_NOISE, dens.PRINT_DETAILS,
             dens.FFT_PRECISION)
    dens.FFT_PRECISION = precision
    dens.USE_BETTER_RESOLUTION = bool(c["in_better"])
    dens.RANDOM_NOISE = int(c["in_noise"])
    dens.PRINT_DETAILS = False
    try:
        if int(c["in_seed"]) >= 0:
            np.random.seed(int(c["in_seed"]))
        r = np.array(c["in_coords"])
        out = str(tmp_path / "sf")
        dens.compute_sf(r, np.array(c["in_dims"]), c["typ"], out, rad, c["in_ucell"], c["sres"])
        return r, dict(np.load(out + ".npz"))
    finally:
        (dens.USE_BETTER_RESOLUTION, dens.RANDOM_NOISE, dens.PRINT_DETAILS,
         dens.FFT_PRECISION) = saved


@pytest.mark.parametrize("precision", ["auto", "fp32", "fp64"])
@pytest.mark.parametrize("name", golden_cases())
def test_dropin_matches_reference_goldens(name, precision, rad, gpu, tmp_path):
    """FP32 storage (FP64 arithmetic in every FFT pass, rounding residuals of
    the z and y passes), FP64 throughout and the default "auto" choice between
    them, all at the same 1e-5 bound on these tiny sparse goldens (numpy's
    single-precision pocketfft alone reaches 1.2e-5 .. 5e-5 on them)."""
    c = load_golden(name)
    r, out = _run_dropin(c, rad, tmp_path, precision=precision)
    assert sorted(out) == sorted(["sf", "sfplt", "L", "N", "kgrid", "kgridplt"])
    assert r.dtype == c["out_r"].dtype and np.array_equal(r, c["out_r"])
    assert np.array_equal(out["N"], c["out_N"]) and out["N"].dtype == c["out_N"].dtype
    assert out["L"].dtype == c["out_L"].dtype and np.array_equal(out["L"], c["out_L"])
    assert np.array_equal(out["kgrid"], c["out_kgrid"])
    assert np.array_equal(out["kgridplt"][..., :3], c["out_kgridplt"][..., :3])
    err = assert_sq_close(out["sf"], c["out_sf"])
    # sfplt drops the q = 0 peak (its centre is the mean of three neighbours),
    # so its own maximum is far below max(sf): the same voxels as for sf
    floor = SQ_FLOOR * float(np.max(c["out_sf"])) / float(np.max(c["out_sfplt"]))
    assert_sq_close(out["sfplt"], c["out_sfplt"], floor=floor)
    print(f"{name} {precision}: S(q) max rel err {err.max():.2e}")
    # the derived grids are exact functions of our own sf
    assert np.array_equal(out["sfplt"], O.get_dplot(out["sf"]))
    assert np.array_equal(out["kgridplt"][..., 3], out["sfplt"])


@pytest.mark.parametrize("fp64", [False, True])
@pytest.mark.parametrize("name", ["orth_corners_f64", "mono120_f32", "na_wide",
                                  "aniso_sres_array", "mixed_f32_24", "clipped_windows"])
def test_deposit_fold_and_home_cells(name, fp64, rad, gpu):
    import torch
    c = load_golden(name)
    ucell = c["in_ucell"]
    args = (c["in_coords"], c["in_dims"], c["typ"], rad, c["sres"], bool(c["in_better"]), ucell)
    s, plan, coords, types, scale = _gpu(*args, fp64=fp64)
    _, o, ref = _oracle_frames(*args)
    for t in range(coords.shape[0]):
        home = torch.empty((s.natoms, 3), dtype=torch.int32, device=plan.device)
        d = plan.deposit_fold(coords[t], types, scale[t], [float(x) for x in s.L], s.arith,
                              s.wrap, home=home)
        plan.check_errors()
        assert np.array_equal(home.cpu().numpy(), ref[t][1]), "home cells differ"
        assert_density_close(d.cpu().numpy().astype(np.float64), ref[t][0], fp64)


def _corner_case(rng, box, n_each, names, theta_deg=None, dtype=np.float64):
    """Atoms within a few Angstrom of all 8 corners plus a few interior ones."""
    pts = []
    for cx in (0.05, 0.9):
        for cy in (0.1, 0.95):
            for cz in (0.02, 0.97):
                for _ in range(n_each):
                    pts.append(np.array([cx, cy, cz]) * box + rng.normal(0, 0.3, 3))
    pts += list(rng.uniform(0, 1, (5, 3)) * box)
    r = np.mod(np.array(pts), box)[None].astype(dtype)
    typ = [names[i % len(names)] for i in range(r.shape[1])]
    if theta_deg is None:
        ucell = np.eye(3)
    else:
        th = math.radians(theta_deg)
        ucell = np.array([[1, 0, 0], [np.cos(th), np.sin(th), 0], [0, 0, 1]])
    return r, np.array([box]), typ, ucell


@pytest.mark.parametrize("fp64", [False, True])
@pytest.mark.parametrize("theta", [None, 120.0, 90.0])
@pytest.mark.parametrize("sres", [0.5, 0.3])
def test_corner_fold_quirk(theta, sres, fp64, rad, gpu):
    rng = np.random.default_rng(11)
    box = np.array([7.0, 8.0, 9.0])
    r, dims, typ, ucell = _corner_case(rng, box, 2, ["C", "O", "H", "N"], theta)
    args = (r, dims, typ, rad, sres, True, ucell)
    s, plan, coords, types, scale = _gpu(*args, fp64=fp64)
    _, o, ref = _oracle_frames(*args)
    d = plan.deposit_fold(coords[0], types, scale[0], [float(x) for x in s.L], s.arith, s.wrap)
    plan.check_errors()
    assert_density_close(d.cpu().numpy().astype(np.float64), ref[0][0], fp64)


@pytest.mark.parametrize("fp64", [False, True])
@pytest.mark.parametrize("theta", [None, 120.0])
def test_corner_fold_spill_beyond_window_tiles(theta, fp64, rad, gpu, tmp_path):
    """Corner-slot contributions (dens.py:107) that land in z tiles the atom's
    own window does not reach: N = (40, 44, 42), Nborder = 17 (Na), atoms at
    all 8 corners (316 such images, counted on the oracle's set-up).  Folded
    density against the oracle, then S(q) of 3 frames through the drop-in's
    pipelined frame loop (the staged spill alternates between frame slots)."""
    from test_adjoint_ref import spill_images, spill_system
    import dens
    coords, dims, typ = spill_system(11, frames=3)
    assert spill_images(coords[0], typ, rad, 1.0, dims) > 100
    ucell = np.eye(3) if theta is None else np.array(
        [[1, 0, 0], [np.cos(math.radians(theta)), np.sin(math.radians(theta)), 0], [0, 0, 1]])
    args = (coords, dims, typ, rad, 1.0, False, ucell)
    s, plan, c, types, scale = _gpu(*args, fp64=fp64)
    _, o, ref = _oracle_frames(*args)
    d = plan.deposit_fold(c[0], types, scale[0], [float(x) for x in s.L], s.arith, s.wrap)
    plan.check_errors()
    assert_density_close(d.cpu().numpy().astype(np.float64), ref[0][0], fp64)
    want = O.compute_sf(np.array(coords), np.array(dims), typ, None, rad, ucell, 1.0,
                        better=False)["sf"]
    saved = (dens.PRINT_DETAILS, dens.FFT_PRECISION, dens.USE_BETTER_RESOLUTION)
    dens.PRINT_DETAILS, dens.USE_BETTER_RESOLUTION = False, False
    dens.FFT_PRECISION = "fp64" if fp64 else "fp32"
    try:
        dens.compute_sf(np.array(coords), dims, typ, str(tmp_path / "s"), rad, ucell, 1.0)
    finally:
        dens.PRINT_DETAILS, dens.FFT_PRECISION, dens.USE_BETTER_RESOLUTION = saved
    got = np.load(str(tmp_path / "s.npz"))["sf"]
    assert_sq_close(got, want)


@pytest.mark.parametrize("fp64", [False, True])
@pytest.mark.parametrize("theta", [None, 120.0])
def test_corner_fold_spill_many_corner_atoms(theta, fp64, rad, gpu, tmp_path):
    """640 atoms at the 8 corners: more corner atoms than the spill pass lists
    per frame (512), so its candidates come from the per-tile enumeration.
    Folded density of frame 0, then S(q) of 2 frames through the frame loop.

    The density sits in a few hundred voxels, so FP32 storage of the exact
    density alone is worth several 1e-6 on the faintest voxels above the
    floor (printed): an FP32 plan is held to 5 times that floor, and "auto"
    must route this input to FP64 and meet 1e-5."""
    from test_adjoint_ref import spill_images, spill_system
    import dens
    coords, dims, typ = spill_system(13, frames=2, per_corner=80)
    assert spill_images(coords[0], typ, rad, 1.0, dims) > 5000  # 635 corner atoms
    ucell = np.eye(3) if theta is None else np.array(
        [[1, 0, 0], [np.cos(math.radians(theta)), np.sin(math.radians(theta)), 0], [0, 0, 1]])
    args = (coords, dims, typ, rad, 1.0, False, ucell)
    s, plan, c, types, scale = _gpu(*args, fp64=fp64)
    _, o, ref = _oracle_frames(*args)
    d = plan.deposit_fold(c[0], types, scale[0], [float(x) for x in s.L], s.arith, s.wrap)
    plan.check_errors()
    assert_density_close(d.cpu().numpy().astype(np.float64), ref[0][0], fp64)
    w = O.compute_sf(np.array(coords), np.array(dims), typ, None, rad, ucell, 1.0,
                     better=False, keep=True)
    want = w["sf"]

    def dropin(precision, name):
        saved = (dens.PRINT_DETAILS, dens.FFT_PRECISION, dens.USE_BETTER_RESOLUTION)
        dens.PRINT_DETAILS, dens.USE_BETTER_RESOLUTION = False, False
        dens.FFT_PRECISION = precision
        try:
            dens.compute_sf(np.array(coords), dims, typ, str(tmp_path / name), rad, ucell, 1.0)
        finally:
            dens.PRINT_DETAILS, dens.FFT_PRECISION, dens.USE_BETTER_RESOLUTION = saved
        return np.load(str(tmp_path / (name + ".npz")))["sf"]

    if fp64:
        assert_sq_close(dropin("fp64", "s"), want)
        return
    s32 = sum(np.abs(np.fft.rfftn(x.astype(np.float32).astype(np.float64))) ** 2 for x in w["d1"])
    mask = want > SQ_FLOOR * want.max()
    floor32 = float(np.max(np.abs(s32[mask] - want[mask]) / want[mask]))
    err = assert_sq_close(dropin("fp32", "s"), want, rel=5.0 * floor32)
    print(f"FP32 storage of the exact density: {floor32:.2e}; FP32 plan: {err.max():.2e}")
    assert floor32 > 2e-6
    err = assert_sq_close(dropin("auto", "a"), want)
    assert dens.LAST_PRECISION["fp64"], dens.LAST_PRECISION
    assert dens.LAST_PRECISION["frame0_fp32_err"] > dens.AUTO_FP32_MAX_ERR


@pytest.mark.parametrize("fp64", [False, True])
def test_empty_system_deposits_zeros(fp64, gpu):
    """No atoms, Nborder 17 on a 32^3 grid (every tile a spill tile): the
    density is exactly zero, twice in a row on the same plan."""
    import torch
    from mdsf import engine
    plan = engine.Plan([32, 32, 32], 17, [1.0, 1.0, 1.0], np.eye(3), [1.0], [0.5], [[4, 4, 4]],
                       [0], fft64=fp64)
    coords = torch.zeros((0, 3), dtype=torch.float32, device=plan.device)
    types = torch.zeros((0,), dtype=torch.int32, device=plan.device)
    scale = torch.ones(3, dtype=torch.float64, device=plan.device)
    for _ in range(2):
        d = plan.deposit_fold(coords, types, scale, [32.0, 32.0, 32.0], 0, (0, 0))
        plan.check_errors()
        assert int(torch.count_nonzero(d)) == 0


def test_occupancy_histogram_bit_exact(rad, gpu):
    import torch
    from mdsf import synth
    coords, dims, typ = synth.water(seed=5, nwater=300, box=18.0, nframes=3)
    s, plan, c, types, scale = _gpu(coords, dims, typ, rad, 0.5, True, np.eye(3))
    occ = torch.zeros(tuple(int(x) for x in s.N), dtype=torch.int32, device=plan.device)
    plan.accumulate_frames(c, types, scale, [float(x) for x in s.L], s.arith, s.wrap,
                           occupancy=occ)
    plan.check_errors()
    _, o, ref = _oracle_frames(coords, dims, typ, rad, 0.5, True, np.eye(3))
    want = np.zeros(tuple(int(x) for x in s.N), dtype=np.int64)
    for _, ir in ref:
        np.add.at(want, tuple((ir % o["N"]).T), 1)
    assert np.array_equal(occ.cpu().numpy().astype(np.int64), want)


@pytest.mark.parametrize("fft64", [False, True])
@pytest.mark.parametrize("shape", [(8, 8, 8), (16, 12, 10), (64, 64, 64), (96, 160, 192),
                                   (38, 14, 22), (30, 18, 50), (2, 4, 2), (256, 256, 256),
                                   # register-staged z pass, Nz/2 = 64 / 128 / 256, with a
                                   # partial last block of rows
                                   (6, 10, 128), (6, 6, 256), (2, 6, 512)])
def test_rfft3_matches_numpy(shape, fft64, gpu):
    import torch
    from mdsf import engine
    rng = np.random.default_rng(7)
    x = rng.random(shape).astype(np.float32) - 0.5
    plan = engine.Plan(shape, 0, [1.0] * 3, np.eye(3), [], [], [], [], fft64=fft64)
    got = plan.rfft3(torch.from_numpy(x)).cpu().numpy()
    want = np.fft.rfftn(x.astype(np.float64))
    # error against the rms spectrum amplitude (Parseval scale of a zero-mean input)
    scale = np.sqrt(np.mean(np.abs(want) ** 2))
    err = np.abs(got - want).max() / scale
    assert err < (1e-12 if fft64 else 2e-6) * np.log2(x.size), err


def test_accumulate_density_power_sum(gpu):
    import torch
    from mdsf import engine
    rng = np.random.default_rng(8)
    shape = (24, 20, 16)
    plan = engine.Plan(shape, 0, [1.0] * 3, np.eye(3), [], [], [], [])
    want = np.zeros((24, 20, 9))
    for _ in range(3):
        x = rng.random(shape).astype(np.float32)
        plan.accumulate_density(torch.from_numpy(x))
        want += O.frame_power(x.astype(np.float64))
    assert_sq_close(plan.sf().cpu().numpy(), want)


def test_determinism_bitwise(rad, gpu):
    from mdsf import synth
    coords, dims, typ = synth.water(seed=9, nwater=400, box=20.0, nframes=2)
    outs = []
    for _ in range(2):
        s, plan, c, types, scale = _gpu(np.array(coords), dims, typ, rad, 0.4, True, np.eye(3))
        plan.accumulate_frames(c, types, scale, [float(x) for x in s.L], s.arith, s.wrap)
        plan.check_errors()
        outs.append(plan.sf().cpu().numpy())
    assert np.array_equal(outs[0], outs[1])


def test_water_64_end_to_end(rad, gpu, tmp_path):
    """The CPU-runnable water workload (64^3, 2,700 atoms), 2 frames, through
    .gro/.trr files and load_traj.process_gro_mdtraj as the driver does
    (main_gromacs.py:197-212), vs the oracle."""
    import dens
    import load_traj
    from mdsf import synth
    gro, trr = synth.water_files(str(tmp_path), seed=1001, nwater=900, box=30.0, nframes=2)
    load_traj.process_gro_mdtraj(gro, trr, str(tmp_path / "traj"))
    traj = np.load(str(tmp_path / "traj.npz"))
    coords, dims, typ = traj["coords"], traj["dims"], traj["typ"]
    ref_coords, ref_dims, _ = synth.water(seed=1001, nwater=900, box=30.0, nframes=2)
    assert coords.dtype == np.float32 and np.array_equal(coords, ref_coords)
    assert np.array_equal(dims, ref_dims)
    typ = [str(x) for x in typ]
    ro = np.array(coords)
    want = O.compute_sf(ro, np.array(dims), typ, None, rad, np.eye(3), 0.46875)
    r = np.array(coords)
    dens.PRINT_DETAILS = False
    dens.compute_sf(r, dims, typ, str(tmp_path / "w"), rad, np.eye(3), 0.46875)
    got = dict(np.load(str(tmp_path / "w.npz")))
    assert list(got["N"]) == [64, 64, 64]
    assert np.array_equal(r, ro)
    assert_sq_close(got["sf"], want["sf"])


def test_atom_far_outside_deposits_nothing(rad, gpu, tmp_path):
    """An atom 3.5 boxes out stays 2.5 boxes out after the single wrap
    (dens.py:206-212); its window starts past the padded grid, so the
    reference's slices (dens.py:292-308) come out empty: no error, no
    density.  S(q) equals the oracle's, and that of the first atom alone."""
    import dens
    r = np.array([[[1.0, 1.0, 1.0], [35.0, 2.0, 2.0]]])
    dims = np.array([[10.0, 10.0, 10.0]])
    want = O.compute_sf(np.array(r), dims, ["O", "H"], None, rad, np.eye(3), 0.5)["sf"]
    alone = O.compute_sf(np.array(r[:, :1]), dims, ["O"], None, rad, np.eye(3), 0.5)["sf"]
    dens.PRINT_DETAILS = False
    dens.compute_sf(r, dims, ["O", "H"], str(tmp_path / "e"), rad, np.eye(3), 0.5)
    got = np.load(str(tmp_path / "e.npz"))["sf"]
    assert r[0, 1, 0] == 25.0
    assert_sq_close(got, want)
    assert np.array_equal(want, alone)


def test_type_index_error_raises(rad, gpu):
    import torch
    s, plan, coords, types, scale = _gpu(np.array([[[1.0, 1.0, 1.0]]]), np.array([[8.0] * 3]),
                                         ["O"], rad, 0.5, True, np.eye(3))
    bad = torch.tensor([5], dtype=torch.int32, device=plan.device)
    plan.accumulate_frames(coords, bad, scale, [8.0] * 3, s.arith, s.wrap)
    with pytest.raises(KeyError):
        plan.check_errors()


def test_unknown_name_raises_keyerror(rad, gpu, tmp_path):
    import dens
    dens.PRINT_DETAILS = False
    with pytest.raises(KeyError):
        dens.compute_sf(np.ones((1, 2, 3)), np.ones((1, 3)) * 9, ["OW", "HW1"],
                        str(tmp_path / "k"), rad, np.eye(3), 0.5)


def test_get_dplot_on_gpu_exact(gpu):
    import dens
    rng = np.random.default_rng(4)
    sf = rng.random((12, 10, 9))
    assert np.array_equal(dens.get_dplot(sf), O.get_dplot(sf))


def test_ewald_slices_match_scipy(rad, gpu):
    import torch
    from scipy.interpolate import RegularGridInterpolator
    from mdsf import engine
    rng = np.random.default_rng(12)
    n = (16, 14, 18)
    sf = rng.random((n[0], n[1], n[2] // 2 + 1)) + 1.0
    plan = engine.Plan(n, 0, [1.0] * 3, np.eye(3), [], [], [], [])
    L = [9.0, 8.0, 10.0]
    _, _, kp = plan.outputs(torch.from_numpy(sf).to(plan.device), L, False)
    wl = 1.542
    xy, xz, yz = engine.ewald_slices(kp, wl)
    D = kp.cpu().numpy()
    X, Y, Z = D[:, 0, 0, 0], D[0, :, 0, 1], D[0, 0, :, 2]
    ES = RegularGridInterpolator((X, Y, Z), D[..., 3], bounds_error=False)
    K = 2.0 * math.pi / wl
    th = np.arctan(np.sqrt(X[:, None] ** 2.0 + Y[None, :] ** 2.0) / K)
    pts = np.stack([X[:, None] * np.cos(th), Y[None, :] * np.cos(th), K * (1.0 - np.cos(th))], -1)
    want_xy = ES(pts.reshape(-1, 3)).reshape(len(X), len(Y))
    th = np.arctan(np.sqrt(X[:, None] ** 2.0 + Z[None, :] ** 2.0) / K)
    pts = np.stack([X[:, None] * np.cos(th), np.broadcast_to(K * (1.0 - np.cos(th)), th.shape),
                    Z[None, :] * np.cos(th)], -1)
    want_xz = ES(pts.reshape(-1, 3)).reshape(len(X), len(Z))
    th = np.arctan(np.sqrt(Y[:, None] ** 2.0 + Z[None, :] ** 2.0) / K)
    pts = np.stack([np.broadcast_to(K * (1.0 - np.cos(th)), th.shape), Y[:, None] * np.cos(th),
                    Z[None, :] * np.cos(th)], -1)
    want_yz = ES(pts.reshape(-1, 3)).reshape(len(Y), len(Z))
    for got, want in ((xy, want_xy), (xz, want_xz), (yz, want_yz)):
        g = got.cpu().numpy()
        assert np.array_equal(np.isnan(g), np.isnan(want))
        ok = ~np.isnan(want)
        np.testing.assert_allclose(g[ok], want[ok], rtol=1e-12, atol=0)


def test_gentraj_features_end_to_end(rad, gpu, tmp_path):
    """A seeded pore + layer system in a 60-degree monoclinic cell (the
    generator after gentraj.py), written as .gro/.trr, loaded by
    load_traj.process_gro_mdtraj, moved to the monoclinic frame as
    main_gromacs.py:204-207 does, then compute_sf vs the oracle."""
    import dens
    import load_traj
    from mdsf import gentraj
    pts, cell = gentraj.generate(seed=11, npoints=400, nframes=2, box=(2.5, 2.5, 2.5),
                                 angles=(90.0, 90.0, 60.0), pores=True, radius=0.4, rows=2,
                                 layers=True, dbwl=0.5, layer_width=0.1)
    prefix = str(tmp_path / "g")
    gentraj.write(prefix, pts, cell, name="C")
    load_traj.process_gro_mdtraj(prefix + ".gro", prefix + ".trr", str(tmp_path / "traj"))
    traj = np.load(str(tmp_path / "traj.npz"))
    coords, dims, typ = np.array(traj["coords"]), traj["dims"], [str(x) for x in traj["typ"]]
    theta = math.radians(60.0)
    coords[..., 1] = coords[..., 1] / np.sin(theta)
    coords[..., 0] = coords[..., 0] - coords[..., 1] * np.cos(theta)
    ucell = np.array([[1, 0, 0], [np.cos(theta), np.sin(theta), 0], [0, 0, 1]])
    ro = np.array(coords)
    want = O.compute_sf(ro, np.array(dims), typ, None, rad, ucell, 0.5)
    r = np.array(coords)
    dens.PRINT_DETAILS = False
    dens.compute_sf(r, dims, typ, str(tmp_path / "s"), rad, ucell, 0.5)
    got = dict(np.load(str(tmp_path / "s.npz")))
    assert np.array_equal(r, ro)
    assert np.array_equal(got["N"], want["N"])
    assert_sq_close(got["sf"], want["sf"])


@pytest.mark.parametrize("precision,water", [("auto", "drop"), ("fp32", "drop"), ("auto", "map")])
def test_reference_hii_frame_end_to_end(precision, water, rad, gpu, tmp_path):
    """The reference's own test fixture test/test_system.gro (one frame of an
    inverted-hexagonal phase: 55,680 atoms in a 120-degree cell), read through
    the drop-in load_traj as main_gromacs.py:197 does.  The 480 water atoms
    (HW1 / HW2 / OW, which radii.txt lacks, so the reference raises KeyError on
    them) are left out ("drop") or renamed O / H ("map", all 55,680 atoms).  The rest is moved to the monoclinic frame as
    main_gromacs.py:204-207 does, then compute_sf vs the oracle at
    Sres = 1 Angstrom (88 x 88 x 84, Nborder 17).  The frame has 138 atom
    names (4 parameter sets) and atoms exactly on the cell faces (home cell N).

    Both "auto" and "fp32" (FP32 storage) meet 1e-5 above 1e-6 max(S); the
    error of scipy's complex64 rfftn of the oracle's own density is printed
    for scale."""
    import gzip
    import shutil
    import scipy.fft
    import dens
    import load_traj
    from conftest import GOLDEN
    gro = str(tmp_path / "test_system.gro")
    with gzip.open(os.path.join(GOLDEN, "test_system.gro.gz"), "rb") as fi, open(gro, "wb") as fo:
        shutil.copyfileobj(fi, fo)
    load_traj.process_gro_mdtraj(gro, gro, str(tmp_path / "traj"))
    traj = np.load(str(tmp_path / "traj.npz"))
    if water == "map":
        names = {"OW": "O", "HW1": "H", "HW2": "H"}
        typ = [names.get(str(n), str(n)) for n in traj["typ"]]
        coords = np.array(traj["coords"])
        assert len(typ) == 55680 and all(n in rad for n in typ)
    else:
        keep = np.array([str(n) in rad for n in traj["typ"]])
        assert keep.sum() == 55200
        coords = np.array(traj["coords"][:, keep])
        typ = [str(n) for n in traj["typ"][keep]]
    dims = traj["dims"]
    theta = math.radians(120.0)
    coords[..., 1] = coords[..., 1] / np.sin(theta)
    coords[..., 0] = coords[..., 0] - coords[..., 1] * np.cos(theta)
    ucell = np.array([[1, 0, 0], [np.cos(theta), np.sin(theta), 0], [0, 0, 1]])
    ro = np.array(coords)
    want = O.compute_sf(ro, np.array(dims), typ, None, rad, ucell, 1.0,
                        better=dens.USE_BETTER_RESOLUTION, keep=True)
    saved = (dens.PRINT_DETAILS, dens.FFT_PRECISION)
    dens.PRINT_DETAILS, dens.FFT_PRECISION = False, precision
    try:
        r = np.array(coords)
        dens.compute_sf(r, dims, typ, str(tmp_path / "s"), rad, ucell, 1.0)
    finally:
        dens.PRINT_DETAILS, dens.FFT_PRECISION = saved
    got = dict(np.load(str(tmp_path / "s.npz")))
    assert list(got["N"]) == [88, 88, 84] and want["Nb"] == 17
    assert np.array_equal(r, ro)
    err = assert_sq_close(got["sf"], want["sf"])
    if precision == "fp32":
        # for scale: numpy's single-precision transform of the same density
        f32 = scipy.fft.rfftn(want["d1"][0].astype(np.float32))
        assert f32.dtype == np.complex64
        s32 = np.abs(f32.astype(np.complex128)) ** 2
        mask = want["sf"] > SQ_FLOOR * want["sf"].max()
        floor32 = float(np.max(np.abs(s32[mask] - want["sf"][mask]) / want["sf"][mask]))
        print(f"scipy float32 rfftn of the same density: max rel err {floor32:.2e}")
    print(f"test_system.gro ({precision}, water {water}): S(q) max rel err {err.max():.2e}, "
          f"median {np.median(err):.2e}")
