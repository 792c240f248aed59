"""Drop-in ``dens`` module: the S(q) path of joeyelk/MD-Structure-Factor on an MI355X.

Same names, signatures, module globals and ``.npz`` outputs as the reference
``dens.py``; callers (``main_gromacs.py:155-158,183-186,209-212``) keep
``import dens`` with this directory on ``sys.path``.  The frame loop
(rescale, wrap, Gaussian deposition, border fold, R2C FFT, |F|^2 frame sum)
and the output grids run as HIP kernels in ``libmdsf.so``; this module does
the reference's host-side set-up arithmetic and the ``.npz`` write.

Precision.  "fp32" plans store the density and the spectrum between FFT
passes in FP32, evaluate every butterfly, twiddle and the real-to-complex
split in FP64, keep 4-bit residuals of the y pass's FP32 roundings for the x
pass, and sum |F|^2 over frames in FP64.  "fp64" plans store FP64
throughout (S(q) within about 1e-13 of FP64 numpy).  The default, "auto",
runs frame 0 both ways and keeps FP32 when its S(q) agrees with FP64 within
AUTO_FP32_MAX_ERR (5e-6) on every voxel above 1e-6 of the maximum, which
leaves the 1e-5 bound a factor of two for the frame sum; FP32 storage of a
density concentrated in a few hundred voxels (atoms packed at the cell
corners) is already coarser than that, and such inputs run in FP64.  Home
cells, grid sizes and k-grids are exact.
"""
import concurrent.futures
import math

import numpy as np

from mdsf import engine, npzio

USE_BETTER_RESOLUTION = True          # dens.py:10
PRINT_DETAILS = True                  # dens.py:11
RANDOM_NOISE = 0                      # dens.py:12
Nspatialgrid = np.asarray([0.0, 0.0, 0.0])   # dens.py:15
theta = np.pi / 3                     # dens.py:16 (set by callers, not read)
PRECISION = 1.0E-24                   # dens.py:18
dtyp = np.float64                     # dens.py:20 (FP64 S(q) accumulation)

# Storage precision of the GPU path: "fp32", "fp64" or "auto" (frame 0 in
# both; FP32 if its S(q) is within AUTO_FP32_MAX_ERR of FP64 on the voxels
# above 1e-6 of the maximum, else FP64; RANDOM_NOISE fields run in FP32).
# S(q) is always summed in FP64.
FFT_PRECISION = "auto"
AUTO_FP32_MAX_ERR = 5e-6
# Outcome of the last compute_sf call's precision choice: {"fp64": bool,
# "frame0_fp32_err": max relative S(q) error of frame 0 in FP32, or None}.
LAST_PRECISION = {}

# Threads deflating the .npz members (None: up to 16, one per core); the
# archive is what np.savez_compressed writes, compressed in parallel.
NPZ_THREADS = None
# zlib level of those members (6: what np.savez_compressed uses; 1 writes
# several times faster and somewhat larger files that np.load reads the same)
NPZ_LEVEL = 6

# Frames per checkpoint block (0: off).  With n > 0, compute_sf runs the
# frames in blocks of n, adds each block's FP64 sum to a running total in
# frame order and, after each block, saves (frames done, the FP64 total) to
# out_filename + ".ckpt.npz"; a later call on the same inputs (coordinates,
# boxes, types, Sres, cell, grid) resumes after the last block saved, and
# the result is bitwise that of an uninterrupted call with the same n.  It
# can differ in the last bits from a call without checkpoints (blocks summed
# apart, then added).  The file is removed once the output is written.  The
# reference saves only at the end (dens.py:346); RANDOM_NOISE runs ignore it.
CHECKPOINT_FRAMES = 0

# Wall-clock seconds of the phases of the last compute_sf call: "setup",
# "upload", "plan" (plan creation, with the "auto" precision check on frame
# 0), "frames", "outputs" (S(q) and the output grids on the GPU), "write"
# (device-to-host copies overlapped with the deflate) (informational).
LAST_TIMINGS = {}


def load_radii(filename):
    """``{name: (Nel, sigma_angstrom)}`` from ``Nel name radius_pm`` lines (dens.py:23-35)."""
    out = {}
    with open(filename) as f:
        for line in f.readlines():
            cols = line.strip().split()
            out[cols[1]] = (float(cols[0]), float(cols[2]) / 100.0)
    return out


def get_borders(ad, dr, keys):
    """Window half-widths in grid steps per type (dens.py:38-43)."""
    return engine.borders(ad, dr, keys)


def rescale(coords, dims):
    """Rescale ``coords`` (frames, atoms, 3) in place so that every frame has the
    mean box; returns ``(coords, avgdims)`` (dens.py:46-62).  The products run
    on the GPU in numpy's arithmetic (float32 only when coords and dims are)."""
    import torch
    if coords.dtype not in (np.float32, np.float64):
        raise TypeError("coordinates must be float32 or float64")
    avgdims = np.average(dims, axis=0)
    a = avgdims / dims
    arith = engine._lib.F32 if (coords.dtype == np.float32 and a.dtype == np.float32) \
        else engine._lib.F64
    dev = engine._device(None)
    c = torch.from_numpy(np.ascontiguousarray(coords)).to(dev)
    sc = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)
    engine.rescale_device(c, sc, arith)
    coords[...] = c.cpu().numpy()
    return coords, avgdims


def _fold_tables(des, ori, ndim):
    """(N, Nborder) per axis from the reference's fold tables, which compute_sf
    builds as des = [0, Nb, N - Nb, N], ori = [0, Nb, Nb + N, 2 Nb + N]
    (dens.py:258-263); other tables are rejected."""
    n, nb = [], []
    for d in range(ndim):
        de = [int(v) for v in np.asarray(des[d]).reshape(-1)]
        oi = [int(v) for v in np.asarray(ori[d]).reshape(-1)]
        ok = len(de) == 4 and len(oi) == 4
        b, nn = (oi[1], oi[2] - oi[1]) if ok else (0, 0)
        if not ok or de != [0, b, nn - b, nn] or oi != [0, b, b + nn, 2 * b + nn]:
            raise ValueError("only the fold tables compute_sf builds are supported: "
                             "des = [0, Nb, N - Nb, N], ori = [0, Nb, Nb + N, 2 Nb + N]")
        n.append(nn)
        nb.append(b)
    return n, nb


def remap_grid_tcl(d0, des, ori):
    """Fold the padded grid ``d0`` onto the periodic cell with per-axis tables
    (dens.py:86-108, corner z slot included), on the GPU; returns FP64."""
    import torch
    n, nb = _fold_tables(des, ori, 3)
    dev = engine._device(None)
    d = torch.from_numpy(np.ascontiguousarray(d0, dtype=np.float64)).to(dev)
    return engine.fold_padded(d, n, nb).cpu().numpy()


def remap_grid(d0, des, ori):
    """The cubic form of remap_grid_tcl: one table for every axis (dens.py:65-83)."""
    return remap_grid_tcl(d0, [des] * 3, [ori] * 3)


def get_dplot(dmag):
    """Half spectrum -> centred plottable grid (dens.py:142-163), on the GPU."""
    import torch
    dmag = np.asarray(dmag, dtype=np.float64)
    nx, ny, h = dmag.shape
    n = [nx, ny, 2 * (h - 1)]
    plan = _output_plan(n)
    sf = torch.from_numpy(np.ascontiguousarray(dmag)).to(plan.device)
    sfplt, _, _ = plan.outputs(sf, [1.0, 1.0, 1.0], False)
    return sfplt.cpu().numpy()


def _frame0_plans(s, coords, types, scale, box):
    """FP32 and FP64 plans that have both run frame 0 (on a copy of it), and
    the max relative error of the FP32 S(q) against the FP64 one on the
    voxels above 1e-6 of the maximum (0 without such voxels)."""
    import torch
    plans, out = [], []
    try:
        for fp64 in (False, True):
            plans.append(engine.Plan.from_setup(s, device=coords.device, fft64=fp64))
            plans[-1].accumulate_frames(coords[:1].clone(), types, scale[:1], box, s.arith,
                                        s.wrap)
            plans[-1].check_errors()
            out.append(plans[-1].sf())
    except BaseException:
        for p in plans:
            p.close()
        raise
    s32, s64 = out
    mask = s64 > 1e-6 * s64.max()
    err = float(torch.max(torch.abs(s32[mask] - s64[mask]) / s64[mask])) if bool(mask.any()) else 0.0
    return plans[0], plans[1], err


def _make_plan(s, coords, types, scale, box):
    """The run's plan in the precision FFT_PRECISION selects (module docstring)."""
    LAST_PRECISION.clear()
    if FFT_PRECISION not in ("auto", "fp32", "fp64"):
        raise ValueError("FFT_PRECISION must be 'auto', 'fp32' or 'fp64'")
    if FFT_PRECISION == "auto" and RANDOM_NOISE == 0 and s.nframes > 0 and s.natoms > 0:
        p32, p64, err = _frame0_plans(s, coords, types, scale, box)
        fp64 = err > AUTO_FP32_MAX_ERR
        keep, drop = (p64, p32) if fp64 else (p32, p64)
        drop.close()
        keep.reset()
        LAST_PRECISION.update(fp64=fp64, frame0_fp32_err=err)
        return keep
    fp64 = FFT_PRECISION == "fp64"
    LAST_PRECISION.update(fp64=fp64, frame0_fp32_err=None)
    return engine.Plan.from_setup(s, device=coords.device, fft64=fp64)


def _output_plan(n):
    return engine.Plan(n, 0, [1.0, 1.0, 1.0], np.eye(3), [], [], [], [])


def _print_details(L, Sres, N):
    print("=" * 50)
    print("=" * 50)
    print("Unit cell has dimensions of %s" % L)
    print("requested resolution: %s Angstroms" % Sres)
    print("Using a spatial grid of: %s" % N)
    print("actual resolution: %s Angstroms" % (L / N))
    print("This corresponds to maximum q vector of %s inverse Angstroms" % (2 * math.pi * N / L))
    print("=" * 50)
    print("=" * 50)


def _checkpoint_key(r, L, typ, Sres, ucell, N):
    """Fingerprint of a compute_sf call's inputs (before the in-place rescale)."""
    import hashlib
    h = hashlib.blake2b(digest_size=16)
    for a in (np.ascontiguousarray(r), np.asarray(L, dtype=np.float64),
              np.asarray(ucell, dtype=np.float64), np.asarray(N, dtype=np.int64)):
        h.update(str(a.dtype).encode() + str(a.shape).encode())
        h.update(a.tobytes())
    h.update(("\0".join(str(t) for t in typ) + "|" + repr(float(Sres))).encode())
    return h.hexdigest()


def _checkpoint_path(out_filename):
    return str(out_filename) + ".ckpt.npz"


def _checkpoint_save(out_filename, key, frames_done, total):
    """Writes the checkpoint through a temporary file and an atomic rename."""
    import os
    path = _checkpoint_path(out_filename)
    tmp = path + ".tmp.npz"
    np.savez(tmp, key=np.frombuffer(key.encode(), dtype=np.uint8),
             frames_done=np.int64(frames_done), sf=total)
    os.replace(tmp, path)


def _checkpoint_load(out_filename, key, shape):
    """(frames done, FP64 partial S(q)) of a checkpoint of these inputs, or
    (0, None) when there is none or it belongs to other inputs."""
    import os
    path = _checkpoint_path(out_filename)
    if not os.path.exists(path):
        return 0, None
    with np.load(path, allow_pickle=False) as z:
        if bytes(z["key"]).decode() != key or tuple(z["sf"].shape) != tuple(shape):
            return 0, None
        return int(z["frames_done"]), np.array(z["sf"], dtype=np.float64)


def compute_sf(r, L, typ, out_filename, rad, ucell=None, Sres=1.0):
    """Compute the 3-D structure factor and save it to ``out_filename + '.npz'``.

    Same contract as the reference (dens.py:166-346): ``r`` (frames, atoms, 3)
    is rescaled and wrapped in place, ``L`` holds per-frame box lengths, ``typ``
    names each atom's entry of ``rad``; the ``.npz`` holds ``sf`` (the SUM over
    frames of |rfftn(rho)|^2), ``sfplt``, ``L``, ``N``, ``kgrid``, ``kgridplt``.
    ``ucell`` defaults to the identity and ``Sres`` to 1 Angstrom.
    """
    import time

    import torch
    t0 = time.perf_counter()
    if ucell is None:
        ucell = np.eye(3)
    s = engine.Setup(r, L, typ, rad, Sres, USE_BETTER_RESOLUTION, ucell)
    ckpt = CHECKPOINT_FRAMES > 0 and RANDOM_NOISE <= 0 and s.nframes > 0
    key = _checkpoint_key(r, L, typ, Sres, ucell, s.N) if ckpt else None
    if PRINT_DETAILS:
        _print_details(s.L, Sres, s.N)
    dev = engine._device(None)
    t1 = time.perf_counter()
    coords = torch.from_numpy(np.ascontiguousarray(r)).to(dev)
    scale = torch.from_numpy(np.ascontiguousarray(s.scale, dtype=np.float64)).to(dev)
    types = torch.from_numpy(s.type_index.astype(np.int32).reshape(-1)).to(dev)
    box = [float(x) for x in s.L]
    torch.cuda.synchronize(dev)
    t2 = time.perf_counter()
    plan = _make_plan(s, coords, types, scale, box)
    torch.cuda.synchronize(dev)
    t2b = time.perf_counter()
    if RANDOM_NOISE > 0:
        plan.rescale_wrap(coords, scale, box, s.arith, s.wrap)
        for _ in range(s.nframes):
            d1 = np.random.rand(int(s.N[0]), int(s.N[1]), int(s.N[2]))   # dens.py:280
            plan.accumulate_density(torch.from_numpy(d1))   # cast to the plan precision
    elif ckpt:
        done, part = _checkpoint_load(out_filename, key, plan.half_shape)
        total = torch.from_numpy(part).to(dev) if part is not None else None
        if done > 0:   # frames already summed: only their in-place rescale
            plan.rescale_wrap(coords[:done], scale[:done], box, s.arith, s.wrap)
        k = done
        while k < s.nframes:
            c = min(int(CHECKPOINT_FRAMES), s.nframes - k)
            plan.reset()
            plan.accumulate_frames(coords[k:k + c], types, scale[k:k + c], box, s.arith, s.wrap)
            plan.check_errors()
            block = plan.sf()
            total = block if total is None else total + block
            k += c
            _checkpoint_save(out_filename, key, k, total.cpu().numpy())
    else:
        plan.accumulate_frames(coords, types, scale, box, s.arith, s.wrap)
    plan.check_errors()
    t3 = time.perf_counter()
    r[...] = coords.cpu().numpy()
    sf_dev = total if ckpt else plan.sf()
    sfplt, kgrid, kgridplt = plan.outputs(sf_dev, box, s.box_f32)
    stream = torch.cuda.current_stream(dev)

    def fetch(t):
        with torch.cuda.stream(stream):
            return t.cpu().numpy()

    # the device-to-host copies run on a helper thread in the order the
    # archive is written, while the writer deflates the members already copied
    with concurrent.futures.ThreadPoolExecutor(1) as pool:
        out = dict(sf=pool.submit(fetch, sf_dev), sfplt=pool.submit(fetch, sfplt), L=s.L, N=s.N,
                   kgrid=pool.submit(fetch, kgrid), kgridplt=pool.submit(fetch, kgridplt))
        t4 = time.perf_counter()
        npzio.savez_compressed(out_filename, threads=NPZ_THREADS, level=NPZ_LEVEL, **out)
    if ckpt:
        import os
        os.remove(_checkpoint_path(out_filename))
    plan.close()
    del sfplt, kgrid, kgridplt, sf_dev, out
    t5 = time.perf_counter()
    LAST_TIMINGS.clear()
    LAST_TIMINGS.update(setup=t1 - t0, upload=t2 - t1, plan=t2b - t2, frames=t3 - t2b,
                        outputs=t4 - t3,
                        write=t5 - t4)
    if ckpt:
        LAST_TIMINGS["resumed_from"] = done
