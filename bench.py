"""Benchmark: frames/s of the S(q) hot path at 256^3 with 50,000 atoms on one MI355X.

Workload (synthetic, seeded): inverted-hexagonal phase in a 104 x 90 x 80 A
box, 4 pore columns lined with C/O/Na headgroups, C/H tails elsewhere
(50,000 atoms), frames = base + N(0, 0.2 A); grid 256^3 (Sres = L/256 per
axis), FP32 deposition and FFT, FP64 frame sum of |F|^2.

One step = one frame through the whole path (rescale/wrap, binning, factor
tables, deposition + fold, z/y/x FFT passes, |F|^2 accumulation), with the
trajectory already resident in HBM.  The timed region is the library's own
pipelined frame loop without stage events; a second pass runs 30 frames one
at a time (after 100 untimed ones) with the library's stage timing (the
deposition kernel and each FFT pass kernel timed by events recorded with
their own dispatch) for the roofline and the per-pass FFT figures, whatever
the step count.  The CPU
baseline runs one whole frame through the FP64 numpy oracle (the reference's
per-atom op sequence) on one host core, and, as a stronger baseline, one frame
per host core at once in separate processes (`cpu_baseline_all_cores`, up to
16); the precision leg compares the FP32 and FP64 plans with that oracle frame.

    python bench.py [--gpus 1] [--steps 200] [--warmup 10] [--no-cpu]
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "md-structure-factor_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
METRIC = "frames/s for S(q) at 256^3, 50k atoms"
# HBM bytes per deposition launch on this workload, from two rocprofv3 --pmc
# runs (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch) scaled by the same
# counters on a 256 MiB device copy (FETCH x 2.000, WRITE x 1.000)
DEPOSIT_TRAFFIC = 143718400


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def fft_bytes(n, fp64=False):
    """Algorithmic HBM bytes of the three passes: density R (4 / 8 bytes), spectrum
    complex R between passes, FP64 complex Nyquist plane (kz = Nz/2), a residual
    byte per y-pass output value (FP32 plans), FP64 S(q) read and written by x."""
    nx, ny, nz = n
    m = nz // 2
    c = 16 if fp64 else 8        # complex value between passes
    res = 0 if fp64 else 1       # y pass residual byte
    plane = nx * ny              # Nyquist plane points
    vol = nx * ny * m            # half-spectrum points below kz = Nz/2
    return {
        # read the density rows, write spec[x][kz][y] and the Nyquist plane
        "z": (c // 2) * nx * ny * nz + c * vol + 16 * plane,
        # read spec, write rho[kz][ky][x] + residuals; the Nyquist plane in place
        "y": 2 * c * vol + res * vol + 32 * plane,
        # read rho + residuals, read and write sf; the Nyquist plane into sf_nyq
        "x": c * vol + res * vol + 16 * vol + 16 * plane + 16 * plane,
        # x with frames in pairs (the frame loop's form), per frame: the frame's
        # rho + residuals + Nyquist plane, half of one read and write of sf / sf_nyq
        "x_pair": c * vol + res * vol + 8 * vol + 16 * plane + 8 * plane,
    }


def voxel_updates(setup):
    """Voxel updates per frame: sum over atoms of the window volume prod_d 2A_d (dens.py:292-308)."""
    per_type = np.prod(2 * setup.half, axis=1)
    return int(np.sum(per_type * setup.counts))


def _cpu_name():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(coords0, dims, typ, rad, sres):
    """One whole frame of the workload through the FP64 numpy oracle (the
    reference's per-atom op sequence, fold, rfftn and |F|^2, dens.py:283-318)
    on one host core: frames/s = 1 / its wall time.  Returns the baseline
    record and the oracle's S(q) of that frame (the precision leg's reference)."""
    from oracle import dens_oracle as O
    r = np.array(coords0[None], copy=True)
    o = O.setup(r, np.array(dims[:1]), typ, rad, sres, better=False)
    n, nb = o["N"], o["Nb"]
    padded = tuple(int(n[i]) + 2 * nb for i in range(3))
    t0 = time.perf_counter()
    d0, _ = O.deposit_frame(r[0], list(typ), rad, o["bdict"], o["dr"], nb, np.eye(3), padded)
    t1 = time.perf_counter()
    d1 = O.fold(d0, *O.fold_tables(n, nb))
    del d0
    t2 = time.perf_counter()
    sf = O.frame_power(d1)
    t3 = time.perf_counter()
    frame_s = t3 - t0
    rec = {
        "value": 1.0 / frame_s,
        "unit": "frames/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"frame 0 of the workload, whole ({len(typ)} atoms, {list(int(x) for x in n)} "
                   f"grid), through the FP64 numpy oracle: deposition {t1 - t0:.1f} s, fold "
                   f"{t2 - t1:.2f} s, rfftn + |F|^2 {t3 - t2:.2f} s"),
        "host": f"{_cpu_name()}; nproc={os.cpu_count()}; {platform.machine()}",
    }
    return rec, sf


def _oracle_frame_seconds(job):
    """Worker of the all-cores baseline: one frame through the FP64 numpy
    oracle (deposition, fold, rfftn + |F|^2) in this process, on one thread;
    returns its wall time."""
    r, dims, typ, rad, sres = job
    from threadpoolctl import threadpool_limits
    from oracle import dens_oracle as O
    with threadpool_limits(limits=1):
        o = O.setup(r, dims, typ, rad, sres, better=False)
        n, nb = o["N"], o["Nb"]
        padded = tuple(int(n[i]) + 2 * nb for i in range(3))
        t0 = time.perf_counter()
        d0, _ = O.deposit_frame(r[0], list(typ), rad, o["bdict"], o["dr"], nb, np.eye(3), padded)
        d1 = O.fold(d0, *O.fold_tables(n, nb))
        del d0
        O.frame_power(d1)
        return time.perf_counter() - t0


def cpu_baseline_all_cores(coords, dims, typ, rad, sres, max_procs=16):
    """The stronger CPU baseline: P frames of the workload through the FP64
    numpy oracle at once, one process and one thread per frame, P = the
    host cores this process may use (at most max_procs, at most the frames);
    frames/s = P / the wall time of the P frames (worker start-up excluded)."""
    import multiprocessing as mp
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count() or 1
    nproc = max(1, min(max_procs, ncpu, coords.shape[0]))
    jobs = [(np.array(coords[i][None], copy=True), np.array(dims[i:i + 1]), typ, rad, sres)
            for i in range(nproc)]
    with mp.get_context("spawn").Pool(nproc) as pool:
        # start the workers (imports) before the clock; both waits are bounded
        pool.map_async(abs, range(nproc)).get(timeout=120)
        t0 = time.perf_counter()
        per = pool.map_async(_oracle_frame_seconds, jobs, chunksize=1).get(timeout=600)
        wall = time.perf_counter() - t0
    return {
        "value": nproc / wall,
        "unit": "frames/s",
        "cores": nproc,
        "kind": "port",
        "sample": (f"frames 0..{nproc - 1} of the workload, one per process, through the FP64 "
                   f"numpy oracle: {wall:.1f} s wall, {min(per):.1f}-{max(per):.1f} s per frame"),
        "host": f"{_cpu_name()}; usable cores={ncpu}; {platform.machine()}",
    }


def copy_calibration(nbytes=256 << 20):
    import torch
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": nbytes, "reps": 3}), flush=True)


def isolated_stages(plan, d_coords, d_types, d_scale, box, s, nframes, stream, warm=100):
    """Each stage alone: one-frame calls of the library's frame loop with stage
    timing on, queued back to back (each call joins back into the stream the
    next one starts from, so every stage of a frame runs by itself: the
    prepare stage, then the deposition, then the FFT passes, and the next
    frame's prepare after them).  Prepare: events on its stream around the
    whole stage.  The deposition's tiled kernel and each FFT pass's kernel
    record their events with their own dispatch (begin and end of the kernel:
    the durations a rocprofv3 kernel trace reports for them).  `warm`
    untimed one-frame calls first, so that a short run is timed at the same
    steady clocks as a long one; the frames are taken in turn from the ones
    on the device."""
    import torch
    nf = int(d_coords.shape[0])
    for t in range(warm):
        i = t % nf
        plan.accumulate_frames(d_coords[i:i + 1], d_types, d_scale[i:i + 1], box, s.arith, s.wrap)
    torch.cuda.synchronize()
    plan.set_stage_timing(1)
    torch.cuda.synchronize()
    for t in range(nframes):
        i = t % nf
        plan.accumulate_frames(d_coords[i:i + 1], d_types, d_scale[i:i + 1], box, s.arith,
                               s.wrap)
    st = plan.stage_times()
    plan.set_stage_timing(0)
    torch.cuda.synchronize()
    plan.check_errors()
    return {"prepare": st["prepare"], "deposit_kernel": st["deposit"], "fft_z": st["fft_z"],
            "fft_y": st["fft_y"], "fft_x_acc": st["fft_x_acc"]}


def paired_x(plan, d_coords, d_types, d_scale, box, s, npairs):
    """The x pass as the frame loop runs it, two frames per launch: two-frame
    calls queued back to back with stage timing.  The pair's launch is timed as
    the second frame's x pass (with the kernel's own dispatch) and the first
    frame's as zero, so the average is the per-frame cost.  The launch runs
    alone: the next call starts after the previous one has joined its stream."""
    import torch
    nf = int(d_coords.shape[0])
    plan.set_stage_timing(1)
    torch.cuda.synchronize()
    for k in range(npairs):
        i = (2 * k) % (nf - 1)
        plan.accumulate_frames(d_coords[i:i + 2], d_types, d_scale[i:i + 2], box, s.arith, s.wrap)
    st = plan.stage_times()
    plan.set_stage_timing(0)
    torch.cuda.synchronize()
    plan.check_errors()
    return st["fft_x_acc"]


def adjoint_leg(rad, iters):
    """fwd+bwd frames/s on the adjoint workload: 20k C/O/H atoms, 60 A box, 128^3,
    16 frames, L = sum (S - S_target)^2; backward recomputes every frame."""
    import torch
    from mdsf import synth
    from mdsf.autograd import SFModel
    coords, target, dims, typ = synth.adjoint()
    nframes = coords.shape[0]
    model = SFModel(coords, dims, typ, rad, synth.adjoint_sres())
    dev = model.plan.device
    with torch.no_grad():
        st = model(torch.tensor(target, device=dev))
    x = torch.tensor(coords, device=dev, requires_grad=True)

    def fwd_bwd():
        x.grad = None
        loss = torch.sum((model(x) - st) ** 2)
        loss.backward()

    def fwd():
        with torch.no_grad():
            model(x)

    fwd_bwd()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(iters):
        fwd_bwd()
    torch.cuda.synchronize()
    t_fb = (time.perf_counter() - t0) / iters
    fwd()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(iters):
        fwd()
    torch.cuda.synchronize()
    t_f = (time.perf_counter() - t0) / iters
    return {"workload": "adjoint_128^3_20k_atoms_16_frames", "frames": nframes,
            "fwd_bwd_frames_per_s": nframes / t_fb, "fwd_frames_per_s": nframes / t_f,
            "ms_per_fwd_bwd": t_fb * 1e3, "iters": iters, "dtype": "f32",
            "grad_norm": float(x.grad.norm())}


def sq_rel_err(got, want, floor=1e-6):
    """max / median relative error of S(q) on voxels above floor * max (and the voxel share)."""
    import torch
    m = torch.max(want)
    mask = want > floor * m
    err = torch.abs(got[mask] - want[mask]) / want[mask]
    return {"max": float(err.max()), "median": float(err.median()),
            "voxels": int(mask.sum()), "of": int(want.numel())}


def precision_leg(coords, s, oracle_sf=None, nframes=2):
    """S(q) of the FP32 plan (the measured path) and of the FP64 plan against
    the FP64 numpy oracle on frame 0 (when ``oracle_sf`` is given), and the
    FP32 plan against the FP64 plan over the first ``nframes`` frames."""
    import torch
    from mdsf import engine
    dev = torch.device("cuda", 0)
    types = torch.from_numpy(s.type_index.astype(np.int32).reshape(-1)).to(dev)

    def run(fp64, frames):
        plan = engine.Plan.from_setup(s, fft64=fp64)
        c = torch.from_numpy(np.ascontiguousarray(coords[:frames])).to(dev)
        scale = torch.from_numpy(np.ascontiguousarray(s.scale[:frames], dtype=np.float64)).to(dev)
        plan.accumulate_frames(c, types, scale, [float(x) for x in s.L], s.arith, s.wrap)
        plan.check_errors()
        out = plan.sf()
        plan.close()
        return out

    out = {"floor": "voxels above 1e-6 * max S"}
    if oracle_sf is not None:
        want = torch.from_numpy(oracle_sf).to(dev)
        for fp64 in (False, True):
            out[("fp64" if fp64 else "fp32") + "_vs_oracle"] = sq_rel_err(run(fp64, 1), want)
        out["oracle"] = "FP64 numpy oracle, frame 0 (the cpu_baseline run)"
    res = sq_rel_err(run(False, nframes), run(True, nframes))
    res["frames"] = nframes
    out["fp32_vs_fp64_plan"] = res
    return out


def batch_leg(rad, nframes):
    """Batched form: 32 independent cubic systems (grids {64..192}^3, boxes
    30-80 A, 5k-60k H/C/O/Na atoms, fresh frames) run from one HIP graph with
    autotuned FFT choices; reports system-frames/s of a replay."""
    import torch
    from mdsf import engine, synth
    t0 = time.perf_counter()
    systems = synth.batch(nframes=nframes)
    setups = [engine.Setup(c, d, t, rad, sres, False, np.eye(3)) for c, d, t, sres in systems]
    batch = engine.Batch(setups)
    t_setup = time.perf_counter() - t0
    dev = batch.device
    keep = []
    for i, ((c, d, t, sres), s) in enumerate(zip(systems, setups)):
        coords = torch.from_numpy(np.ascontiguousarray(c)).to(dev)
        types = torch.from_numpy(s.type_index.astype(np.int32).reshape(-1)).to(dev)
        scale = torch.from_numpy(np.ascontiguousarray(s.scale, dtype=np.float64)).to(dev)
        batch.set_frames(i, coords, types, scale)
        keep.append((coords, types, scale))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    batch.run()                      # capture + first replay
    batch.check_errors()
    t_first = time.perf_counter() - t0
    stream = torch.cuda.current_stream(dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    batch.run()
    e1.record(stream)
    torch.cuda.synchronize()
    batch.check_errors()
    ms = e0.elapsed_time(e1)
    frames = sum(int(c.shape[0]) for c, _, _, _ in systems)
    updates = sum(voxel_updates(s) * int(c.shape[0]) for (c, _, _, _), s in zip(systems, setups))
    grids = sorted({int(s.N[0]) for s in setups})
    out = {"workload": f"batch_{len(systems)}_systems_{nframes}_frames",
           "systems": len(systems), "frames_per_system": nframes, "grids": grids,
           "atoms": [int(s.natoms) for s in setups],
           "system_frames_per_s": frames / (ms * 1e-3), "ms_per_replay": ms,
           "voxel_updates_per_s": updates / (ms * 1e-3),
           "first_run_s_incl_capture": t_first, "setup_s_incl_autotune": t_setup,
           "dtype": "f32", "graph": "one hipGraph launch per replay"}
    batch.close()
    return out


def dropin_leg(rad, nframes):
    """Wall time of the drop-in dens.compute_sf on the 256^3 workload (host
    arrays in, .npz out), by phase: set-up, upload, frame loop, outputs, write."""
    import tempfile

    import dens
    from mdsf import synth
    coords, dims, typ = synth.hii(seed=2002, nframes=nframes)
    dens.PRINT_DETAILS = False
    saved = dens.USE_BETTER_RESOLUTION
    dens.USE_BETTER_RESOLUTION = False
    try:
        with tempfile.TemporaryDirectory() as d:
            t0 = time.perf_counter()
            dens.compute_sf(coords, dims, typ, os.path.join(d, "sf"), rad, np.eye(3),
                            synth.hii_sres())
            total = time.perf_counter() - t0
            size = os.path.getsize(os.path.join(d, "sf.npz"))
    finally:
        dens.USE_BETTER_RESOLUTION = saved
    out = {"workload": f"compute_sf hii_phase_256^3_50k_atoms {nframes} frames",
           "total_s": total, "npz_MB": size / 1e6,
           "frames_per_s_in_frame_loop": nframes / dens.LAST_TIMINGS["frames"]}
    out.update({k + "_s": v for k, v in dens.LAST_TIMINGS.items()})
    return out


def cylindrical_leg(plan, box, ucell):
    """The r-theta-z average of PLOT_RAD_NEW (plot2d.py:559-655) on the 256^3
    kgridplt of the timed frames (GPU outputs, then engine.rz_average)."""
    import torch
    from mdsf import engine
    _, _, kp = plan.outputs(plan.sf(), box, False)
    out = engine.rz_average(kp, ucell)          # first call: bins, library load
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = engine.rz_average(kp, ucell)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    samples = int(np.sum([max(int(engine.THETA_BINS_PER_INV_ANG * 2.0 * np.pi * r), 1)
                          for r in out["rarr"]])) * len(out["zar"])
    return {"workload": "rz_average of the hii 256^3 kgridplt (400 r bins x 255 z)",
            "ms": dt * 1e3, "interpolations": samples,
            "interpolations_per_s": samples / dt, "ucell": "orthogonal (the workload's cell)"}


def mono_leg(coords, dims, typ, rad, sres, steps, warmup):
    """The same frames in the reference driver's default cell (main_gromacs.py
    -ct 120, :61,79-80): coordinates moved to the monoclinic frame as
    main_gromacs.py:204-207 does (float32), ucell = [[1,0,0],[cos,sin,0],[0,0,1]]
    (the xy-coupled deposition); frames/s of the pipelined frame loop."""
    import torch
    from mdsf import engine
    th = np.radians(120.0)
    c = np.array(coords[:max(steps, warmup)], copy=True)
    c[..., 1] = c[..., 1] / np.sin(th)
    c[..., 0] = c[..., 0] - c[..., 1] * np.cos(th)
    ucell = np.array([[1, 0, 0], [np.cos(th), np.sin(th), 0], [0, 0, 1]])
    s = engine.Setup(c, dims[:c.shape[0]], typ, rad, sres, False, ucell)
    plan = engine.Plan.from_setup(s, autotune=True)
    dev = plan.device
    d_c = torch.from_numpy(c).to(dev)
    d_t = torch.from_numpy(s.type_index.astype(np.int32).reshape(-1)).to(dev)
    d_s = torch.from_numpy(np.ascontiguousarray(s.scale, dtype=np.float64)).to(dev)
    box = [float(x) for x in s.L]
    plan.accumulate_frames(d_c[:warmup], d_t, d_s[:warmup], box, s.arith, s.wrap)
    plan.reset()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream(dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    plan.accumulate_frames(d_c[:steps], d_t, d_s[:steps], box, s.arith, s.wrap)
    e1.record(stream)
    torch.cuda.synchronize()
    plan.check_errors()
    ms = e0.elapsed_time(e1)
    iso = isolated_stages(plan, d_c, d_t, d_s, box, s, 30, stream)
    plan.close()
    return {"workload": "hii_phase_256^3_50k_atoms in the 120-degree cell (main_gromacs.py "
                        "-ct 120 transform)", "frames": steps, "frames_per_s": steps / (ms * 1e-3),
            "ms_per_frame": ms / steps,
            "deposition": "xy-coupled (MODE_XY): k_deposit_lists_xy, the tile lists with the "
                          "cross factor 2^(c b_x b_y) per MFMA operand element",
            "stage_ms_isolated": {k: round(v, 4) for k, v in iso.items()}}


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU baseline leg (and the precision leg's oracle comparison)")
    ap.add_argument("--no-cpu-all", action="store_true",
                    help="skip the all-cores CPU baseline (one oracle frame per host core)")
    ap.add_argument("--fft", choices=["fp32", "fp64"], default="fp32")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the fwd+bwd leg")
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the S(q) error report")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched leg")
    ap.add_argument("--no-isolated", action="store_true",
                    help="skip the one-stage-at-a-time timing pass")
    ap.add_argument("--no-mono", action="store_true", help="skip the 120-degree cell leg")
    ap.add_argument("--no-autotune", action="store_true",
                    help="use the default FFT launch choices instead of timing them at plan creation")
    ap.add_argument("--batch-frames", type=int, default=50)
    ap.add_argument("--no-dropin", action="store_true", help="skip the drop-in compute_sf leg")
    ap.add_argument("--dropin-frames", type=int, default=200)
    ap.add_argument("--adj-iters", type=int, default=5)
    ap.add_argument("--no-cylindrical", action="store_true",
                    help="skip the cylindrical-average (PLOT_RAD_NEW) leg")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 256 MiB device-to-device copy 3 times (the reference "
                         "kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the S(q) path runs on a single MI355X here "
              "(frames would shard across GPUs with one FP64 all-reduce of S(q)); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    import dens
    from mdsf import engine, synth

    dev = torch.device("cuda", 0)
    rad = dens.load_radii(os.path.join(PKG, "radii.txt"))
    nframes = max(args.steps, args.warmup, 1)
    log(f"generating workload: {nframes} frames, 50,000 atoms")
    coords, dims, typ = synth.hii(seed=2002, nframes=nframes)
    sres = synth.hii_sres()
    cpu_all = None
    if not args.no_cpu and not args.no_cpu_all:
        # before this process touches the GPU: the worker processes start
        # from a process without a GPU context
        log("cpu baseline, all cores: one frame per process through the FP64 numpy oracle")
        cpu_all = cpu_baseline_all_cores(coords, dims, typ, rad, sres)
    torch.cuda.set_device(dev)
    s = engine.Setup(coords, dims, typ, rad, sres, False, np.eye(3))
    assert list(s.N) == [256, 256, 256], s.N
    plan = engine.Plan.from_setup(s, fft64=(args.fft == "fp64"), autotune=not args.no_autotune)
    d_coords = torch.from_numpy(coords).to(dev)
    d_types = torch.from_numpy(s.type_index.astype(np.int32).reshape(-1)).to(dev)
    d_scale = torch.from_numpy(np.ascontiguousarray(s.scale, dtype=np.float64)).to(dev)
    box = [float(x) for x in s.L]
    stream = torch.cuda.current_stream(dev)
    # The timed region is the library's own frame loop (mdsf_accumulate_frames):
    # frame t + 1 prepared on this stream while frame t is deposited and earlier
    # frames transformed on the plan's two other streams (no stage events).
    log(f"warmup {args.warmup} frames")
    plan.accumulate_frames(d_coords[:args.warmup], d_types, d_scale[:args.warmup], box, s.arith,
                           s.wrap)
    plan.reset()
    torch.cuda.synchronize()
    log(f"timing {args.steps} frames")
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    plan.accumulate_frames(d_coords[:args.steps], d_types, d_scale[:args.steps], box, s.arith,
                           s.wrap)
    end.record(stream)
    torch.cuda.synchronize()
    plan.check_errors()
    total_ms = start.elapsed_time(end)
    log(f"{args.steps / (total_ms * 1e-3):.1f} frames/s")
    iso = None
    if not args.no_isolated:
        iso = isolated_stages(plan, d_coords, d_types, d_scale, box, s, 30, stream)
        log("per-stage avg ms (isolated): " + json.dumps({k: round(v, 4) for k, v in iso.items()}))
        if d_coords.shape[0] >= 2:
            iso["fft_x_pair"] = paired_x(plan, d_coords, d_types, d_scale, box, s, 15)
            log(f"x pass per frame, frames in pairs: {iso['fft_x_pair']:.4f} ms")

    n = [int(x) for x in s.N]
    fb = fft_bytes(n, args.fft == "fp64")
    updates = voxel_updates(s)
    # the roofline of the dominant kernel: its average duration alone (HIP
    # events around it on the stream it runs on, one frame at a time), which
    # the kernel durations of a rocprofv3 trace of the same run match
    roofline = None
    if iso is not None:
        rdep_ms = iso["deposit_kernel"]
        rdep_tflops = 2.0 * updates / (rdep_ms * 1e-3) / 1e12
        roofline = {"bound": "mfma",
                    "kernel": "k_deposit_lists<16> (deposition + fold, FP32 MFMA, from the "
                              "per-tile atom lists the prepare stage builds)",
                    "achieved": round(rdep_tflops, 3), "peak": FP32_PEAK_TFLOPS,
                    "unit": "TFLOP/s", "frac": round(rdep_tflops / FP32_PEAK_TFLOPS, 4),
                    "traffic": DEPOSIT_TRAFFIC,
                    "basis": f"{updates} voxel updates/frame x 2 flop (one FMA each) / avg "
                             f"deposition kernel {rdep_ms * 1e3:.1f} us (alone, one frame at a "
                             f"time, HIP events recorded with the kernel's own dispatch)"}

    out = {
        "metric": METRIC,
        "value": args.steps / (total_ms * 1e-3),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.fft == "fp32" else "f32 deposit / f64 fft",
        "data": "synthetic",
        "config": {"workload": "hii_phase_256^3_50k_atoms", "grid": n, "atoms": int(s.natoms),
                   "box_A": [float(x) for x in s.L], "fft_precision": args.fft,
                   "sf_accumulation": "f64"},
        "roofline": roofline,
        "deposit": {"voxel_updates_per_frame": updates},
    }
    if iso is not None:
        out["deposit"]["updates_per_s"] = round(updates / (iso["deposit_kernel"] * 1e-3), 1)
        out["deposit"]["kernel_ms"] = round(iso["deposit_kernel"], 4)
        out["stage_ms_isolated"] = {k: round(v, 4) for k, v in iso.items()}
        out["fft_passes"] = {
            k: {"achieved": round(fb[k] / (iso[st] * 1e-3) / 1e9, 1),
                "frac": round(fb[k] / (iso[st] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                "bytes": fb[k], "avg_ms": round(iso[st], 4)}
            for k, st in (("z", "fft_z"), ("y", "fft_y"), ("x", "fft_x_acc"), ("x_pair", "fft_x_pair"))
            if st in iso}
        out["fft_passes_note"] = ("each pass alone (one frame at a time; HIP events recorded "
                                  "with the pass kernel's own dispatch); x_pair: the x pass as "
                                  "the frame loop runs it, two frames per launch, per frame; "
                                  "algorithmic bytes / time vs 8 TB/s")
        if "fft_x_pair" in iso:
            out["fft_ms_per_frame"] = {
                "one_frame": round(iso["fft_z"] + iso["fft_y"] + iso["fft_x_acc"], 4),
                "frame_loop": round(iso["fft_z"] + iso["fft_y"] + iso["fft_x_pair"], 4)}
    oracle_sf = None
    if not args.no_cpu:
        log("cpu baseline: frame 0 through the FP64 numpy oracle on one core")
        out["cpu_baseline"], oracle_sf = cpu_baseline(coords[0], dims, typ, rad, sres)
        if cpu_all is not None:
            out["cpu_baseline_all_cores"] = cpu_all
    if not args.no_precision:
        log("precision: fp32 / fp64 plans vs the oracle (frame 0), fp32 vs fp64 (2 frames)")
        out["precision"] = precision_leg(coords, s, oracle_sf)
    del oracle_sf
    if not args.no_mono:
        log("120-degree cell leg")
        out["mono120"] = mono_leg(coords, dims, typ, rad, sres, args.steps, args.warmup)
    if not args.no_cylindrical:
        log("cylindrical average of the 256^3 kgridplt")
        out["cylindrical"] = cylindrical_leg(plan, box, np.eye(3))
    if not args.no_batch:
        log(f"batched leg: 32 systems x {args.batch_frames} frames")
        out["batched"] = batch_leg(rad, args.batch_frames)
    if not args.no_dropin:
        log("drop-in compute_sf leg (host arrays in, .npz out)")
        out["dropin"] = dropin_leg(rad, args.dropin_frames)
    if not args.no_adjoint:
        log("fwd+bwd leg (adjoint workload)")
        out["fwd_bwd"] = adjoint_leg(rad, args.adj_iters)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
