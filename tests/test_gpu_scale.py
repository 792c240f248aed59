"""FP32 against FP64 at full size (GPU), and the Ewald slice at 512^3.

The FP64 plan is the reference here: FP64 deposition, FFT and frame sum,
pinned to the FP64 numpy oracle at small sizes by ``test_gpu_parity.py``.
The FP32 path (FP32 deposition and FFT, FP64 twiddles and frame sum) must
stay within max relative error 1e-5 of S(q) over voxels above 1e-6 * max S;
the measured max and median errors are printed.
"""
import time

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _sq(coords, dims, typ, rad, sres, fp64):
    from mdsf import engine
    s = engine.Setup(coords, dims, typ, rad, sres, False, np.eye(3))
    plan = engine.Plan.from_setup(s, fft64=fp64)
    dev = plan.device
    c = torch.from_numpy(np.ascontiguousarray(coords)).to(dev)
    types = torch.from_numpy(s.type_index.astype(np.int32).reshape(-1)).to(dev)
    scale = torch.from_numpy(np.ascontiguousarray(s.scale, dtype=np.float64)).to(dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan.accumulate_frames(c, types, scale, [float(x) for x in s.L], s.arith, s.wrap)
    plan.check_errors()
    dt = time.perf_counter() - t0
    return s, plan, plan.sf(), dt


def _err(got, want, floor=1e-6):
    m = torch.max(want)
    mask = want > floor * m
    e = torch.abs(got[mask] - want[mask]) / want[mask]
    return float(e.max()), float(e.median()), int(mask.sum())


def test_fp32_vs_fp64_hii_256(rad, gpu):
    from mdsf import synth
    coords, dims, typ = synth.hii(seed=2002, nframes=2)
    _, p32, s32, t32 = _sq(coords, dims, typ, rad, synth.hii_sres(), False)
    _, p64, s64, t64 = _sq(coords, dims, typ, rad, synth.hii_sres(), True)
    emax, emed, nv = _err(s32, s64)
    print(f"\nHII 256^3 2 frames: fp32 vs fp64 S(q) max rel {emax:.3e} median {emed:.3e} "
          f"over {nv} voxels; fp32 {t32 * 1e3:.1f} ms, fp64 {t64 * 1e3:.1f} ms")
    assert emax <= 1e-5


def test_hii_frame_vs_oracle(rad, gpu):
    """One frame of the 256^3 HII workload (Nborder 55: the corner-slot bands
    of dens.py:107 span 4 z tiles, 85 corner atoms) through the FP32 and FP64
    plans against the FP64 numpy oracle (dens.py:283-318, about 20 s on one core)."""
    from oracle import dens_oracle as O
    from mdsf import synth
    coords, dims, typ = synth.hii(seed=2002, nframes=1)
    t0 = time.perf_counter()
    want = O.compute_sf(np.array(coords), np.array(dims), typ, None, rad, np.eye(3),
                        synth.hii_sres(), better=False)["sf"]
    t_oracle = time.perf_counter() - t0
    want = torch.from_numpy(want).to(gpu)
    for fp64 in (False, True):
        s, plan, got, _ = _sq(coords, dims, typ, rad, synth.hii_sres(), fp64)
        assert s.nborder == 55
        emax, emed, nv = _err(got, want)
        print(f"\nHII 256^3 frame 0, {'fp64' if fp64 else 'fp32'} plan vs oracle: max rel "
              f"{emax:.3e} median {emed:.3e} over {nv} voxels (oracle {t_oracle:.1f} s)")
        assert emax <= (1e-9 if fp64 else 1e-5)
        plan.close()


def test_large_512_precision_and_ewald(rad, gpu):
    """1M H/C/N/O atoms, 200 A box, 512^3, one frame; then the Ewald slices at 1.542 A."""
    from mdsf import engine, synth
    coords, dims, typ = synth.large(nframes=1)
    s, p32, s32, t32 = _sq(coords, dims, typ, rad, synth.large_sres(), False)
    assert list(s.N) == [512, 512, 512]
    _, p64, s64, t64 = _sq(coords, dims, typ, rad, synth.large_sres(), True)
    emax, emed, nv = _err(s32, s64)
    print(f"\n512^3 1M atoms: fp32 vs fp64 S(q) max rel {emax:.3e} median {emed:.3e} over {nv} "
          f"voxels; fp32 {t32 * 1e3:.1f} ms/frame, fp64 {t64 * 1e3:.1f} ms/frame")
    assert emax <= 1e-5
    del s64
    p64.close()
    sfplt, kgrid, kgridplt = p32.outputs(s32, [float(x) for x in s.L], s.box_f32)
    del kgrid
    t0 = time.perf_counter()
    xy, xz, yz = engine.ewald_slices(kgridplt, 1.542)
    torch.cuda.synchronize()
    print(f"Ewald slices (510^2, 510x511 x2) in {(time.perf_counter() - t0) * 1e3:.1f} ms")
    assert tuple(xy.shape) == (510, 510) and tuple(xz.shape) == (510, 511)
    # the sphere stays inside the grid near q = 0: the centre of each section is finite
    assert torch.isfinite(xy[255, 255]) and torch.isfinite(xz[255, 255])
