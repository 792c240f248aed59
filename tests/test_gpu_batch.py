"""Batched systems (one HIP graph).

* The small batch (5 systems, 32-64^3) against stand-alone plans with the
  same autotuned kernels: the graph must reproduce their S(q) bit for bit,
  replay after replay, with eager work in between, and never modify its input
  coordinates.
* The 32-system workload of the bench (grids {64, 96, 128, 160, 192}^3,
  boxes 30-80 A, 5k-60k H/C/O/Na atoms: radix-3 and radix-5 grids, Nborder up
  to 107, hundreds to thousands of corner atoms per frame, all 16 graph
  branches), first 3 frames, plus two reduced systems (a few hundred atoms,
  Na at the cell corners, Nborder > 32): every system's S(q) against a
  stand-alone FP64 plan, and the two reduced ones against the FP64 numpy
  oracle (dens.py:277-318 per system).  Bound: max relative error 1e-5 over
  voxels above 1e-6 max(S).
"""
import numpy as np
import pytest
import torch

from oracle import dens_oracle as O

pytestmark = pytest.mark.gpu

SQ_REL = 1e-5
SQ_FLOOR = 1e-6


def _systems():
    from mdsf import synth
    return synth.batch(seed=11, nsys=5, nframes=3, grids=(32, 48, 64), box=(20.0, 30.0),
                       natoms=(300, 1500))


def _bind(batch, systems, setups):
    dev = batch.device
    bound = []
    for i, ((c, d, t, sres), s) in enumerate(zip(systems, setups)):
        coords = torch.from_numpy(np.ascontiguousarray(c)).to(dev)
        types = torch.from_numpy(s.type_index.astype(np.int32).reshape(-1)).to(dev)
        scale = torch.from_numpy(np.ascontiguousarray(s.scale, dtype=np.float64)).to(dev)
        batch.set_frames(i, coords, types, scale)
        bound.append((coords, types, scale))
    return bound


def _rel_err(got, want, floor=SQ_FLOOR):
    mask = want > floor * want.max()
    e = np.abs(got[mask] - want[mask]) / want[mask]
    return float(e.max()), float(np.median(e))


def test_batch_matches_single_plans_and_replays(rad, gpu):
    from mdsf import engine
    systems = _systems()
    setups = [engine.Setup(c, d, t, rad, sres, False, np.eye(3)) for c, d, t, sres in systems]
    for s in setups:
        assert int(s.N[0]) in (32, 48, 64)
    batch = engine.Batch(setups)
    dev = batch.device
    bound = _bind(batch, systems, setups)
    # eager references: stand-alone autotuned plans on copies of the frames
    refs = []
    for (coords, types, scale), s in zip(bound, setups):
        plan = engine.Plan.from_setup(s, autotune=True)
        plan.accumulate_frames(coords.clone(), types, scale, [float(x) for x in s.L], s.arith,
                               s.wrap)
        plan.check_errors()
        refs.append(plan.sf().cpu().numpy())
        plan.close()
    for rep in range(3):
        batch.run()
        batch.check_errors()
        for i, ref in enumerate(refs):
            got = batch.sf(i).cpu().numpy()
            assert np.array_equal(got, ref), (rep, i, np.max(np.abs(got - ref)))
        # eager work on the same device between replays
        torch.cuda.synchronize()
        junk = torch.randn(1 << 20, device=dev).sum()
        assert torch.isfinite(junk)
    for (coords, _, _), (c, _, _, _) in zip(bound, systems):
        assert np.array_equal(coords.cpu().numpy(), c)
    batch.close()


def _reduced(system, keep, seed, frames=2):
    """The first `keep` atoms and `frames` frames of a batch system, with 8 of
    its Na atoms moved to within 1.5 A of the 8 cell corners (windows off the
    core on all axes)."""
    c, d, t, sres = system
    rng = np.random.default_rng(seed)
    typ = list(t[:keep])
    d = d[:frames]
    coords = np.array(c[:frames, :keep], dtype=np.float32)
    L = float(d[0, 0])
    na = [i for i, n in enumerate(typ) if n == "Na"][:8]
    assert len(na) == 8
    for k, i in enumerate(na):
        corner = np.array([(k >> 2) & 1, (k >> 1) & 1, k & 1], dtype=np.float64) * L
        p = np.mod(corner + rng.uniform(-1.5, 1.5, 3), L)
        coords[:, i] = p.astype(np.float32)
    return coords, d, typ, sres


def test_bench_batch_vs_fp64_plans_and_oracle(rad, gpu):
    from mdsf import engine, synth
    full = synth.batch(nframes=50)           # the bench's 32 systems (seed 3003)
    systems = [(np.ascontiguousarray(c[:3]), d[:3], t, sres) for c, d, t, sres in full]
    del full
    # two reduced systems: a radix-3 (96^3) and a radix-5 (160^3) grid
    pick = {}
    for i, (c, d, t, sres) in enumerate(systems):
        n = int(float(d[0, 0]) / sres)
        # Na half-width > 32 cells: L < 50 A at 96^3, L < 80 A at 160^3; the
        # 160^3 box is kept above 40 A so that the oracle's padded grid stays small
        lo, hi = (30.0, 50.0) if n == 96 else (40.0, 60.0)
        if n in (96, 160) and n not in pick and lo < float(d[0, 0]) < hi:
            pick[n] = i
    assert sorted(pick) == [96, 160]
    reduced = [_reduced(systems[pick[n]], 200, 5 + n) for n in (96, 160)]
    systems += reduced
    setups = [engine.Setup(c, d, t, rad, sres, False, np.eye(3)) for c, d, t, sres in systems]
    grids = sorted({int(s.N[0]) for s in setups})
    assert grids == [64, 96, 128, 160, 192]
    assert all(s.nborder > 32 for s in setups[-2:])
    batch = engine.Batch(setups)
    bound = _bind(batch, systems, setups)
    batch.run()
    batch.check_errors()
    worst = 0.0
    for i, ((coords, types, scale), s) in enumerate(zip(bound, setups)):
        got = batch.sf(i).cpu().numpy()
        plan = engine.Plan.from_setup(s, fft64=True)
        plan.accumulate_frames(coords.clone(), types, scale, [float(x) for x in s.L], s.arith,
                               s.wrap)
        plan.check_errors()
        ref = plan.sf().cpu().numpy()
        plan.close()
        emax, emed = _rel_err(got, ref)
        worst = max(worst, emax)
        print(f"system {i}: N={int(s.N[0])} atoms={s.natoms} Nb={s.nborder} "
              f"fp32 batch vs fp64 plan max {emax:.2e} median {emed:.2e}", flush=True)
        assert emax <= SQ_REL, (i, emax)
    for k, (c, d, t, sres) in enumerate(reduced):
        i = len(systems) - 2 + k
        want = O.compute_sf(np.array(c), np.array(d), t, None, rad, np.eye(3), sres,
                            better=False)["sf"]
        got = batch.sf(i).cpu().numpy()
        emax, emed = _rel_err(got, want)
        print(f"reduced system {i}: N={int(setups[i].N[0])} Nb={setups[i].nborder} "
              f"batch vs oracle max {emax:.2e} median {emed:.2e}", flush=True)
        assert emax <= SQ_REL, (i, emax)
    batch.close()
