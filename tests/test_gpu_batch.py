"""Batched systems (one HIP graph) against the same systems run one by one.

Each system of the batch has an autotuned plan; a stand-alone plan with the
same autotuned choices runs the same kernels, so the batch must reproduce
its S(q) bit for bit.  The graph is replayed several times with eager work
in between, and every replay is checked (a replay must not depend on state
left by an earlier one); the batch never modifies its input coordinates.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _systems():
    from mdsf import synth
    return synth.batch(seed=11, nsys=5, nframes=3, grids=(32, 48, 64), box=(20.0, 30.0),
                       natoms=(300, 1500))


def test_batch_matches_single_plans_and_replays(rad, gpu):
    from mdsf import engine
    systems = _systems()
    setups = [engine.Setup(c, d, t, rad, sres, False, np.eye(3)) for c, d, t, sres in systems]
    for s in setups:
        assert int(s.N[0]) in (32, 48, 64)
    batch = engine.Batch(setups)
    dev = batch.device
    bound = []
    for i, ((c, d, t, sres), s) in enumerate(zip(systems, setups)):
        coords = torch.from_numpy(np.ascontiguousarray(c)).to(dev)
        types = torch.from_numpy(s.type_index.astype(np.int32).reshape(-1)).to(dev)
        scale = torch.from_numpy(np.ascontiguousarray(s.scale, dtype=np.float64)).to(dev)
        batch.set_frames(i, coords, types, scale)
        bound.append((coords, types, scale))
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
