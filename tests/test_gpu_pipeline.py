"""The pipelined frame loop (three streams, two frame slots) against the same
frames run stage by stage on one stream: S(q) must be bit-identical (every
kernel is deterministic and the S(q) sum keeps frame order), for both plan
precisions and for an odd frame count; the stage timing API reports positive
per-stage averages."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fp64", [False, True])
@pytest.mark.parametrize("nframes", [1, 2, 5])
def test_pipelined_loop_matches_sequential(nframes, fp64, rad, gpu):
    from mdsf import engine, synth
    coords, dims, typ = synth.uniform(7, 6000, 36.0, ["H", "C", "O", "Na"], nframes)
    s = engine.Setup(coords, dims, typ, rad, 36.0 / 80.5, False, np.eye(3))
    assert list(s.N) == [80, 80, 80]
    dev = torch.device("cuda", 0)
    types = torch.from_numpy(s.type_index.astype(np.int32).reshape(-1)).to(dev)
    scale = torch.from_numpy(np.ascontiguousarray(s.scale, dtype=np.float64)).to(dev)
    box = [float(x) for x in s.L]
    pa = engine.Plan.from_setup(s, fft64=fp64)
    pa.set_stage_timing(True)
    ca = torch.from_numpy(coords.copy()).to(dev)
    pa.accumulate_frames(ca, types, scale, box, s.arith, s.wrap)
    pa.check_errors()
    got = pa.sf().cpu().numpy()
    times = pa.stage_times()
    assert all(v > 0 for v in times.values()), times
    pb = engine.Plan.from_setup(s, fft64=fp64)
    cb = torch.from_numpy(coords.copy()).to(dev)
    for t in range(nframes):
        pb.set_frame_buffer(t & 1)
        pb.prepare_frame(cb[t], types, scale[t], box, s.arith, s.wrap)
        pb.deposit()
        for i in range(3):
            pb.fft_pass(i)
    pb.check_errors()
    want = pb.sf().cpu().numpy()
    assert np.array_equal(got, want)
    assert torch.equal(ca, cb)   # the in-place rescale / wrap is the same
    pa.close()
    pb.close()
