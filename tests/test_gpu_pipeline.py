"""The pipelined frame loop (three streams, three frame slots) against the
same frames run stage by stage on one stream: S(q) must be bit-identical
(every kernel is deterministic and the S(q) sum keeps frame order), for both
plan precisions and for an odd frame count, on a grid whose x pass takes two
frames per launch (64: the FP64 sums read and written once per pair, with the
additions in frame order; 4 frames: one pair, then the last two unpaired) and
one whose x pass does not (80); the stage timing API reports positive
per-stage averages."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("grid", [80, 64])
@pytest.mark.parametrize("fp64", [False, True])
@pytest.mark.parametrize("nframes", [1, 2, 4, 5])
def test_pipelined_loop_matches_sequential(nframes, fp64, grid, rad, gpu):
    from mdsf import engine, synth
    coords, dims, typ = synth.uniform(7, 6000, 36.0, ["H", "C", "O", "Na"], nframes)
    s = engine.Setup(coords, dims, typ, rad, 36.0 / (grid + 0.5), False, np.eye(3))
    assert list(s.N) == [grid, grid, grid]
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


@pytest.mark.parametrize("fp64", [False, True])
def test_pipelined_adjoint_matches_single_frame_entry_points(fp64, rad, gpu):
    """mdsf_adjoint_frames (two frame slots, three streams) against the same
    frames through mdsf_prepare_frame / mdsf_deposit / mdsf_density_adjoint /
    mdsf_deposit_adjoint on one stream: the same kernels in the same order per
    frame, so the gradients must be bit-identical (tiled MFMA adjoint for FP32,
    gather for FP64)."""
    from mdsf import engine, synth
    nframes = 3
    coords, dims, typ = synth.uniform(8, 3000, 24.0, ["H", "C", "O", "Na"], nframes)
    s = engine.Setup(coords, dims, typ, rad, 24.0 / 48.5, False, np.eye(3))
    dev = torch.device("cuda", 0)
    types = torch.from_numpy(s.type_index.astype(np.int32).reshape(-1)).to(dev)
    scale = torch.from_numpy(np.ascontiguousarray(s.scale, dtype=np.float64)).to(dev)
    box = [float(x) for x in s.L]
    rng = np.random.default_rng(9)
    pa = engine.Plan.from_setup(s, fft64=fp64)
    dsf = torch.from_numpy(rng.normal(size=pa.half_shape)).to(dev)
    ca = torch.from_numpy(coords.copy()).to(dev)
    got = pa.adjoint_frames(ca, types, scale, box, s.arith, s.wrap, dsf).cpu().numpy()
    pa.check_errors()
    pb = engine.Plan.from_setup(s, fft64=fp64)
    cb = torch.from_numpy(coords.copy()).to(dev)
    want = np.empty_like(got)
    rho = torch.empty(tuple(int(v) for v in s.N), dtype=pb.real_dtype, device=dev)
    for t in range(nframes):
        pb.set_frame_buffer(t & 1)
        pb.prepare_frame(cb[t], types, scale[t], box, s.arith, s.wrap)
        pb.deposit(out=rho)
        grho = pb.density_adjoint(rho, dsf)
        g = pb.deposit_adjoint(grho)
        want[t] = (g * scale[t][None, :]).cpu().numpy()
    pb.check_errors()
    assert np.array_equal(ca.cpu().numpy(), cb.cpu().numpy())
    assert np.array_equal(got, want), float(np.max(np.abs(got - want)))


def test_stage_timing_accumulates_over_calls(rad, gpu):
    """Stage timing of one-frame calls queued back to back: the timed frames
    add up until stage_times reads them, every stage has a positive time, and
    a read with no frame timed since the last one is an error."""
    from mdsf import engine, synth
    coords, dims, typ = synth.uniform(11, 3000, 30.0, ["H", "C", "O"], 4)
    s = engine.Setup(coords, dims, typ, rad, 30.0 / 64.5, False, np.eye(3))
    dev = torch.device("cuda", 0)
    types = torch.from_numpy(s.type_index.astype(np.int32).reshape(-1)).to(dev)
    scale = torch.from_numpy(np.ascontiguousarray(s.scale, dtype=np.float64)).to(dev)
    box = [float(x) for x in s.L]
    c = torch.from_numpy(coords.copy()).to(dev)
    p = engine.Plan.from_setup(s)
    p.set_stage_timing(1)
    for t in range(4):
        p.accumulate_frames(c[t:t + 1], types, scale[t:t + 1], box, s.arith, s.wrap)
    times = p.stage_times()
    assert all(v > 0 for v in times.values()), times
    with pytest.raises(ValueError):  # MDSF_ERR_ARG: nothing timed since the last read
        p.stage_times()
    p.set_stage_timing(0)
    p.check_errors()
    p.close()


@pytest.mark.parametrize("fp64", [False, True])
def test_split_prepare_stage(fp64, rad, gpu):
    """A grid of 2,048 tiles (256 x 256 x 128), where the frame loop runs the
    prepare stage in branches (tile lists on a third stream, the deposition
    waiting for them directly; FP64 plans have no lists): S(q) of 4 frames
    bitwise equal with and without stage timing (which records the list
    branch's end) and equal to the stage-by-stage loop; every stage time is
    positive."""
    from mdsf import engine, synth
    coords, dims, typ = synth.uniform(21, 3000, 40.0, ["H", "C", "O"], 4)
    s = engine.Setup(coords, dims, typ, rad, np.array([40 / 256.5, 40 / 256.5, 40 / 128.5]),
                     False, np.eye(3))
    assert list(s.N) == [256, 256, 128]
    dev = torch.device("cuda", 0)
    types = torch.from_numpy(s.type_index.astype(np.int32).reshape(-1)).to(dev)
    scale = torch.from_numpy(np.ascontiguousarray(s.scale, dtype=np.float64)).to(dev)
    box = [float(x) for x in s.L]
    out = []
    for timing in (0, 1):
        p = engine.Plan.from_setup(s, fft64=fp64)
        p.set_stage_timing(timing)
        c = torch.from_numpy(coords.copy()).to(dev)
        p.accumulate_frames(c, types, scale, box, s.arith, s.wrap)
        p.check_errors()
        out.append(p.sf().cpu().numpy())
        if timing:
            times = p.stage_times()
            assert all(v > 0 for v in times.values()), times
        p.close()
    assert np.array_equal(out[0], out[1])
    pb = engine.Plan.from_setup(s, fft64=fp64)
    cb = torch.from_numpy(coords.copy()).to(dev)
    for t in range(4):
        pb.set_frame_buffer(t % 3)
        pb.prepare_frame(cb[t], types, scale[t], box, s.arith, s.wrap)
        pb.deposit()
        for i in range(3):
            pb.fft_pass(i)
    pb.check_errors()
    assert np.array_equal(out[0], pb.sf().cpu().numpy())
    pb.close()
