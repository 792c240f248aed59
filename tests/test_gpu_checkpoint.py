"""dens.compute_sf with CHECKPOINT_FRAMES: a call stopped after a saved block
and called again resumes there, and its output is bitwise that of an
uninterrupted call with the same block size (the blocks' FP64 sums are added
in the same order); against a call without checkpoints, S(q) agrees to FP64
rounding.  The interruption is an exception raised on the host right after a
checkpoint is written."""
import os

import numpy as np
import pytest

from test_gpu_xy_lists import _xy_system

pytestmark = pytest.mark.gpu


def _run(coords, dims, typ, rad, out, every):
    import dens
    saved = (dens.PRINT_DETAILS, dens.CHECKPOINT_FRAMES)
    dens.PRINT_DETAILS, dens.CHECKPOINT_FRAMES = False, every
    try:
        r = np.array(coords)
        dens.compute_sf(r, dims, typ, out, rad, np.eye(3), 0.5)
    finally:
        dens.PRINT_DETAILS, dens.CHECKPOINT_FRAMES = saved
    with np.load(out + ".npz") as z:
        return r, {k: np.array(z[k]) for k in z.files}


def test_checkpoint_resume_bitwise(rad, gpu, tmp_path, monkeypatch):
    import dens
    coords, dims, typ = _xy_system(71, frames=7, box=(20.0, 22.0, 24.0), n=150)
    full_r, full = _run(coords, dims, typ, rad, str(tmp_path / "full"), 2)
    assert not os.path.exists(str(tmp_path / "full") + ".ckpt.npz")
    assert dens.LAST_TIMINGS["resumed_from"] == 0
    out = str(tmp_path / "cut")
    real_save = dens._checkpoint_save
    calls = []

    def stop_after_second(*a):
        real_save(*a)
        calls.append(a[2])
        if len(calls) == 2:
            raise RuntimeError("interrupted")

    monkeypatch.setattr(dens, "_checkpoint_save", stop_after_second)
    with pytest.raises(RuntimeError, match="interrupted"):
        _run(coords, dims, typ, rad, out, 2)
    monkeypatch.setattr(dens, "_checkpoint_save", real_save)
    assert calls == [2, 4] and os.path.exists(out + ".ckpt.npz")
    assert not os.path.exists(out + ".npz")
    cut_r, cut = _run(coords, dims, typ, rad, out, 2)
    assert dens.LAST_TIMINGS["resumed_from"] == 4
    assert not os.path.exists(out + ".ckpt.npz")
    assert np.array_equal(cut_r, full_r)
    for k in full:
        assert np.array_equal(cut[k], full[k]), k
    plain_r, plain = _run(coords, dims, typ, rad, str(tmp_path / "plain"), 0)
    assert np.array_equal(plain_r, full_r)
    np.testing.assert_allclose(full["sf"], plain["sf"], rtol=1e-12, atol=1e-12 * plain["sf"].max())
