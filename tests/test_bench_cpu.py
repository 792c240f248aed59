"""The bench's all-cores CPU baseline: one oracle frame per worker process
(spawned, one thread each), on a small HII system so that it runs in
seconds."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def test_cpu_baseline_all_cores(rad):
    import bench
    from mdsf import synth
    coords, dims, typ = synth.hii(seed=5, nframes=3, natoms=600, layers=5, per_layer=5)
    rec = bench.cpu_baseline_all_cores(coords, dims, typ, rad, 1.5, max_procs=2)
    assert rec["cores"] == 2 and rec["kind"] == "port" and rec["unit"] == "frames/s"
    assert np.isfinite(rec["value"]) and rec["value"] > 0
    assert rec["sample"].startswith("frames 0..1 of the workload")
