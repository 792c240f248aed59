"""Checkpoint files of dens.compute_sf (CHECKPOINT_FRAMES): the fingerprint
of a call's inputs, and the save / load round trip (host only)."""
import numpy as np


def test_checkpoint_round_trip(tmp_path):
    import dens
    rng = np.random.default_rng(3)
    r = rng.uniform(0, 10, (4, 7, 3)).astype(np.float32)
    L = np.full((4, 3), 10.0)
    typ = ["C", "O", "H", "N", "C", "O", "H"]
    key = dens._checkpoint_key(r, L, typ, 0.5, np.eye(3), [20, 20, 20])
    total = rng.normal(size=(20, 20, 11))
    out = str(tmp_path / "run")
    assert dens._checkpoint_load(out, key, total.shape) == (0, None)
    dens._checkpoint_save(out, key, 2, total)
    done, got = dens._checkpoint_load(out, key, total.shape)
    assert done == 2 and got.dtype == np.float64 and np.array_equal(got, total)
    # another call's inputs do not resume from it
    r2 = r.copy()
    r2[3, 6, 2] += 1e-3
    for other in (dens._checkpoint_key(r2, L, typ, 0.5, np.eye(3), [20, 20, 20]),
                  dens._checkpoint_key(r, L, typ, 0.6, np.eye(3), [20, 20, 20]),
                  dens._checkpoint_key(r, L, typ[::-1], 0.5, np.eye(3), [20, 20, 20]),
                  dens._checkpoint_key(r, L, typ, 0.5, np.eye(3), [20, 20, 22])):
        assert other != key
        assert dens._checkpoint_load(out, other, total.shape) == (0, None)
    assert dens._checkpoint_load(out, key, (20, 20, 12)) == (0, None)
