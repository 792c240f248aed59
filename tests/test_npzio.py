"""Parallel .npz writer: np.load / zipfile read back what np.savez_compressed would hold."""
import zipfile

import numpy as np
import pytest


@pytest.mark.parametrize("force_zip64", [False, True])
def test_parallel_npz_round_trip(tmp_path, force_zip64, monkeypatch):
    from mdsf import npzio
    monkeypatch.setattr(npzio, "CHUNK", 1 << 16)   # many chunks per member
    rng = np.random.default_rng(0)
    arrays = dict(sf=rng.random((20, 18, 11)), sfplt=rng.random((18, 16, 21)),
                  L=np.array([10.0, 11.0, 12.0], dtype=np.float32), N=np.array([20, 18, 20]),
                  kgrid=rng.random((20, 18, 11, 4)), empty=np.zeros((0, 3)),
                  c=np.arange(7) + 1j * np.arange(7), names=np.array(["O", "H", "Na"]),
                  fortran=np.asfortranarray(rng.random((5, 6))), scalar=np.float64(3.5))
    path = npzio.savez_compressed(tmp_path / "x", threads=4, force_zip64=force_zip64, **arrays)
    with zipfile.ZipFile(path) as z:
        assert z.testzip() is None
        assert sorted(z.namelist()) == sorted(k + ".npy" for k in arrays)
    d = np.load(path)
    for k, v in arrays.items():
        assert d[k].dtype == np.asarray(v).dtype and d[k].shape == np.asarray(v).shape
        assert np.array_equal(d[k], v)
    ref = tmp_path / "ref.npz"
    np.savez_compressed(ref, **arrays)
    r = np.load(ref)
    for k in arrays:
        assert np.array_equal(r[k], d[k])


def test_future_members_resolve_in_order(tmp_path):
    import concurrent.futures
    from mdsf import npzio
    rng = np.random.default_rng(3)
    a = rng.random((50, 40))
    b = np.arange(1000, dtype=np.int32)
    with concurrent.futures.ThreadPoolExecutor(1) as pool:
        fa = pool.submit(lambda: a.copy())
        fb = pool.submit(lambda: b.copy())
        path = npzio.savez_compressed(str(tmp_path / "f"), threads=4, a=fa, L=np.float32(2.5), b=fb)
    d = np.load(path)
    assert d.files == ["a", "L", "b"]
    assert np.array_equal(d["a"], a) and np.array_equal(d["b"], b) and d["L"] == np.float32(2.5)
