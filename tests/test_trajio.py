"""Trajectory I/O without mdtraj (load_traj.py:11-20, 90-111).

* ``load_gro`` against the reference's own test (test/test_load_traj.py:16-25)
  on its fixture ``water.gro`` (copied to tests/golden/).
* .trr frames against bytes assembled here field by field from the GROMACS
  trr layout (magic, version string, 13 header ints, t / lambda, box, x) in
  single and double precision; reader / writer round trips.
* ``process_gro_mdtraj``: the .npz keys, dtypes and values the reference
  writes (coords = nm x 10 in float32, dims = cell lengths x 10, typ = name).
"""
import os
import struct

import numpy as np
import pytest

from conftest import GOLDEN

WATER_GRO = os.path.join(GOLDEN, "water.gro")


def test_load_gro_names_match_reference_test():
    import load_traj
    names = load_traj.load_gro(WATER_GRO)
    assert names[:6] == ["OW1", "HW2", "HW3", "OW1", "HW2", "HW3"]


def test_read_gro_fixture():
    from mdsf import trajio
    xyz, box, names, res = trajio.read_gro(WATER_GRO)
    assert xyz.shape == (1, 6, 3) and xyz.dtype == np.float32
    assert np.allclose(xyz[0, 0], [0.126, 1.624, 1.679])
    assert np.allclose(np.diag(box[0]), [1.8206] * 3) and res[0] == "WATER"


def _frame_bytes(real, natoms, step, t, box, x):
    f = ">f" if real == 4 else ">d"
    b = struct.pack(">i", 1993) + struct.pack(">i", 13) + struct.pack(">i", 12) + b"GMX_trn_file"
    b += struct.pack(">13i", 0, 0, 9 * real, 0, 0, 0, 0, 3 * natoms * real, 0, 0, natoms, step, 0)
    b += struct.pack(f, t) + struct.pack(f, 0.0)
    b += b"".join(struct.pack(f, v) for v in np.asarray(box).reshape(-1))
    b += b"".join(struct.pack(f, v) for v in np.asarray(x).reshape(-1))
    return b


@pytest.mark.parametrize("real", [4, 8])
def test_read_trr_against_assembled_bytes(tmp_path, real):
    from mdsf import trajio
    rng = np.random.default_rng(0)
    frames = [rng.uniform(0, 3, (5, 3)) for _ in range(3)]
    boxes = [np.diag([3.0, 3.1, 3.2]) + (t == 2) * np.array([[0, 0, 0], [-1.5, 0, 0], [0, 0, 0]])
             for t in range(3)]
    path = tmp_path / "a.trr"
    with open(path, "wb") as fh:
        for t in range(3):
            fh.write(_frame_bytes(real, 5, 10 * t, 0.5 * t, boxes[t], frames[t]))
    xyz, box, times, steps = trajio.read_trr(path)
    dt = np.float32 if real == 4 else np.float64
    for t in range(3):
        assert np.array_equal(xyz[t], frames[t].astype(dt).astype(np.float32))
        assert np.array_equal(box[t], boxes[t].astype(dt).astype(np.float64))
    assert list(steps) == [0, 10, 20] and np.allclose(times, [0.0, 0.5, 1.0])
    # the writer produces the same bytes
    out = tmp_path / "b.trr"
    trajio.write_trr(out, np.stack(frames), np.stack(boxes), times=[0.0, 0.5, 1.0],
                     steps=[0, 10, 20], double=(real == 8))
    assert out.read_bytes() == path.read_bytes()


def test_gro_write_read_round_trip(tmp_path):
    from mdsf import trajio
    rng = np.random.default_rng(1)
    xyz = np.round(rng.uniform(0, 5, (2, 7, 3)), 3).astype(np.float32)
    box = np.array([[5.0, 0, 0], [-2.5, 4.33013, 0], [0, 0, 6.0]])
    names = ["C", "O", "H", "H", "NA", "C1", "HW2"]
    trajio.write_gro(tmp_path / "s.gro", xyz, box, names)
    rx, rb, rn, _ = trajio.read_gro(tmp_path / "s.gro")
    assert np.array_equal(rx, xyz) and rn == names
    assert np.allclose(rb[0], box) and np.allclose(trajio.cell_lengths(rb[0]), [5.0, 5.0, 6.0],
                                                    atol=1e-5)


def test_process_gro_mdtraj_npz(tmp_path):
    import load_traj
    from mdsf import trajio
    rng = np.random.default_rng(2)
    names = ["O", "H", "H"] * 4
    xyz = rng.uniform(0, 2.5, (3, 12, 3)).astype(np.float32)
    box = np.diag([2.5, 2.5, 2.5])
    trajio.write_gro(tmp_path / "top.gro", xyz[:1], box, names)
    trajio.write_trr(tmp_path / "traj.trr", xyz, box)
    load_traj.process_gro_mdtraj(str(tmp_path / "top.gro"), str(tmp_path / "traj.trr"),
                                 str(tmp_path / "out"))
    d = np.load(tmp_path / "out.npz")
    assert set(d.files) == {"dims", "coords", "name", "mass", "typ"}
    assert d["coords"].dtype == np.float32 and np.array_equal(d["coords"], xyz * 10)
    assert d["dims"].dtype == np.float32 and np.allclose(d["dims"], 25.0)
    assert list(d["typ"]) == names and list(d["name"]) == names
    assert np.allclose(d["mass"][:3], [15.999, 1.008, 1.008])
