"""GROMACS .gro and .trr files without mdtraj (what ``load_traj.py`` reads through it).

* ``.gro``: fixed-column text (title, atom count, one line per atom
  ``%5d%-5s%5s%5d%8.3f%8.3f%8.3f[velocities]``, box line).  Several frames may
  follow each other.  The box line holds v1(x) v2(y) v3(z) [v1(y) v1(z) v2(x)
  v2(z) v3(x) v3(y)] in nm.
* ``.trr``: XDR (big-endian) frames: magic 1993, the version string
  "GMX_trn_file", the block sizes (ir, e, box, vir, pres, top, sym, x, v, f),
  natoms, step, nre, t, lambda, then the box (3x3), virial, pressure,
  positions, velocities and forces whose sizes are non-zero, each in single
  or double precision (the box / position sizes tell which).

Readers return float32 nm positions, as the trajectory libraries do; the
caller applies the Angstrom conversion (load_traj.py:95-96).  Unit-cell
lengths are the norms of the box vectors.
"""
import struct

import numpy as np

TRR_MAGIC = 1993
TRR_VERSION = b"GMX_trn_file"


# ------------------------------------------------------------------------ .gro
def gro_atom_names(path):
    """Atom names, columns 11-15 of each atom line (load_traj.py:11-20)."""
    with open(path) as f:
        lines = f.readlines()
    return [line[10:15].strip() for line in lines[2:-1]]


def _box_vectors(fields):
    v = [float(x) for x in fields]
    box = np.zeros((3, 3))
    box[0, 0], box[1, 1], box[2, 2] = v[0], v[1], v[2]
    if len(v) >= 9:
        box[0, 1], box[0, 2], box[1, 0], box[1, 2], box[2, 0], box[2, 1] = v[3:9]
    return box


def read_gro(path):
    """All frames of a .gro file: (xyz float32 (F, N, 3) nm, box (F, 3, 3) nm,
    names, residue names)."""
    frames, boxes = [], []
    names = resnames = None
    with open(path) as f:
        lines = f.read().splitlines()
    i = 0
    while i < len(lines):
        if not lines[i].strip() and i + 1 >= len(lines):
            break
        n = int(lines[i + 1].strip())
        atoms = lines[i + 2:i + 2 + n]
        if len(atoms) != n:
            raise ValueError(f"{path}: truncated frame at line {i + 1}")
        xyz = np.empty((n, 3), dtype=np.float32)
        for k, line in enumerate(atoms):
            xyz[k, 0] = float(line[20:28])
            xyz[k, 1] = float(line[28:36])
            xyz[k, 2] = float(line[36:44])
        if names is None:
            names = [line[10:15].strip() for line in atoms]
            resnames = [line[5:10].strip() for line in atoms]
        frames.append(xyz)
        boxes.append(_box_vectors(lines[i + 2 + n].split()))
        i += n + 3
    return np.stack(frames), np.stack(boxes), names, resnames


def write_gro(path, xyz, box, names, resnames=None, resids=None, title="mdsf"):
    """Write frames (F, N, 3) nm with orthorhombic or triclinic boxes (F, 3, 3) nm."""
    xyz = np.asarray(xyz)
    if xyz.ndim == 2:
        xyz = xyz[None]
    box = np.asarray(box, dtype=np.float64)
    if box.ndim == 2:
        box = np.broadcast_to(box, (xyz.shape[0], 3, 3))
    n = xyz.shape[1]
    resnames = resnames or ["MOL"] * n
    resids = resids if resids is not None else [1] * n
    with open(path, "w") as f:
        for t in range(xyz.shape[0]):
            f.write(f"{title}\n{n:5d}\n")
            for k in range(n):
                f.write("%5d%-5s%5s%5d%8.3f%8.3f%8.3f\n" % (
                    resids[k] % 100000, resnames[k][:5], names[k][:5], (k + 1) % 100000,
                    xyz[t, k, 0], xyz[t, k, 1], xyz[t, k, 2]))
            b = box[t]
            vals = [b[0, 0], b[1, 1], b[2, 2]]
            off = [b[0, 1], b[0, 2], b[1, 0], b[1, 2], b[2, 0], b[2, 1]]
            if any(v != 0.0 for v in off):
                vals += off
            f.write(" ".join("%10.5f" % v for v in vals) + "\n")


# ------------------------------------------------------------------------ .trr
_HDR_INTS = ("ir", "e", "box", "vir", "pres", "top", "sym", "x", "v", "f", "natoms", "step",
             "nre")


def _trr_header(buf, pos):
    magic, = struct.unpack_from(">i", buf, pos)
    if magic != TRR_MAGIC:
        raise ValueError(f"not a .trr frame at byte {pos} (magic {magic})")
    pos += 4
    slen, = struct.unpack_from(">i", buf, pos)     # string length incl. the terminator
    pos += 4
    n, = struct.unpack_from(">i", buf, pos)         # XDR string: length, bytes, pad to 4
    pos += 4 + ((n + 3) // 4) * 4
    vals = struct.unpack_from(">13i", buf, pos)
    pos += 13 * 4
    h = dict(zip(_HDR_INTS, vals))
    natoms = h["natoms"]
    if h["box"]:
        real = h["box"] // 9
    elif h["x"]:
        real = h["x"] // (3 * natoms)
    elif h["v"]:
        real = h["v"] // (3 * natoms)
    else:
        real = h["f"] // (3 * natoms) if natoms else 4
    if real not in (4, 8):
        raise ValueError(f"unsupported real size {real} in .trr header")
    fmt = ">f4" if real == 4 else ">f8"
    t, lam = np.frombuffer(buf, dtype=fmt, count=2, offset=pos)
    pos += 2 * real
    h.update(real=real, fmt=fmt, t=float(t), lam=float(lam))
    return h, pos


def read_trr(path):
    """All frames of a .trr file: (xyz float32 (F, N, 3) nm, box (F, 3, 3) nm,
    time (F,), step (F,)).  Frames without positions are skipped."""
    with open(path, "rb") as f:
        buf = f.read()
    pos = 0
    xyzs, boxes, times, steps = [], [], [], []
    while pos < len(buf):
        h, pos = _trr_header(buf, pos)
        box = np.zeros((3, 3))
        for key in ("box", "vir", "pres"):
            if h[key]:
                m = np.frombuffer(buf, dtype=h["fmt"], count=9, offset=pos).reshape(3, 3)
                if key == "box":
                    box = m.astype(np.float64)
                pos += h[key]
        xyz = None
        for key in ("x", "v", "f"):
            if h[key]:
                a = np.frombuffer(buf, dtype=h["fmt"], count=3 * h["natoms"], offset=pos)
                if key == "x":
                    xyz = a.reshape(h["natoms"], 3).astype(np.float32)
                pos += h[key]
        if xyz is not None:
            xyzs.append(xyz)
            boxes.append(box)
            times.append(h["t"])
            steps.append(h["step"])
    if not xyzs:
        raise ValueError(f"{path}: no frames with positions")
    return np.stack(xyzs), np.stack(boxes), np.asarray(times), np.asarray(steps)


def write_trr(path, xyz, box, times=None, steps=None, double=False):
    """Write positions (F, N, 3) nm and boxes (F, 3, 3) nm as a .trr (no v / f)."""
    xyz = np.asarray(xyz)
    if xyz.ndim == 2:
        xyz = xyz[None]
    box = np.asarray(box, dtype=np.float64)
    if box.ndim == 2:
        box = np.broadcast_to(box, (xyz.shape[0], 3, 3))
    nf, n = xyz.shape[0], xyz.shape[1]
    real = 8 if double else 4
    fmt = ">f8" if double else ">f4"
    with open(path, "wb") as f:
        for t in range(nf):
            ver = TRR_VERSION
            head = struct.pack(">iii", TRR_MAGIC, len(ver) + 1, len(ver))
            head += ver + b"\0" * ((4 - len(ver) % 4) % 4)
            sizes = [0, 0, 9 * real, 0, 0, 0, 0, 3 * n * real, 0, 0, n,
                     int(steps[t]) if steps is not None else t, 0]
            head += struct.pack(">13i", *sizes)
            tt = float(times[t]) if times is not None else float(t)
            f.write(head)
            f.write(np.array([tt, 0.0], dtype=fmt).tobytes())
            f.write(np.ascontiguousarray(box[t], dtype=fmt).tobytes())
            f.write(np.ascontiguousarray(xyz[t], dtype=fmt).tobytes())


# --------------------------------------------------------------------- general
def cell_lengths(box):
    """Norms of the box vectors (the unit-cell lengths trajectory libraries report)."""
    box = np.asarray(box, dtype=np.float64)
    return np.sqrt(np.sum(box * box, axis=-1))


def read_frames(path):
    """Positions (F, N, 3) float32 nm and boxes (F, 3, 3) nm from a .trr or .gro file."""
    low = str(path).lower()
    if low.endswith(".trr"):
        xyz, box, _, _ = read_trr(path)
    elif low.endswith(".gro"):
        xyz, box, _, _ = read_gro(path)
    else:
        raise NotImplementedError(f"{path}: only .trr and .gro trajectories are read")
    return xyz, box
