"""Drop-in ``load_traj`` module: trajectory -> .npz without mdtraj.

Same names and outputs as the reference ``load_traj.py`` for the functions
its driver uses (``main_gromacs.py:197``): ``load_gro`` (load_traj.py:11-20)
and ``process_gro_mdtraj`` (load_traj.py:90-111).  Topologies are .gro files;
trajectories are .trr (XDR, single or double precision) or .gro files, read
by ``mdsf.trajio``.  The MDAnalysis-based ``process`` / ``process_gro`` of the
reference are not provided.
"""
import numpy as np

from mdsf import trajio

# masses of the elements an atom name's leading letters name (g/mol); the
# .npz ``mass`` key is informational (nothing on the S(q) path reads it)
_MASS = {"H": 1.008, "C": 12.011, "N": 14.007, "O": 15.999, "F": 18.998, "NA": 22.990,
         "MG": 24.305, "P": 30.974, "S": 32.06, "CL": 35.45, "K": 39.098, "CA": 40.078,
         "FE": 55.845, "ZN": 65.38, "BR": 79.904, "I": 126.904}


def _mass(name):
    letters = "".join(ch for ch in name.upper() if ch.isalpha())
    if letters[:2] in ("NA", "CL", "MG", "ZN", "FE", "BR") and letters[:2] == name.upper()[:2]:
        return _MASS[letters[:2]]
    return _MASS.get(letters[:1], 0.0)


def load_gro(gro):
    """Atom names of a .gro file (columns 11-15 of each atom line), load_traj.py:11-20."""
    return trajio.gro_atom_names(gro)


def process_gro_mdtraj(topology_filename, trajectory_filename, output_filename):
    """Trajectory -> ``output_filename``.npz with ``dims``, ``coords`` (Angstrom,
    float32: nm x 10), ``name``, ``mass`` and ``typ`` (= ``name``), as
    load_traj.py:90-111 writes them."""
    print("processing ", trajectory_filename)
    xyz, box = trajio.read_frames(trajectory_filename)
    _, _, names, _ = trajio.read_gro(topology_filename) if str(topology_filename).lower().endswith(
        ".gro") else (None, None, None, None)
    if names is None:
        raise NotImplementedError("topologies are read from .gro files")
    if xyz.shape[1] != len(names):
        raise ValueError(f"topology has {len(names)} atoms, trajectory {xyz.shape[1]}")
    coords = xyz * 10                                          # load_traj.py:95
    dims = trajio.cell_lengths(box).astype(np.float32) * 10    # load_traj.py:96
    name = np.array(names)
    mass = np.array([_mass(n) for n in names])
    print("saving ", output_filename)
    np.savez_compressed(output_filename, dims=dims, coords=coords, name=name, mass=mass, typ=name)
    print("done saving")
