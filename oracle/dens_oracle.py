"""CPU oracle: FP64 numpy restatement of the reference S(q) path (``dens.py``).

TEST INFRASTRUCTURE ONLY.  The HIP path is tested against this module: only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it.  The product (``md-structure-factor_amd/``) never
calls it and has no CPU fallback.

Every function restates one piece of joeyelk/MD-Structure-Factor ``dens.py``
(cited ``file:line``), keeping the reference's arithmetic in the same order and
dtypes, including its quirks:

* the frame *sum* of |F|^2 (``dens.py:318``), not a mean;
* the corner-fold z index taken from the y side (``dens.py:107``);
* the asymmetric window ``[ir-A, ir+A)`` (``dens.py:292-297``);
* exact zeros wrapped to ``L`` (``dens.py:211-212``);
* the >=1e6-atom wrap branch that only wraps atom indices ``[0, frames)``
  (``dens.py:213-221``);
* numpy-2 (NEP 50) promotion: float32 trajectories stay float32 through
  rescale/wrap, ``dr`` and ``ir`` are float64.

How each function is pinned (see ``tests/test_oracle.py``):

* ``compute_sf`` end to end, and hence ``rescale``, ``grid_shape``, ``wrap``,
  ``borders``, ``deposit_frame``, ``fold``, ``get_dplot`` and ``kgrids`` as
  composed there: the golden vectors in ``tests/golden/*.npz``, written by the
  reference's own ``compute_sf`` (``tests/golden/make_golden.py``).
* ``fold_destination`` (closed form of the fold): brute-force agreement with
  ``fold`` on random padded grids, plus mass conservation.
* ``deposit_frame``: the single-atom closed form (product of 1-D Gaussians on an
  orthorhombic grid) and whole-cell translation invariance.
* ``dplot_index`` (closed form of ``get_dplot``): agreement with ``get_dplot``.
* ``load_radii``: known entries of ``radii.txt``.
"""
import math

import numpy as np

PRECISION = 1.0e-24          # dens.py:18
BUFFSIZE = 1000000           # dens.py:206


# --------------------------------------------------------------------- inputs
def load_radii(filename):
    """``{name: (Nel, sigma_angstrom)}`` from ``Nel name radius_pm`` lines (dens.py:23-35)."""
    table = {}
    with open(filename) as fh:
        for raw in fh.readlines():
            fields = raw.strip().split()
            table[fields[1]] = (float(fields[0]), float(fields[2]) / 100.0)
    return table


def borders(rad, dr, names):
    """Window half-widths in grid steps, per type, as float64 3-vectors (dens.py:38-43)."""
    out = {}
    for name in names:
        nel, sig = rad[name]
        out[name] = sig * np.sqrt(np.log(nel / PRECISION)) / dr
    return out


def rescale(coords, dims):
    """In-place per-frame rescale to the mean box; returns the mean box (dens.py:46-62)."""
    avg = np.average(dims, axis=0)
    factor = avg / dims
    for it in range(coords.shape[0]):
        for i in range(3):
            coords[it, :, i] *= factor[it, i]
    return coords, avg


def grid_shape(L, sres, better):
    """Grid points per axis (dens.py:181-189): truncate, optionally bump, make even."""
    n = (L / sres).astype(int)
    if better:
        for i in range(3):
            if L[i] / n[i] > sres:
                n[i] += 1
    for i in range(3):
        n[i] += n[i] % 2
    return n


def wrap(r, L):
    """In-place periodic wrap (dens.py:204-221), including the buffered branch quirk."""
    zero = [0.0, 0.0, 0.0]
    if r.shape[1] < BUFFSIZE:
        for it in range(r.shape[0]):
            r[it, ...] = np.where(r[it, ...] < L, r[it, ...], r[it, ...] - L)
            r[it, ...] = np.where(r[it, ...] > zero, r[it, ...], r[it, ...] + L)
    else:
        lo, hi = wrap_atom_range(r.shape[0], r.shape[1])
        for it in range(r.shape[0]):
            if hi > lo:
                r[it, lo:hi, :] = np.where(r[it, lo:hi, :] < L, r[it, lo:hi, :], r[it, lo:hi, :] - L)
                r[it, lo:hi, :] = np.where(r[it, lo:hi, :] > zero, r[it, lo:hi, :], r[it, lo:hi, :] + L)
    return r


def wrap_atom_range(nframes, natoms):
    """Atom index range the reference's wrap touches (dens.py:209-221).

    Below ``BUFFSIZE`` atoms every atom is wrapped.  At or above it the loop
    bound is the *frame* count and only a chunk whose end overshoots it is
    wrapped, i.e. atoms ``[last_chunk_start, nframes)``.
    """
    if natoms < BUFFSIZE:
        return 0, natoms
    lo = hi = 0
    for imin in range(0, nframes, BUFFSIZE):
        imax = imin + BUFFSIZE
        if imax > nframes:
            lo, hi = imin, nframes
    return lo, hi


# ----------------------------------------------------------------- deposition
def home_cells(rframe, dr):
    """Integer home cell of every atom, ``ir = (r / dr).astype(int)`` (dens.py:285):
    FP64 division (``dr`` is FP64), truncation toward zero."""
    return (rframe / dr).astype(int)


def deposit_frame(rframe, typ, rad, bdict, dr, nb, ucell, shape_padded, out=None):
    """Sum of truncated Gaussians on the padded FP64 grid for one frame (dens.py:283-309).

    ``grid[i,j,k] = ((i-Nb)dr0, (j-Nb)dr1, (k-Nb)dr2)`` (dens.py:244-256) is
    formed on the window only; the values are the same FP64 products.
    Returns the padded grid and the per-atom integer home cells ``ir``.
    """
    d0 = np.zeros(shape_padded) if out is None else out
    irs = home_cells(rframe, dr)
    # padded indices per axis; the reference slices grid and d0 with the same
    # [i0:i1] (dens.py:292-308), so numpy's slice rules (negative indices
    # count from the end, bounds clamp) pick the same cells of both
    pidx = [np.arange(int(shape_padded[a])) for a in range(3)]
    for im in range(rframe.shape[0]):
        ir = irs[im]
        ax, ay, az = bdict[typ[im]].astype(int)
        x0, x1 = ir[0] - ax + nb, ir[0] + ax + nb
        y0, y1 = ir[1] - ay + nb, ir[1] + ay + nb
        z0, z1 = ir[2] - az + nb, ir[2] + az + nb
        gx = (pidx[0][x0:x1] - nb) * dr[0]
        gy = (pidx[1][y0:y1] - nb) * dr[1]
        gz = (pidx[2][z0:z1] - nb) * dr[2]
        win = np.empty((len(gx), len(gy), len(gz), 3))
        win[..., 0] = gx[:, None, None]
        win[..., 1] = gy[None, :, None]
        win[..., 2] = gz[None, None, :]
        buf = rframe[im, :] - win
        buf2 = np.einsum('ml,ijkm', ucell, buf)
        dist2 = buf2[..., 0] ** 2 + buf2[..., 1] ** 2 + buf2[..., 2] ** 2
        sig = rad[typ[im]][1]
        nel = rad[typ[im]][0]
        d0[x0:x1, y0:y1, z0:z1] += (nel / np.power(sig, 3.)) * np.exp(-dist2 / (2 * np.power(sig, 2.)))
    return d0, irs


def fold_tables(n, nb):
    """Destination / origin slice bounds per axis (dens.py:258-263)."""
    des = [[0, nb, int(n[i]) - nb, int(n[i])] for i in range(3)]
    ori = [[0, nb, nb + int(n[i]), 2 * nb + int(n[i])] for i in range(3)]
    return des, ori


def fold(d0, des, ori):
    """Fold the padded borders onto the periodic cell (dens.py:86-108), corner bug kept."""
    c = [slice(ori[a][1], ori[a][2]) for a in range(3)]
    d1 = np.copy(d0[c[0], c[1], c[2]])
    for o1 in (0, 2):
        t1 = 2 - o1
        s1 = [slice(ori[a][o1], ori[a][o1 + 1]) for a in range(3)]
        e1 = [slice(des[a][t1], des[a][t1 + 1]) for a in range(3)]
        d1[:, :, e1[2]] += d0[c[0], c[1], s1[2]]
        d1[:, e1[1], :] += d0[c[0], s1[1], c[2]]
        d1[e1[0], :, :] += d0[s1[0], c[1], c[2]]
        for o2 in (0, 2):
            t2 = 2 - o2
            s2 = [slice(ori[a][o2], ori[a][o2 + 1]) for a in range(3)]
            e2 = [slice(des[a][t2], des[a][t2 + 1]) for a in range(3)]
            d1[:, e1[1], e2[2]] += d0[c[0], s1[1], s2[2]]
            d1[e1[0], :, e2[2]] += d0[s1[0], c[1], s2[2]]
            d1[e1[0], e2[1], :] += d0[s1[0], s2[1], c[2]]
            for o3 in (0, 2):
                s3z = slice(ori[2][o3], ori[2][o3 + 1])
                # dens.py:107: the z destination reuses the *y* slot (id2), not id3
                d1[e1[0], e2[1], slice(des[2][t2], des[2][t2 + 1])] += d0[s1[0], s2[1], s3z]
    return d1


def fold_destination(p, n, nb):
    """Closed form of ``fold`` for one padded index triple -> periodic index triple.

    Per axis, side(p) is L (p < Nb), C (Nb <= p < Nb+N) or H; the standard
    destination is (p - Nb) mod N.  When no axis is C (a corner) the z
    destination is ``t_z + (Nz - Nb if side_y == L else 0)`` with ``t_z = pz``
    (z on L) or ``pz - Nb - Nz`` (z on H): ``dens.py:107`` picks the z slot from
    the y side.
    """
    side = []
    std = []
    for a in range(3):
        if p[a] < nb:
            side.append("L")
            std.append(p[a] - nb + n[a])
        elif p[a] < nb + n[a]:
            side.append("C")
            std.append(p[a] - nb)
        else:
            side.append("H")
            std.append(p[a] - nb - n[a])
    if "C" not in side:
        tz = p[2] if side[2] == "L" else p[2] - nb - n[2]
        std[2] = tz + (n[2] - nb if side[1] == "L" else 0)
    return tuple(std)


# ------------------------------------------------------------------- spectrum
def frame_power(d1):
    """|rfftn(d1)|^2 as the reference forms it (dens.py:313-316)."""
    f = np.fft.rfftn(d1)
    m = f.conjugate()
    return np.real(np.multiply(m, f, out=m))


def get_dplot(dmag):
    """Half spectrum -> centred plottable grid (dens.py:142-163)."""
    sx, sy, sz = dmag.shape
    out = np.empty([sx, sy, sz * 2 - 1])
    hx, hy = int(sx / 2), int(sy / 2)
    z1, z2 = int(sz) - 1, int(sz * 2) - 2
    out[hx:sx, hy:sy, z1:z2] = dmag[0:hx, 0:hy, 0:z1]
    out[0:hx, hy:sy, z1:z2] = dmag[hx:sx, 0:hy, 0:z1]
    out[hx:sx, 0:hy, z1:z2] = dmag[0:hx, hy:sy, 0:z1]
    out[0:hx, 0:hy, z1:z2] = dmag[hx:sx, hy:sy, 0:z1]
    out[1:sx - 1, 1:sy - 1, :z1] = out[:1:-1, :1:-1, z2:z1:-1]
    cx, cy, cz = int(out.shape[0] / 2), int(out.shape[1] / 2), int(out.shape[2] / 2)
    out[cx, cy, cz] = 1.0 / 3.0 * (out[cx + 1, cy, cz] + out[cx, cy + 1, cz] + out[cx, cy, cz + 1])
    return out[1:-1, 1:-1, 1:-1]


def dplot_index(a, b, c, n):
    """Closed form of ``get_dplot``: source index in ``sf`` for output (a, b, c).

    Returns None for the centre voxel, which is the mean of its +x, +y, +z
    neighbours.
    """
    nx, ny, nz = n
    kx, ky, kz = a + 1 - nx // 2, b + 1 - ny // 2, c + 1 - nz // 2
    if kx == 0 and ky == 0 and kz == 0:
        return None
    if kz >= 0:
        return (kx % nx, ky % ny, kz)
    return ((-kx) % nx, (-ky) % ny, -kz)


def kgrids(shape, L, sfplt):
    """``kgrid`` and ``kgridplt`` (dens.py:325-344), in the dtype arithmetic of ``L``."""
    kgrid = np.zeros((shape[0], shape[1], shape[2], 4))
    for ix in range(int(kgrid.shape[0] / 2)):
        kgrid[ix, :, :, 0] = ix * 2.0 * math.pi / L[0]
        kgrid[kgrid.shape[0] - 1 - ix, :, :, 0] = -(ix + 0.5) * 2.0 * math.pi / L[0]
    for iy in range(int(kgrid.shape[1] / 2)):
        kgrid[:, iy, :, 1] = iy * 2.0 * math.pi / L[1]
        kgrid[:, kgrid.shape[1] - 1 - iy, :, 1] = -(iy + 0.5) * 2.0 * math.pi / L[1]
    for iz in range(kgrid.shape[2]):
        kgrid[:, :, iz, 2] = iz * 2.0 * math.pi / L[2]
    kp = np.zeros((shape[0] - 2, shape[1] - 2, shape[2] * 2 - 3, 4))
    for ix in range(kp.shape[0]):
        kp[ix, :, :, 0] = (ix - kp.shape[0] / 2) * 2.0 * math.pi / L[0]
    for ix in range(kp.shape[1]):
        kp[:, ix, :, 1] = (ix - kp.shape[1] / 2) * 2.0 * math.pi / L[1]
    for ix in range(kp.shape[2]):
        kp[:, :, ix, 2] = (ix - kp.shape[2] / 2) * 2.0 * math.pi / L[2]
    kp[:, :, :, 3] = sfplt
    return kgrid, kp


# ------------------------------------------------------------------ top level
def setup(r, L, typ, rad, sres, better=True):
    """Host-side set-up of ``compute_sf`` (dens.py:179-263), mutating ``r`` in place."""
    r, L = rescale(r, L)
    n = grid_shape(L, sres, better)
    dr = np.divide(L, n)
    wrap(r, L)
    bdict = borders(rad, dr, set(typ))
    nb = int(max(max(v) for v in bdict.values()))
    return dict(r=r, L=L, N=n, dr=dr, bdict=bdict, Nb=nb)


def compute_sf(r, L, typ, out_filename, rad, ucell, sres, *, better=True,
               random_noise=0, frames=None, keep=False):
    """Restatement of ``dens.compute_sf`` (dens.py:166-346).

    Mutates ``r`` like the reference.  Writes ``out_filename + '.npz'`` with the
    reference's keys when ``out_filename`` is given; always returns a dict of the
    outputs (plus ``ir`` per frame and the folded densities when ``keep``).
    ``frames`` restricts the frame loop to the given indices (for timing a
    bounded sample); the default is every frame, as in the reference.
    """
    s = setup(r, L, typ, rad, sres, better)
    r, L, n, dr, bdict, nb = s["r"], s["L"], s["N"], s["dr"], s["bdict"], s["Nb"]
    padded = tuple(int(n[i]) + 2 * nb for i in range(3))
    des, ori = fold_tables(n, nb)
    sf = np.zeros([int(n[0]), int(n[1]), int(n[2] / 2) + 1])
    irs, d1s = [], []
    loop = range(r.shape[0]) if frames is None else frames
    for it in loop:
        if random_noise > 0:
            d1 = np.random.rand(int(n[0]), int(n[1]), int(n[2]))
        else:
            d0, ir = deposit_frame(r[it], typ, rad, bdict, dr, nb, ucell, padded)
            d1 = fold(d0, des, ori)
            if keep:
                irs.append(ir)
        if keep:
            d1s.append(np.array(d1))
        sf += frame_power(d1)
    sfplt = get_dplot(sf)
    kgrid, kgridplt = kgrids(sf.shape, L, sfplt)
    res = dict(sf=sf, sfplt=sfplt, L=L, N=n, kgrid=kgrid, kgridplt=kgridplt)
    if out_filename:
        np.savez_compressed(out_filename, **res)
    if keep:
        res.update(ir=np.asarray(irs), d1=np.asarray(d1s), Nb=nb, dr=dr, bdict=bdict)
    return res
