"""Differentiable FP64 torch restatement of the S(q) forward path (test infrastructure).

Used only by the tests, as the autograd reference for the hand-written adjoint
in libmdsf.  It follows the same steps as the reference ``dens.compute_sf``:
rescale (dens.py:46-62), wrap (dens.py:206-212), home cells and windows
(dens.py:285-297), Gaussians with ``buf @ ucell`` (dens.py:299-308), the
border fold with the reference's corner z slot (dens.py:86-108), ``rfftn`` and
the frame sum of |F|^2 (dens.py:313-318).  Its forward values are pinned
against the numpy oracle in ``test_adjoint_ref.py``; its gradients are what
torch.autograd derives from them.
"""
import numpy as np
import torch

from oracle import dens_oracle as O


def prepare(dims, typ, rad, sres, better=False):
    dims = np.asarray(dims, dtype=np.float64)
    L = np.average(dims, axis=0)
    scale = L / dims
    n = O.grid_shape(L, sres, better)
    dr = np.divide(L, n)
    names = sorted(set(typ))
    bdict = O.borders(rad, dr, names)
    nb = int(max(max(v) for v in bdict.values()))
    return dict(L=L, scale=scale, N=[int(x) for x in n], dr=dr, bdict=bdict, Nb=nb)


def fold_index(u, n, nb):
    """Folded flat index of padded grid coordinates ``u = p - Nb`` (3 int64 tensors)."""
    side, std = [], []
    for a in range(3):
        s = torch.where(u[a] < 0, 0, torch.where(u[a] >= n[a], 2, 1))
        side.append(s)
        std.append(torch.remainder(u[a], n[a]))
    corner = (side[0] != 1) & (side[1] != 1) & (side[2] != 1)
    z_alt = torch.where((side[1] == 2) & (side[2] == 0), u[2] + nb,
                        torch.where((side[1] == 0) & (side[2] == 2), u[2] - nb, std[2]))
    z = torch.where(corner, z_alt, std[2])
    return (std[0] * n[1] + std[1]) * n[2] + z


def _clipped_atom(rho, r_a, ir, A, name, rad, geo, U):
    """One atom whose window leaves the padded grid: the reference slices the
    same padded range of its grid and of d0 (dens.py:292-308), so each axis
    keeps the cells numpy's slice rules keep, at their own grid coordinates."""
    n, nb, dr = geo["N"], geo["Nb"], geo["dr"]
    us = []
    for d in range(3):
        ln = n[d] + 2 * nb
        p = np.arange(ln)[int(ir[d]) - A[d] + nb:int(ir[d]) + A[d] + nb]
        if len(p) == 0:
            return rho
        us.append(torch.tensor(p - nb, dtype=torch.int64))
    ux, uy, uz = torch.meshgrid(us[0], us[1], us[2], indexing="ij")
    g = torch.stack([ux.to(torch.float64) * dr[0], uy.to(torch.float64) * dr[1],
                     uz.to(torch.float64) * dr[2]], dim=-1)
    d2 = (((r_a - g) @ U) ** 2).sum(-1)
    nel, sig = rad[name]
    val = (nel / sig ** 3) * torch.exp(-d2 / (2 * sig ** 2))
    return rho.index_add(0, fold_index((ux, uy, uz), n, nb).reshape(-1), val.reshape(-1))


def frame_density(r, typ, rad, geo, ucell):
    """Folded density (Nx, Ny, Nz) of one wrapped frame ``r`` (natoms, 3), differentiable in r."""
    n, nb, dr = geo["N"], geo["Nb"], geo["dr"]
    dr_t = torch.tensor(dr, dtype=torch.float64)
    U = torch.tensor(np.asarray(ucell, dtype=np.float64))
    rho = torch.zeros(n[0] * n[1] * n[2], dtype=torch.float64)
    typ = np.asarray(typ)
    for name in sorted(set(typ.tolist())):
        idx = np.nonzero(typ == name)[0]
        A = geo["bdict"][name].astype(int)
        ir_all = (r[idx].detach() / dr_t).to(torch.int64)   # truncation, as astype(int)
        lo = ir_all - torch.tensor(A) + nb
        hi = ir_all + torch.tensor(A) + nb
        inside = ((lo >= 0) & (hi <= torch.tensor(n) + 2 * nb)).all(dim=1).numpy()
        for k in np.nonzero(~inside)[0]:
            rho = _clipped_atom(rho, r[idx[k]], ir_all[k].numpy(), A, name, rad, geo, U)
        idx = idx[inside]
        if len(idx) == 0:
            continue
        rt = r[idx]
        ir = (rt.detach() / dr_t).to(torch.int64)          # truncation, as astype(int)
        offs = [torch.arange(-A[d], A[d], dtype=torch.int64) for d in range(3)]
        ux = ir[:, 0, None, None, None] + offs[0][None, :, None, None]
        uy = ir[:, 1, None, None, None] + offs[1][None, None, :, None]
        uz = ir[:, 2, None, None, None] + offs[2][None, None, None, :]
        ux, uy, uz = torch.broadcast_tensors(ux, uy, uz)
        g = torch.stack([ux.to(torch.float64) * dr_t[0], uy.to(torch.float64) * dr_t[1],
                         uz.to(torch.float64) * dr_t[2]], dim=-1)
        buf = rt[:, None, None, None, :] - g
        buf2 = buf @ U
        d2 = (buf2 ** 2).sum(-1)
        nel, sig = rad[name]
        val = (nel / sig ** 3) * torch.exp(-d2 / (2 * sig ** 2))
        flat = fold_index((ux, uy, uz), n, nb)
        rho = rho.index_add(0, flat.reshape(-1), val.reshape(-1))
    return rho.reshape(n)


def structure_factor(coords, dims, typ, rad, sres, ucell=None, better=False):
    """S(q) = sum over frames of |rfftn(rho)|^2 for original (unscaled) ``coords`` (F, N, 3)."""
    geo = prepare(dims, typ, rad, sres, better)
    ucell = np.eye(3) if ucell is None else ucell
    L = torch.tensor(geo["L"], dtype=torch.float64)
    scale = torch.tensor(geo["scale"], dtype=torch.float64)
    S = None
    for t in range(coords.shape[0]):
        r = coords[t] * scale[t]
        r = torch.where(r < L, r, r - L)
        r = torch.where(r > 0, r, r + L)
        f = torch.fft.rfftn(frame_density(r, typ, rad, geo, ucell))
        p = f.real ** 2 + f.imag ** 2
        S = p if S is None else S + p
    return S
