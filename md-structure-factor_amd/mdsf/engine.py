"""Host side of the S(q) path on one GPU.

``setup`` restates the host arithmetic of ``dens.compute_sf`` (mean box,
rescale factors, grid size, borders, window half-widths, wrap range) in numpy
with the reference's dtypes, and ``Plan`` drives ``libmdsf`` with torch device
memory and the current torch stream.  All per-atom and per-voxel work runs in
the HIP kernels.
"""
import numpy as np
import torch

from . import _lib

PRECISION = 1.0e-24   # dens.py:18
BUFFSIZE = 1000000    # dens.py:206


def grid_shape(L, sres, better):
    """Grid points per axis (dens.py:181-189)."""
    n = (L / sres).astype(int)
    if better:
        for i in range(3):
            if L[i] / n[i] > sres:
                n[i] += 1
    for i in range(3):
        n[i] += n[i] % 2
    return n


def borders(rad, dr, names):
    """``{name: sigma*sqrt(ln(Nel/1e-24))/dr}`` (dens.py:38-43); KeyError for unknown names."""
    return {a: rad[a][1] * np.sqrt(np.log(rad[a][0] / PRECISION)) / dr for a in names}


def wrap_range(nframes, natoms):
    """Atoms the wrap touches (dens.py:209-221): all below 1e6 atoms, else the
    chunk ``[imin, nframes)`` of the loop that iterates frame indices."""
    if natoms < BUFFSIZE:
        return 0, natoms
    lo = hi = 0
    for imin in range(0, nframes, BUFFSIZE):
        if imin + BUFFSIZE > nframes:
            lo, hi = imin, nframes
    return lo, hi


class Setup:
    """Everything ``compute_sf`` derives on the host before the frame loop."""

    def __init__(self, r, dims, typ, rad, sres, better, ucell):
        if r.ndim != 3 or r.shape[2] != 3:
            raise ValueError("coordinates must have shape (frames, atoms, 3)")
        if r.dtype not in (np.float32, np.float64):
            raise TypeError("coordinates must be float32 or float64 (in-place rescale)")
        dims = np.asarray(dims)
        self.L = np.average(dims, axis=0)                     # dens.py:52
        self.scale = self.L / dims                            # dens.py:53
        self.N = grid_shape(self.L, sres, better)             # dens.py:181-189
        self.dr = np.divide(self.L, self.N)                   # dens.py:202
        typ = np.asarray(typ)
        self.names, self.name_index = np.unique(typ, return_inverse=True)
        self.names = [str(x) for x in self.names]
        self.name_index = self.name_index.reshape(-1)
        self.bdict = borders(rad, self.dr, self.names)        # dens.py:223
        if len(self.bdict) == 0:
            raise ValueError("no atoms")
        self.nborder = int(max(max(v) for v in self.bdict.values()))  # dens.py:226-231
        # The density is one sum over atoms (dens.py:283-308): an atom enters it only
        # through its amplitude Nel/sigma^3, its 2 sigma^2 and its window half-widths,
        # so names whose three agree exactly deposit as one device type.  Force-field
        # systems name every atom of a residue apart (100+ names) with a handful of
        # distinct parameter sets.
        keys, self.type_of_name, self.type_names = {}, [], []
        for a in self.names:
            key = (float(rad[a][0] / np.power(rad[a][1], 3.)), float(2 * np.power(rad[a][1], 2.)),
                   tuple(int(h) for h in self.bdict[a].astype(int)))
            if key not in keys:
                keys[key] = len(keys)
                self.type_names.append([])
            self.type_of_name.append(keys[key])
            self.type_names[keys[key]].append(a)
        self.type_of_name = np.asarray(self.type_of_name, dtype=np.int64)
        self.type_index = self.type_of_name[self.name_index]
        self.amp = np.array([k[0] for k in keys])
        self.two_sig2 = np.array([k[1] for k in keys])
        self.half = np.array([k[2] for k in keys], dtype=np.int64)
        self.counts = np.bincount(self.type_index, minlength=len(keys))
        self.nframes, self.natoms = r.shape[0], r.shape[1]
        self.wrap = wrap_range(self.nframes, self.natoms)
        # numpy rescales / wraps in float32 only when both coords and dims are float32
        self.arith = _lib.F32 if (r.dtype == np.float32 and self.scale.dtype == np.float32) else _lib.F64
        self.dtype = _lib.F32 if r.dtype == np.float32 else _lib.F64
        self.box_f32 = self.L.dtype == np.float32
        self.ucell = np.asarray(ucell, dtype=np.float64).reshape(3, 3)
        for d in range(3):
            if self.nborder > self.N[d]:
                raise ValueError(
                    f"Nborder={self.nborder} exceeds the grid N={list(self.N)}: the border fold "
                    "(dens.py:86-108) needs Nborder <= N; use a finer Sres")


def _device(device):
    """``None`` (current device), an index or a ``torch.device`` -> torch.device('cuda', i)."""
    if device is None:
        return torch.device("cuda", torch.cuda.current_device())
    if isinstance(device, torch.device):
        return torch.device("cuda", device.index if device.index is not None
                            else torch.cuda.current_device())
    return torch.device("cuda", int(device))


def grid_desc(n, nborder, dr, ucell, amp, two_sig2, half, counts, flags=0, desc=None):
    """Fill an ``mdsf_grid_desc``; returns (desc, type array) -- keep both alive."""
    ntypes = len(amp)
    tarr = (_lib.TypeDesc * max(ntypes, 1))()
    for t in range(ntypes):
        tarr[t].amp = float(amp[t])
        tarr[t].two_sig2 = float(two_sig2[t])
        for d in range(3):
            tarr[t].half[d] = int(half[t][d])
        tarr[t].count = int(counts[t])
    desc = _lib.GridDesc() if desc is None else desc
    for d in range(3):
        desc.n[d] = int(n[d])
        desc.dr[d] = float(dr[d])
    desc.nborder = int(nborder)
    u = np.asarray(ucell, dtype=np.float64).reshape(9)
    for i in range(9):
        desc.ucell[i] = float(u[i])
    desc.ntypes = ntypes
    desc.types = tarr
    desc.natoms = int(np.sum(counts))
    desc.flags = int(flags)
    return desc, tarr


def setup_desc(s, flags=0, desc=None):
    return grid_desc(s.N, s.nborder, s.dr, s.ucell, s.amp, s.two_sig2, s.half, s.counts, flags,
                     desc)


class Plan:
    """One system (grid, border, cell, types) bound to a libmdsf plan on a GPU.

    ``fft64`` selects FP64 deposition and FFT; ``autotune`` times the FFT tile
    and kernel choices for this grid when the plan is created.
    """

    def __init__(self, n, nborder, dr, ucell, amp, two_sig2, half, counts, device=None,
                 fft64=False, autotune=False, _handle=None):
        self.lib = _lib.load()
        if not torch.cuda.is_available():
            raise RuntimeError("libmdsf needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.device = _device(device)
        self.n = [int(x) for x in n]
        self.natoms = int(np.sum(counts))
        self.fft64 = bool(fft64)
        self._owned = _handle is None
        if _handle is not None:
            self._handle = _lib.c_void_p(_handle)
            return
        flags = (_lib.FFT_FP64 if fft64 else 0) | (_lib.AUTOTUNE if autotune else 0)
        desc, self._tarr = grid_desc(n, nborder, dr, ucell, amp, two_sig2, half, counts, flags)
        self._handle = _lib.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(self.lib.mdsf_plan_create(desc, self._handle))

    @classmethod
    def from_setup(cls, s, device=None, fft64=False, autotune=False):
        return cls(s.N, s.nborder, s.dr, s.ucell, s.amp, s.two_sig2, s.half, s.counts, device,
                   fft64, autotune)

    # ----------------------------------------------------------------- utils
    def _stream(self):
        return _lib.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    @staticmethod
    def _ptr(t):
        return None if t is None else _lib.c_void_p(t.data_ptr())

    def close(self):
        if getattr(self, "_handle", None) is not None and self._handle.value:
            if getattr(self, "_owned", True):
                self.lib.mdsf_plan_destroy(self._handle)
            self._handle = _lib.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def half_shape(self):
        return (self.n[0], self.n[1], self.n[2] // 2 + 1)

    # ------------------------------------------------------------- pipeline
    def reset(self):
        _lib.check(self.lib.mdsf_sf_reset(self._handle, self._stream()))

    def check_errors(self):
        _lib.check(self.lib.mdsf_check_errors(self._handle, self._stream()))

    def accumulate_frames(self, coords, types, scale, box, arith, wrap, occupancy=None,
                          home=None):
        """Frame loop (dens.py:277-321) over device ``coords`` (F, natoms, 3), in place."""
        dtype = _lib.F32 if coords.dtype == torch.float32 else _lib.F64
        _lib.check(self.lib.mdsf_accumulate_frames(
            self._handle, self._ptr(coords), dtype, coords.shape[0], self._ptr(types),
            self._ptr(scale), _lib.doubles(box), arith, int(wrap[0]), int(wrap[1]),
            self._ptr(occupancy), self._ptr(home), self._stream()))

    def rescale_wrap(self, coords, scale, box, arith, wrap):
        dtype = _lib.F32 if coords.dtype == torch.float32 else _lib.F64
        _lib.check(self.lib.mdsf_rescale_wrap(
            self._handle, self._ptr(coords), dtype, coords.shape[0], self._ptr(scale),
            _lib.doubles(box), arith, int(wrap[0]), int(wrap[1]), self._stream()))

    @property
    def real_dtype(self):
        return torch.float64 if self.fft64 else torch.float32

    def deposit_fold(self, coords, types, scale, box, arith, wrap, home=None):
        """One frame (dens.py:283-311): folded density (Nx, Ny, Nz) in plan precision."""
        out = torch.empty(self.n, dtype=self.real_dtype, device=self.device)
        dtype = _lib.F32 if coords.dtype == torch.float32 else _lib.F64
        _lib.check(self.lib.mdsf_deposit_fold(
            self._handle, self._ptr(coords), dtype, self._ptr(types), self._ptr(scale),
            _lib.doubles(box), arith, int(wrap[0]), int(wrap[1]), self._ptr(out),
            self._ptr(home), self._stream()))
        return out

    def prepare_frame(self, coords, types, scale, box, arith, wrap, occupancy=None, home=None):
        """Rescale/wrap one frame in place and bin its atoms (first launches of a frame)."""
        dtype = _lib.F32 if coords.dtype == torch.float32 else _lib.F64
        _lib.check(self.lib.mdsf_prepare_frame(
            self._handle, self._ptr(coords), dtype, self._ptr(types), self._ptr(scale),
            _lib.doubles(box), arith, int(wrap[0]), int(wrap[1]), self._ptr(occupancy),
            self._ptr(home), self._stream()))

    def set_stage_timing(self, enable=True):
        """True / 1: events around every stage; 2: around the deposition only; False: off."""
        mode = int(enable) if not isinstance(enable, bool) else (1 if enable else 0)
        _lib.check(self.lib.mdsf_set_stage_timing(self._handle, mode))

    def stage_times(self):
        """Average ms per frame of the frames accumulate_frames timed since the
        last call (timed frames accumulate over calls):
        {prepare, deposit, fft_z, fft_y, fft_x_acc}."""
        out = (_lib.c_double * 5)()
        _lib.check(self.lib.mdsf_stage_times(self._handle, out))
        return dict(zip(["prepare", "deposit", "fft_z", "fft_y", "fft_x_acc"], list(out)))

    def set_frame_buffer(self, which):
        """Select density / spectrum buffer 0, 1 or 2 for the single-frame entry points."""
        _lib.check(self.lib.mdsf_set_frame_buffer(self._handle, int(which)))

    def deposit(self, out=None):
        _lib.check(self.lib.mdsf_deposit(self._handle, self._ptr(out), self._stream()))

    def fft_pass(self, which, density=None):
        _lib.check(self.lib.mdsf_fft_pass(self._handle, int(which), self._ptr(density),
                                          self._stream()))

    def accumulate_density(self, density):
        d = density.to(device=self.device, dtype=self.real_dtype).contiguous()
        _lib.check(self.lib.mdsf_accumulate_density(self._handle, self._ptr(d), self._stream()))
        return d

    def rfft3(self, density):
        """np.fft.rfftn of a density in the plan's precision (complex128 out)."""
        d = density.to(device=self.device, dtype=self.real_dtype).contiguous()
        out = torch.empty(self.half_shape + (2,), dtype=torch.float64, device=self.device)
        _lib.check(self.lib.mdsf_rfft3(self._handle, self._ptr(d), self._ptr(out), self._stream()))
        return torch.view_as_complex(out)

    # -------------------------------------------------------------- adjoint
    def adjoint_frames(self, coords, types, scale, box, arith, wrap, dsf, grad=None):
        """dL/d(coords) for device ``coords`` (F, natoms, 3) given ``dsf`` = dL/dS
        (Nx, Ny, Nz/2+1) FP64.  ``coords`` is rescaled and wrapped in place (pass a
        copy); returns FP64 (F, natoms, 3).  Each frame is recomputed."""
        dtype = _lib.F32 if coords.dtype == torch.float32 else _lib.F64
        g = dsf.to(device=self.device, dtype=torch.float64).contiguous()
        if tuple(g.shape) != self.half_shape:
            raise ValueError(f"dL/dS must have shape {self.half_shape}")
        if grad is None:
            grad = torch.empty(tuple(coords.shape), dtype=torch.float64, device=self.device)
        _lib.check(self.lib.mdsf_adjoint_frames(
            self._handle, self._ptr(coords), dtype, coords.shape[0], self._ptr(types),
            self._ptr(scale), _lib.doubles(box), arith, int(wrap[0]), int(wrap[1]),
            self._ptr(g), self._ptr(grad), self._stream()))
        return grad

    def density_adjoint(self, density, dsf):
        """dL/d(density) for L(|rfftn(density)|^2) given dL/dS (plan precision)."""
        d = density.to(device=self.device, dtype=self.real_dtype).contiguous()
        g = dsf.to(device=self.device, dtype=torch.float64).contiguous()
        if tuple(g.shape) != self.half_shape:
            raise ValueError(f"dL/dS must have shape {self.half_shape}")
        out = torch.empty(self.n, dtype=self.real_dtype, device=self.device)
        _lib.check(self.lib.mdsf_density_adjoint(self._handle, self._ptr(d), self._ptr(g),
                                                 self._ptr(out), self._stream()))
        return out

    def deposit_adjoint(self, grad_density):
        """dL/dr (natoms, 3) FP64 of the last prepared frame from dL/d(folded density)."""
        gd = grad_density.to(device=self.device, dtype=self.real_dtype).contiguous()
        out = torch.empty((self.natoms, 3), dtype=torch.float64, device=self.device)
        _lib.check(self.lib.mdsf_deposit_adjoint(self._handle, self._ptr(gd), self._ptr(out),
                                                 self._stream()))
        return out

    def sf(self):
        out = torch.empty(self.half_shape, dtype=torch.float64, device=self.device)
        _lib.check(self.lib.mdsf_sf_read(self._handle, self._ptr(out), self._stream()))
        return out

    def outputs(self, sf, box, box_f32):
        """``sfplt``, ``kgrid``, ``kgridplt`` (dens.py:323-344) on the device."""
        nx, ny, nz = self.n
        sfplt = torch.empty((nx - 2, ny - 2, nz - 1), dtype=torch.float64, device=self.device)
        kgrid = torch.empty((nx, ny, nz // 2 + 1, 4), dtype=torch.float64, device=self.device)
        kgridplt = torch.empty((nx - 2, ny - 2, nz - 1, 4), dtype=torch.float64,
                               device=self.device)
        _lib.check(self.lib.mdsf_outputs(
            self._handle, self._ptr(sf), _lib.doubles(box), 1 if box_f32 else 0,
            self._ptr(sfplt), self._ptr(kgrid), self._ptr(kgridplt), self._stream()))
        return sfplt, kgrid, kgridplt


class Batch:
    """Independent systems (each a ``Setup``) run together from one HIP graph.

    Every system gets an autotuned FP32 plan.  ``set_frames`` binds device
    coordinates (never modified by the batch); ``run`` enqueues all systems'
    frame loops as one graph launch (captured on the first run after the
    frames change); ``sf(i)`` reads system i's S(q).
    """

    def __init__(self, setups, device=None):
        self.lib = _lib.load()
        if not torch.cuda.is_available():
            raise RuntimeError("libmdsf needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.device = _device(device)
        self.setups = list(setups)
        descs = (_lib.GridDesc * len(self.setups))()
        keep = []
        for i, s in enumerate(self.setups):
            _, tarr = setup_desc(s, 0, descs[i])
            keep.append(tarr)
        self._handle = _lib.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(self.lib.mdsf_batch_create(descs, len(self.setups), self._handle))
        self.plans = [Plan(s.N, 0, s.dr, s.ucell, [], [], [], s.counts, self.device,
                           _handle=self.lib.mdsf_batch_plan(self._handle, i))
                      for i, s in enumerate(self.setups)]
        self._bound = [None] * len(self.setups)

    def _stream(self):
        return _lib.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def set_frames(self, i, coords, types, scale):
        s = self.setups[i]
        dtype = _lib.F32 if coords.dtype == torch.float32 else _lib.F64
        self._bound[i] = (coords, types, scale)   # keep the device buffers alive
        _lib.check(self.lib.mdsf_batch_set_frames(
            self._handle, i, _lib.c_void_p(coords.data_ptr()), dtype, coords.shape[0],
            _lib.c_void_p(types.data_ptr()), _lib.c_void_p(scale.data_ptr()),
            _lib.doubles([float(x) for x in s.L]), s.arith, int(s.wrap[0]), int(s.wrap[1])))

    def run(self):
        _lib.check(self.lib.mdsf_batch_run(self._handle, self._stream()))

    def check_errors(self):
        _lib.check(self.lib.mdsf_batch_check_errors(self._handle, self._stream()))

    def sf(self, i):
        return self.plans[i].sf()

    def close(self):
        if getattr(self, "_handle", None) is not None and self._handle.value:
            for p in self.plans:
                p._handle = _lib.c_void_p()
            self.lib.mdsf_batch_destroy(self._handle)
            self._handle = _lib.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rescale_device(coords, scale, arith):
    """In-place rescale (dens.py:54-58) of device coordinates (F, natoms, 3) by
    device per-frame factors ``scale`` (F, 3) FP64, in numpy's arithmetic ``arith``."""
    lib = _lib.load()
    dtype = _lib.F32 if coords.dtype == torch.float32 else _lib.F64
    s = _lib.c_void_p(torch.cuda.current_stream(coords.device).cuda_stream)
    _lib.check(lib.mdsf_rescale(_lib.c_void_p(coords.data_ptr()), dtype, int(coords.shape[0]),
                                int(coords.shape[1]), _lib.c_void_p(scale.data_ptr()), int(arith),
                                s))


def fold_padded(d0, n, nborder):
    """Fold of a padded FP64 device grid (n + 2 nborder per axis) onto the
    periodic cell n (remap_grid_tcl, dens.py:86-108, corner z slot included)."""
    lib = _lib.load()
    d = d0.to(dtype=torch.float64).contiguous()
    n = [int(x) for x in n]
    nb = [int(x) for x in nborder]
    if list(d.shape) != [n[i] + 2 * nb[i] for i in range(3)]:
        raise ValueError(f"padded grid shape {tuple(d.shape)} does not match N={n}, "
                         f"Nborder={nb}")
    out = torch.empty(n, dtype=torch.float64, device=d.device)
    s = _lib.c_void_p(torch.cuda.current_stream(d.device).cuda_stream)
    _lib.check(lib.mdsf_fold_padded(_lib.c_void_p(d.data_ptr()), (_lib.c_int32 * 3)(*n),
                                    (_lib.c_int32 * 3)(*nb), _lib.c_void_p(out.data_ptr()), s))
    return out


def ewald_slices(kgridplt, wavelength):
    """Ewald-sphere sections (plot2d.py:218-263) of a device kgridplt (nx, ny, nz, 4)."""
    lib = _lib.load()
    kp = kgridplt.to(dtype=torch.float64).contiguous()
    nx, ny, nz = (int(x) for x in kp.shape[:3])
    xy = torch.empty((nx, ny), dtype=torch.float64, device=kp.device)
    xz = torch.empty((nx, nz), dtype=torch.float64, device=kp.device)
    yz = torch.empty((ny, nz), dtype=torch.float64, device=kp.device)
    s = _lib.c_void_p(torch.cuda.current_stream(kp.device).cuda_stream)
    _lib.check(lib.mdsf_ewald_slices(_lib.c_void_p(kp.data_ptr()), nx, ny, nz, float(wavelength),
                                     _lib.c_void_p(xy.data_ptr()), _lib.c_void_p(xz.data_ptr()),
                                     _lib.c_void_p(yz.data_ptr()), s))
    return xy, xz, yz


# --- cylindrical (r, z) average: PLOT_RAD_NEW's numbers (plot2d.py:559-650) ---
THETA_BINS_PER_INV_ANG = 20.0   # plot2d.py:590
MIN_THETA_BINS = 1              # plot2d.py:591
RBINS = 400                     # plot2d.py:592


def reciprocal_inverse(ucell):
    """``b_inv`` (plot2d.py:595-603): inverse of the stacked reciprocal vectors
    b_i = (a_j x a_k) / (a_i . (a_j x a_k)), no factor 2 pi."""
    u = np.asarray(ucell, dtype=np.float64).reshape(3, 3)
    a1, a2, a3 = u[0], u[1], u[2]
    b1 = np.cross(a2, a3) / np.dot(a1, np.cross(a2, a3))
    b2 = np.cross(a3, a1) / np.dot(a2, np.cross(a3, a1))
    b3 = np.cross(a1, a2) / np.dot(a3, np.cross(a1, a2))
    return np.linalg.inv(np.vstack((b1, b2, b3)))


def rz_bins(X, Y, Z, ucell, rbins=RBINS):
    """r and z bin centres and theta samples per r bin (plot2d.py:605-623)."""
    u = np.asarray(ucell, dtype=np.float64).reshape(3, 3)
    zbins = Z.shape[0]
    xr = (X[-1] - X[0]) * u[0][0]
    yr = (Y[-1] - Y[0]) * u[1][1]
    rmax = min(xr, yr) / 2.0
    rmax *= 0.95
    rarr = np.linspace(0.0, rmax, rbins)
    zar = np.linspace(Z[0], Z[-1], zbins)
    circ = 2. * np.pi * rarr
    nth = np.array([max(int(THETA_BINS_PER_INV_ANG * c), MIN_THETA_BINS) for c in circ],
                   dtype=np.int32)
    return rarr, zar, nth


def alkane_intensity(rfin, z, final, inner=1.4, outer=1.57, angle=120):
    """``normalize_alkanes`` (plot2d.py:874-918) on the device: mean intensity of
    the (r, z) points with inner < |(r, z)| < outer whose angle atan(z / r)
    (degrees) falls in the 90-bin histogram bins [start, nbins - start),
    start = int((angle / 2) / (180 / nbins))."""
    nbins = 90
    dev = final.device
    bins = torch.from_numpy(np.linspace(-90, 90, nbins)).to(dev)
    R = torch.from_numpy(np.asarray(rfin, dtype=np.float64)).to(dev)[:, None]
    Zt = torch.from_numpy(np.asarray(z, dtype=np.float64)).to(dev)[None, :]
    nrm = torch.sqrt(R * R + Zt * Zt)
    ring = (inner < nrm) & (nrm < outer)
    ang = (180 / np.pi) * torch.atan(Zt / R)                  # r = 0: +-90 degrees
    idx = torch.bucketize(ang, bins, right=True) - 1          # np.digitize(...) - 1
    start = int((angle / 2) / (180 / nbins))
    end = nbins - start
    sel = ring & (idx >= start) & (idx < end)
    total = torch.sum(torch.where(sel, final, torch.zeros_like(final)))
    count = torch.sum(sel)
    return total / count


def rz_average(kgridplt, ucell, format=False, rbins=RBINS):
    """PLOT_RAD_NEW's numbers (plot2d.py:575-650) for a device kgridplt (nx, ny, nz, 4):
    the theta average of S over cylinders about q_z (HIP kernel), NaN bins set to
    the smallest finite value, normalised by the mean (``format=False``) or by the
    alkane-ring intensity (``format=True``), and mirrored to negative r.

    Returns a dict of ``oa`` (rbins, nz) and ``final`` (2 rbins - 1, nz) device
    tensors and the ``rarr``, ``zar``, ``rfin``, ``zfin`` (1, nz) axes (numpy)."""
    lib = _lib.load()
    kp = kgridplt.to(dtype=torch.float64).contiguous()
    nx, ny, nz = (int(x) for x in kp.shape[:3])
    X = kp[:, 0, 0, 0].cpu().numpy()
    Y = kp[0, :, 0, 1].cpu().numpy()
    Z = kp[0, 0, :, 2].cpu().numpy()
    for ax in (X, Y, Z):
        if not np.all(np.diff(ax) > 0):
            raise ValueError("kgridplt axes must be strictly ascending")
    rarr, zar, nth = rz_bins(X, Y, Z, ucell, rbins)
    binv = np.ascontiguousarray(reciprocal_inverse(ucell), dtype=np.float64)
    dev = kp.device
    d_r = torch.from_numpy(rarr).to(dev)
    d_z = torch.from_numpy(zar).to(dev)
    d_n = torch.from_numpy(nth).to(dev)
    oa = torch.empty((len(rarr), len(zar)), dtype=torch.float64, device=dev)
    s = _lib.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    _lib.check(lib.mdsf_rz_average(_lib.c_void_p(kp.data_ptr()), nx, ny, nz,
                                   _lib.c_void_p(d_r.data_ptr()), _lib.c_void_p(d_n.data_ptr()),
                                   len(rarr), _lib.c_void_p(d_z.data_ptr()), len(zar),
                                   binv.ctypes.data_as(_lib.ctypes.POINTER(_lib.c_double)),
                                   _lib.c_void_p(oa.data_ptr()), s))
    nan = torch.isnan(oa)
    finite = oa[~nan]
    mn = finite.min() if finite.numel() else torch.tensor(float("nan"), dtype=oa.dtype, device=dev)
    oa = torch.where(nan, mn, oa)                                     # plot2d.py:629-630
    if not format:
        oa = oa / torch.mean(oa)                                      # plot2d.py:632-634
    final = torch.cat([torch.flip(oa, [0]), oa[1:]], 0)              # plot2d.py:637
    rfin = np.append(-rarr[::-1], rarr[1:])                          # plot2d.py:638
    zfin = zar[None, :].copy()                                        # plot2d.py:639
    if format:
        final = final / alkane_intensity(rfin, zfin[0], final, 1.4, 1.57, 120)  # plot2d.py:653-655
    return {"oa": oa, "final": final, "rarr": rarr, "zar": zar, "rfin": rfin, "zfin": zfin}


def interp_points(kgridplt, pts):
    """Trilinear interpolant of kgridplt's S at device points (n, 3), NaN outside
    (RegularGridInterpolator(..., bounds_error=False), plot2d.py:588)."""
    lib = _lib.load()
    kp = kgridplt.to(dtype=torch.float64).contiguous()
    nx, ny, nz = (int(x) for x in kp.shape[:3])
    p = pts.to(device=kp.device, dtype=torch.float64).reshape(-1, 3).contiguous()
    out = torch.empty(p.shape[0], dtype=torch.float64, device=kp.device)
    s = _lib.c_void_p(torch.cuda.current_stream(kp.device).cuda_stream)
    _lib.check(lib.mdsf_interp_points(_lib.c_void_p(kp.data_ptr()), nx, ny, nz,
                                      _lib.c_void_p(p.data_ptr()), int(p.shape[0]),
                                      _lib.c_void_p(out.data_ptr()), s))
    return out.reshape(pts.shape[:-1])
