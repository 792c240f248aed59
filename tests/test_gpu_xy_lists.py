"""xy-coupled cells (the reference driver's default -ct 120 cell and other
monoclinic angles) on the FP32 tile lists, against the oracle.

An FP32 plan whose cell couples x with y (dens.py:301 with
ucell = [[1,0,0],[cos g, sin g, 0],[0,0,1]]) deposits from the per-tile atom
lists when every type's window meets each x and each y tile through one
periodic image (2A <= N - 16) and the cross exponent stays bounded; the
atom's factor over a tile is then X(x) 2^(LY(y) + CX(x) BY(y)).  Otherwise
the enumerating xy kernel runs (the small grids of test_gpu_parity.py and the
mono120 golden).  `xy_lists_taken` restates the plan's choice so that each
test states which kernel it reaches.

Tolerances as in test_gpu_parity.py: densities 1e-5 relative + 2e-6 of the
maximum, S(q) 1e-5 relative above 1e-6 max(S); a second deposition of the
same frame bitwise equal.
"""
import math

import numpy as np
import pytest
import torch

from conftest import load_golden
from oracle import dens_oracle as O
from test_gpu_coverage import _density_check, _many_types_system, _monoclinic, _sq_dropin
from test_gpu_parity import _corner_case, _gpu, assert_sq_close

pytestmark = pytest.mark.gpu

LOG2E = 1.4426950408889634
SRES = 0.3   # N = 120 x 134 x 146, Nborder 17


def xy_lists_taken(s, fp64=False):
    """The plan's choice (csrc/capi.cpp, plan creation) for an engine.Setup."""
    U = np.asarray(s.ucell, dtype=np.float64).reshape(3, 3)
    G = U @ U.T
    if fp64 or G[0, 1] == 0.0 or G[0, 2] != 0.0 or G[1, 2] != 0.0:
        return False
    for t in range(len(s.half)):
        if any(2 * int(s.half[t][d]) > int(s.N[d]) - 16 for d in range(2)):
            return False
        c = 2.0 * abs(G[0, 1]) * LOG2E / s.two_sig2[t]
        if c * (s.half[t][0] + 1) * s.dr[0] * (s.half[t][1] + 1) * s.dr[1] > 100.0:
            return False
    return True


def _xy_system(seed, frames=2, box=(36.0, 40.0, 44.0), n=420):
    """H / C / O / N uniform in the cell, plus two atoms within 1 A of each of
    the 8 corners (corner atoms: their windows leave the cell on all three
    axes, and with Nborder > 16 their corner-slot copies spill into z tiles
    the window does not reach)."""
    rng = np.random.default_rng(seed)
    box = np.asarray(box)
    names = ["H", "C", "O", "N"]
    base = rng.uniform(0, 1, (n, 3)) * box
    corners = [np.array([(k >> 2) & 1, (k >> 1) & 1, k & 1]) * box + rng.uniform(-1, 1, 3)
               for k in range(8) for _ in range(2)]
    base = np.mod(np.vstack([base, corners]), box)
    typ = [names[i % 4] for i in range(base.shape[0])]
    coords = np.stack([np.mod(base + rng.normal(0, 0.05, base.shape), box)
                       for _ in range(frames)]).astype(np.float32)
    return coords, np.tile(box, (frames, 1)), typ


@pytest.mark.parametrize("theta", [120.0, 60.0, 100.0])
def test_xy_lists_density_and_sq(theta, rad, gpu, tmp_path):
    from mdsf import engine
    coords, dims, typ = _xy_system(61)
    ucell = _monoclinic(theta)
    s = engine.Setup(np.array(coords), dims, typ, rad, SRES, False, ucell)
    assert xy_lists_taken(s), "the system should reach the xy list kernel"
    assert s.nborder > 16
    print(f"xy lists theta={theta}: N={list(s.N)} Nborder={s.nborder}", flush=True)
    _density_check(coords, dims, typ, rad, SRES, ucell, False, frames=(0, 1))
    want = O.compute_sf(np.array(coords), np.array(dims), typ, None, rad, ucell, SRES,
                        better=False)["sf"]
    _, got = _sq_dropin(coords, dims, typ, rad, SRES, ucell, "fp32", tmp_path)
    err = assert_sq_close(got["sf"], want)
    print(f"  S(q) max rel {err.max():.2e} median {np.median(err):.2e}", flush=True)


def test_xy_lists_many_parameter_sets(rad, gpu, tmp_path):
    """40 parameter sets (the list kernel's per-type layout sized for 64
    types) in a 60 A cell at Sres 0.6 (N = 100, windows up to 56 cells, corner
    atoms): density of two frames and S(q) against the oracle."""
    from mdsf import engine
    coords, dims, typ = _many_types_system(33, frames=2, box=60.0)
    ucell = _monoclinic(120.0)
    s = engine.Setup(coords, dims, typ, rad, 0.6, False, ucell)
    assert len(s.amp) == 40 and xy_lists_taken(s)
    _density_check(coords, dims, typ, rad, 0.6, ucell, False, frames=(0, 1))
    want = O.compute_sf(np.array(coords), np.array(dims), typ, None, rad, ucell, 0.6,
                        better=False)["sf"]
    _, got = _sq_dropin(coords, dims, typ, rad, 0.6, ucell, "fp32", tmp_path)
    assert_sq_close(got["sf"], want)


def test_xy_lists_adjoint_frames(rad, gpu):
    """Forward (the xy list deposition) and backward (the gather adjoint of
    xy-coupled cells) over two frames of a 120-degree system on the lists,
    against autograd of the torch reference (tolerances of
    test_gpu_adjoint.py: S(q) 1e-4 relative, gradient 2e-4 of its maximum)."""
    import sf_torch_ref as T
    from mdsf.autograd import SFModel
    coords, dims, typ = _xy_system(63, frames=2, n=200)
    coords = coords.astype(np.float64)
    ucell = _monoclinic(120.0)
    model = SFModel(coords, dims, typ, rad, SRES, ucell=ucell, fp64=False)
    from mdsf import engine
    assert xy_lists_taken(engine.Setup(np.array(coords), dims, typ, rad, SRES, False, ucell))
    rng = np.random.default_rng(4)
    x = torch.tensor(coords, device=model.plan.device, requires_grad=True)
    S = model(x)
    w = torch.tensor(rng.normal(size=tuple(S.shape)), device=S.device) / S.detach().max()
    torch.sum(w * S).backward()
    got = x.grad.cpu().numpy()
    xr = torch.tensor(coords, requires_grad=True)
    Sr = T.structure_factor(xr, dims, typ, rad, SRES, ucell)
    np.testing.assert_allclose(S.detach().cpu().numpy(), Sr.detach().numpy(), rtol=1e-4,
                               atol=1e-6 * float(Sr.max()))
    torch.sum(w.cpu() * Sr).backward()
    want = xr.grad.numpy()
    err = float(np.max(np.abs(got - want)) / np.max(np.abs(want)))
    assert err <= 2e-4, err


def test_xy_lists_reproducible(rad, gpu):
    coords, dims, typ = _xy_system(62, frames=1)
    args = (coords, dims, typ, rad, SRES, False, _monoclinic(120.0))
    s, plan, c, types, scale = _gpu(*args)
    assert xy_lists_taken(s)
    box = [float(x) for x in s.L]
    c0 = c[0].clone()
    d1 = plan.deposit_fold(c[0], types, scale[0], box, s.arith, s.wrap).cpu().numpy()
    d2 = plan.deposit_fold(c0, types, scale[0], box, s.arith, s.wrap).cpu().numpy()
    plan.check_errors()
    plan.close()
    assert np.array_equal(d1, d2)


def test_xy_kernel_choice(rad):
    """Both xy kernels are reached, with 16 and with 64 types: the small-grid
    parity cases and the 40-type system of test_gpu_coverage.py keep the
    enumerating kernel (a window wider than N - 16 cells), FP64 plans never
    take lists, and the 120-degree HII workload of the bench takes them."""
    from mdsf import engine, synth
    g = load_golden("mono120_f32")
    s = engine.Setup(np.array(g["in_coords"]), g["in_dims"], g["typ"], rad, g["sres"],
                     bool(g["in_better"]), g["in_ucell"])
    assert not xy_lists_taken(s)
    rng = np.random.default_rng(11)
    r, dims, typ, ucell = _corner_case(rng, np.array([7.0, 8.0, 9.0]), 2, ["C", "O", "H", "N"],
                                       120.0)
    assert not xy_lists_taken(engine.Setup(r, dims, typ, rad, 0.3, True, ucell))
    c40, d40, t40 = _many_types_system(31)
    assert not xy_lists_taken(engine.Setup(c40, d40, t40, rad, 0.6, False, _monoclinic(120.0)))
    coords, dims, typ = _xy_system(61, frames=1)
    s = engine.Setup(np.array(coords), dims, typ, rad, SRES, False, _monoclinic(120.0))
    assert xy_lists_taken(s) and not xy_lists_taken(s, fp64=True)
    c, d, t = synth.hii(seed=2002, nframes=1)
    th = math.radians(120.0)
    c = np.array(c, copy=True)
    c[..., 1] = c[..., 1] / np.sin(th)
    c[..., 0] = c[..., 0] - c[..., 1] * np.cos(th)
    assert xy_lists_taken(engine.Setup(c, d, t, rad, synth.hii_sres(), False, _monoclinic(120.0)))
