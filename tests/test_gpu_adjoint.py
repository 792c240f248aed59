"""GPU adjoint (dL/dS -> dL/dx) against torch autograd and finite differences.

References:
  * ``sf_torch_ref`` (FP64, CPU), whose forward is pinned to the oracle in
    ``test_adjoint_ref.py``: torch.autograd of it is the gradient reference;
  * central finite differences of the FP64 GPU forward (itself pinned to the
    oracle by ``test_gpu_parity.py``) at the adjoint workload's scale.

Tolerances, on max |error| over max |gradient| of each case:
  FP64 plans 1e-9; FP32 plans 2e-4 (FP32 densities, spectra and inner sums,
  FP64 outer sums).  Against finite differences at h = 1e-3 Angstrom:
  FP64 1e-4, FP32 1e-2 (the FP32 forward's S(q) errors enter dL/dS).
"""
import numpy as np
import pytest
import torch

import sf_torch_ref as T
from test_adjoint_ref import corner_system, monoclinic, spill_images, spill_system

pytestmark = pytest.mark.gpu

TOL = {False: 2e-4, True: 1e-9}


def rel_err(got, want):
    return float(np.max(np.abs(got - want)) / np.max(np.abs(want)))


def _plan(coords, dims, typ, rad, sres, ucell, fp64):
    from mdsf import engine
    s = engine.Setup(coords, dims, typ, rad, sres, False, ucell)
    plan = engine.Plan.from_setup(s, fft64=fp64)
    dev = plan.device
    types = torch.from_numpy(s.type_index.astype(np.int32).reshape(-1)).to(dev)
    scale = torch.from_numpy(np.ascontiguousarray(s.scale, dtype=np.float64)).to(dev)
    return s, plan, types, scale


@pytest.mark.parametrize("fp64", [False, True])
@pytest.mark.parametrize("shape", [(12, 10, 8), (16, 16, 16), (32, 24, 20),
                                   # register-staged inverse z pass (FP32), Nz/2 = 64 / 128 / 256,
                                   # partial last block of rows
                                   (6, 10, 128), (6, 6, 256), (2, 6, 512)])
def test_density_adjoint_matches_autograd(shape, fp64, gpu):
    from mdsf import engine
    rng = np.random.default_rng(3)
    rho = rng.normal(size=shape)
    h = shape[2] // 2 + 1
    g = rng.normal(size=(shape[0], shape[1], h))
    x = torch.tensor(rho, requires_grad=True)
    f = torch.fft.rfftn(x)
    loss = torch.sum(torch.tensor(g) * (f.real ** 2 + f.imag ** 2))
    loss.backward()
    want = x.grad.numpy()
    plan = engine.Plan(list(shape), 0, [1.0, 1.0, 1.0], np.eye(3), [], [], [], [], fft64=fp64)
    rho_in = rho if fp64 else rho.astype(np.float32).astype(np.float64)
    got = plan.density_adjoint(torch.tensor(rho_in), torch.tensor(g)).cpu().numpy()
    if not fp64:
        x32 = torch.tensor(rho_in, requires_grad=True)
        f32 = torch.fft.rfftn(x32)
        torch.sum(torch.tensor(g) * (f32.real ** 2 + f32.imag ** 2)).backward()
        want = x32.grad.numpy()
    assert rel_err(got.astype(np.float64), want) <= TOL[fp64] / 10


@pytest.mark.parametrize("fp64", [False, True])
@pytest.mark.parametrize("theta", [None, 120.0])
def test_deposit_adjoint_matches_autograd(theta, fp64, rad, gpu):
    coords, dims, typ = corner_system(7, frames=1)
    ucell = monoclinic(theta)
    s, plan, types, scale = _plan(coords, dims, typ, rad, 0.5, ucell, fp64)
    c = torch.tensor(coords).to(plan.device)
    plan.prepare_frame(c[0], types, scale[0], [float(v) for v in s.L], s.arith, s.wrap)
    rng = np.random.default_rng(1)
    gd = rng.normal(size=tuple(int(v) for v in s.N))
    got = plan.deposit_adjoint(torch.tensor(gd)).cpu().numpy()
    plan.check_errors()
    geo = T.prepare(dims, typ, rad, 0.5)
    r = c[0].cpu().clone().requires_grad_(True)           # wrapped on the GPU
    rho = T.frame_density(r, typ, rad, geo, ucell)
    torch.sum(torch.tensor(gd) * rho).backward()
    assert rel_err(got, r.grad.numpy()) <= TOL[fp64]


@pytest.mark.parametrize("fp64", [False, True])
@pytest.mark.parametrize("theta", [None, 120.0])
def test_deposit_adjoint_corner_spill(theta, fp64, rad, gpu):
    """Corner-slot images (dens.py:107) in z tiles the atom's window does not
    reach (Nborder = 17 at N ~ 40): their gradient terms, against autograd."""
    coords, dims, typ = spill_system(11)
    assert spill_images(coords[0], typ, rad, 1.0, dims) > 100
    ucell = monoclinic(theta)
    s, plan, types, scale = _plan(coords, dims, typ, rad, 1.0, ucell, fp64)
    c = torch.tensor(coords).to(plan.device)
    plan.prepare_frame(c[0], types, scale[0], [float(v) for v in s.L], s.arith, s.wrap)
    rng = np.random.default_rng(5)
    gd = rng.normal(size=tuple(int(v) for v in s.N))
    got = plan.deposit_adjoint(torch.tensor(gd)).cpu().numpy()
    plan.check_errors()
    geo = T.prepare(dims, typ, rad, 1.0)
    r = c[0].cpu().clone().requires_grad_(True)
    rho = T.frame_density(r, typ, rad, geo, ucell)
    torch.sum(torch.tensor(gd) * rho).backward()
    assert rel_err(got, r.grad.numpy()) <= TOL[fp64]


@pytest.mark.parametrize("fp64", [False, True])
@pytest.mark.parametrize("theta", [None, 120.0])
def test_adjoint_frames_match_autograd(theta, fp64, rad, gpu):
    from mdsf.autograd import SFModel
    coords, dims, typ = corner_system(8, frames=3)
    ucell = monoclinic(theta)
    model = SFModel(coords, dims, typ, rad, 0.5, ucell=ucell, fp64=fp64)
    rng = np.random.default_rng(2)
    x = torch.tensor(coords, device=model.plan.device, requires_grad=True)
    S = model(x)
    w = torch.tensor(rng.normal(size=tuple(S.shape)), device=S.device) / S.detach().max()
    torch.sum(w * S).backward()
    got = x.grad.cpu().numpy()
    # the caller's coordinates are untouched (the adjoint works on copies)
    assert np.array_equal(x.detach().cpu().numpy(), coords)
    xr = torch.tensor(coords, requires_grad=True)
    Sr = T.structure_factor(xr, dims, typ, rad, 0.5, ucell)
    np.testing.assert_allclose(S.detach().cpu().numpy(), Sr.detach().numpy(),
                               rtol=1e-4 if not fp64 else 1e-10, atol=1e-6 * float(Sr.max()))
    torch.sum(w.cpu() * Sr).backward()
    assert rel_err(got, xr.grad.numpy()) <= TOL[fp64]


@pytest.mark.parametrize("fp64", [False, True])
@pytest.mark.parametrize("theta", [None, 120.0])
def test_adjoint_frames_corner_spill(theta, fp64, rad, gpu):
    """The pipelined forward and backward frame loops (frame slots alternate)
    on 3 frames with corner-slot spill images, against autograd of the torch
    reference."""
    from mdsf.autograd import SFModel
    coords, dims, typ = spill_system(11, frames=3)
    ucell = monoclinic(theta)
    model = SFModel(coords, dims, typ, rad, 1.0, ucell=ucell, fp64=fp64)
    rng = np.random.default_rng(6)
    x = torch.tensor(coords, device=model.plan.device, requires_grad=True)
    S = model(x)
    w = torch.tensor(rng.normal(size=tuple(S.shape)), device=S.device) / S.detach().max()
    torch.sum(w * S).backward()
    got = x.grad.cpu().numpy()
    xr = torch.tensor(coords, requires_grad=True)
    Sr = T.structure_factor(xr, dims, typ, rad, 1.0, ucell)
    np.testing.assert_allclose(S.detach().cpu().numpy(), Sr.detach().numpy(),
                               rtol=1e-4 if not fp64 else 1e-10, atol=1e-6 * float(Sr.max()))
    torch.sum(w.cpu() * Sr).backward()
    assert rel_err(got, xr.grad.numpy()) <= TOL[fp64]


@pytest.mark.parametrize("fp64", [False, True])
def test_adjoint_clipped_windows(fp64, rad, gpu):
    """The reference's clipped_windows fixture: windows cut at the padded
    grid's high edge (their gradient comes from the clipped-window pass) and
    windows the reference's slices empty (zero gradient), against autograd of
    the torch reference (pinned to the fixture on the CPU)."""
    from conftest import load_golden
    from mdsf.autograd import SFModel
    c = load_golden("clipped_windows")
    coords, dims, typ = c["in_coords"], c["in_dims"], c["typ"]
    model = SFModel(coords, dims, typ, rad, c["sres"], better=bool(c["in_better"]), fp64=fp64)
    rng = np.random.default_rng(12)
    x = torch.tensor(coords, device=model.plan.device, requires_grad=True)
    S = model(x)
    w = torch.tensor(rng.normal(size=tuple(S.shape)), device=S.device) / S.detach().max()
    torch.sum(w * S).backward()
    got = x.grad.cpu().numpy()
    xr = torch.tensor(coords, requires_grad=True)
    Sr = T.structure_factor(xr, dims, typ, rad, c["sres"], np.eye(3), better=bool(c["in_better"]))
    torch.sum(w.cpu() * Sr).backward()
    want = xr.grad.numpy()
    assert rel_err(got, want) <= TOL[fp64]
    cut = [12, 14, 15, 16]        # fixture atoms whose windows the high edge cuts
    assert np.all(np.abs(want[:, cut]).max(axis=-1) > 0)


def test_adjoint_fd_adjoint_workload(rad, gpu):
    """20k C/O/H atoms, 60 A box, 128^3, 16 frames; L = sum (S - S_target)^2.
    FP32 and FP64 adjoints against central FD (h = 1e-3 A) of the FP64 forward
    on 32 sampled coordinates."""
    from mdsf import synth
    from mdsf.autograd import SFModel
    coords, target, dims, typ = synth.adjoint()
    sres = synth.adjoint_sres()
    m64 = SFModel(coords, dims, typ, rad, sres, fp64=True)
    dev = m64.plan.device
    with torch.no_grad():
        St = m64(torch.tensor(target, device=dev))
    xd = torch.tensor(coords, device=dev)

    def loss64(x):
        with torch.no_grad():
            return float(torch.sum((m64(x) - St) ** 2))

    grads = {}
    for fp64 in (True, False):
        model = m64 if fp64 else SFModel(coords, dims, typ, rad, sres, fp64=False)
        x = xd.clone().requires_grad_(True)
        torch.sum((model(x) - St) ** 2).backward()
        grads[fp64] = x.grad.cpu().numpy()
    rng = np.random.default_rng(4006)
    picks = [(int(rng.integers(16)), int(rng.integers(20000)), int(rng.integers(3)))
             for _ in range(32)]
    h = 1e-3
    fd = []
    for (t, a, d) in picks:
        xp, xm = xd.clone(), xd.clone()
        xp[t, a, d] += h
        xm[t, a, d] -= h
        fd.append((loss64(xp) - loss64(xm)) / (2 * h))
    fd = np.array(fd)
    scale = np.max(np.abs(fd))
    e64 = np.max(np.abs(np.array([grads[True][p] for p in picks]) - fd)) / scale
    e32 = np.max(np.abs(np.array([grads[False][p] for p in picks]) - fd)) / scale
    print(f"adjoint vs FD: fp64 {e64:.3e}, fp32 {e32:.3e} (max |dL/dx| sampled {scale:.3e})")
    assert e64 <= 1e-4
    assert e32 <= 1e-2
