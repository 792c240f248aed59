"""S(q) as a differentiable torch function with a hand-written adjoint.

``structure_factor(coords, plan, ...)`` returns the FP64 half-spectrum S(q) =
sum over frames of |rfftn(rho_t)|^2 (the ``sf`` of ``dens.compute_sf``,
dens.py:277-318) for ORIGINAL, un-rescaled coordinates; its backward calls
``mdsf_adjoint_frames``, which recomputes each frame's density and spectrum
(nothing per frame is kept between forward and backward) and returns
dL/d(coords).  The rescale and wrap of the reference are applied to a copy, so
the caller's tensor is not modified.

``SFModel`` binds a system (box per frame, types, radii, resolution, cell) to
a plan once and is then called on coordinate tensors.
"""
import numpy as np
import torch

from . import engine


class StructureFactorFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, coords, plan, types, scale, box, arith, wrap):
        work = coords.detach().clone()
        plan.reset()
        plan.accumulate_frames(work, types, scale, box, arith, wrap)
        plan.check_errors()
        ctx.save_for_backward(coords)
        ctx.args = (plan, types, scale, box, arith, wrap)
        return plan.sf()

    @staticmethod
    def backward(ctx, grad_sf):
        (coords,) = ctx.saved_tensors
        plan, types, scale, box, arith, wrap = ctx.args
        work = coords.detach().clone()
        grad = plan.adjoint_frames(work, types, scale, box, arith, wrap, grad_sf)
        plan.check_errors()
        return grad.to(coords.dtype), None, None, None, None, None, None


def structure_factor(coords, plan, types, scale, box, arith, wrap):
    """Differentiable S(q) of device ``coords`` (frames, atoms, 3) through ``plan``."""
    return StructureFactorFn.apply(coords, plan, types, scale, box, arith, wrap)


class SFModel:
    """A system bound to a GPU plan: ``model(coords) -> S(q)``, differentiable in coords.

    ``coords_like`` (frames, atoms, 3) fixes the dtype and shape the model is
    called with; ``dims`` (frames, 3) are the per-frame boxes, as for
    ``dens.compute_sf``.
    """

    def __init__(self, coords_like, dims, typ, rad, sres, better=False, ucell=None, fp64=False,
                 device=None):
        r = np.asarray(coords_like)
        ucell = np.eye(3) if ucell is None else ucell
        self.setup = s = engine.Setup(r, dims, typ, rad, sres, better, ucell)
        self.plan = engine.Plan.from_setup(s, device=device, fft64=fp64)
        dev = self.plan.device
        self.types = torch.from_numpy(s.type_index.astype(np.int32).reshape(-1)).to(dev)
        self.scale = torch.from_numpy(np.ascontiguousarray(s.scale, dtype=np.float64)).to(dev)
        self.box = [float(x) for x in s.L]

    def __call__(self, coords):
        return structure_factor(coords, self.plan, self.types, self.scale, self.box,
                                self.setup.arith, self.setup.wrap)

    def close(self):
        self.plan.close()
