"""PyTorch autograd binding of the superposition (no reference counterpart).

``Propagator(u, v, ...)`` maps source weights (amplitude A, phase theta, on
the N points u) to the complex field z on the M points v:

    z_m = sum_n A_n e^{i theta_n} e^{i sign 2 pi d_nm / lambda} / d_nm

Forward runs ``holo_plan_transform`` and backward the hand-written adjoint
``holo_plan_adjoint`` (the conjugate-transpose sum from v back to u, with the
phasors recomputed on the fly), both through the C ABI with device pointers
on torch's current stream.  z is returned as a real (M, 2) tensor [re, im];
its gradient (dL/dRe z, dL/dIm z) is exactly what the adjoint consumes.
With ``obliquity=True`` every term also carries |dz_nm| / d_nm (the MATLAB
light-field model's cosine factor); the adjoint keeps it, since it is
symmetric in the pair.
"""
from __future__ import annotations

import torch

from . import FP32_SPLIT, Context, Plan


class _SuperposeFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, amp, theta, prop):
        z = prop.plan_u.transform(amp, theta, prop.v, prop.sign, prop.lam, prop.precision,
                                  obliquity=prop.obliquity)
        ctx.prop = prop
        ctx.save_for_backward(amp, theta)
        return torch.view_as_real(z)

    @staticmethod
    def backward(ctx, gz):
        amp, theta = ctx.saved_tensors
        prop = ctx.prop
        ga, gt = prop.plan_v.adjoint(gz.contiguous(), amp, theta, prop.u, prop.sign, prop.lam,
                                     prop.precision, obliquity=prop.obliquity)
        return ga, gt, None


class Propagator:
    """Differentiable z = superposition(A, theta) from points u to points v."""

    def __init__(self, u: torch.Tensor, v: torch.Tensor, sign: int = +1,
                 lam: float = 632.8e-9, precision: int = FP32_SPLIT, tile: int = 64,
                 ctx: Context = None, obliquity: bool = False):
        if not (u.is_cuda and v.is_cuda):
            raise ValueError("Propagator needs GPU tensors (no CPU fallback)")
        self.u = u.to(torch.float64).contiguous()
        self.v = v.to(torch.float64).contiguous()
        self.sign, self.lam, self.precision = sign, lam, precision
        self.obliquity = obliquity
        self.ctx = ctx or Context(u.device.index or 0)
        self.plan_u = Plan(self.ctx, self.u, tile)   # forward sources
        self.plan_v = Plan(self.ctx, self.v, tile)   # adjoint sources

    def __call__(self, amp: torch.Tensor, theta: torch.Tensor) -> torch.Tensor:
        return _SuperposeFn.apply(amp.to(torch.float64).contiguous(),
                                  theta.to(torch.float64).contiguous(), self)
