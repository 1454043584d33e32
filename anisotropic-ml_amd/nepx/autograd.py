"""torch.autograd.Function over the hand-written adjoint.

NepEnergy.apply(pos, rot, quat, arena) -> total pair energy U (FP64 scalar)
  pos  (N, 3) float64  centre-of-mass positions (requires_grad as needed)
  rot  (N, 3) float64  lab-frame rotation vectors applied to `quat` (pass zeros)
  quat (N, 4)          orientations, scalar first (not differentiated)
backward: dU/dpos = -F and dU/drot = -tau, both from the HIP adjoint
(nepx_descriptor_adjoint or the fused kernel) and the deterministic gather --
no autograd through the descriptor.  dU/drot is the exact derivative at
rot = 0; elsewhere it is the torque at the rotated orientation (the
first-order change of U under a further small rotation).

The configuration is written into the arena, the list is rebuilt for it and
the pair kernels run; U, F and tau stay on the device.
"""
import math

import torch

from . import _lib
from .system import Arena


def rotate_quaternions(quat, rot):
    """q' = exp(rot/2) (x) q for lab-frame rotation vectors rot (N, 3)."""
    ang = torch.linalg.norm(rot, dim=1, keepdim=True)
    half = 0.5 * ang
    axis = torch.where(ang > 0, rot / torch.clamp(ang, min=1e-300), torch.zeros_like(rot))
    dq = torch.cat([torch.cos(half), torch.sin(half) * axis], dim=1)
    w1, v1 = dq[:, :1], dq[:, 1:]
    w2, v2 = quat[:, :1], quat[:, 1:]
    w = w1 * w2 - torch.sum(v1 * v2, dim=1, keepdim=True)
    v = w1 * v2 + w2 * v1 + torch.linalg.cross(v1, v2, dim=1)
    return torch.cat([w, v], dim=1)


def evaluate(arena: Arena, pos, quat, k=0):
    """Energy U (float), forces F and torques T (N, 3) FP64 tensors for system k at (pos, quat)."""
    arena.set_configuration(pos.detach(), quat.detach(), k)
    arena.build_list(True)
    f = arena.flags()
    if f[_lib.FLAG_OVERFLOW]:
        arena.ensure_capacity()
    arena.filter_pairs()
    arena.pair_fused()
    arena.gather()
    F, T = arena.forces(k)
    a, b = arena.starts[k], arena.starts[k] + arena.specs[k].n
    U = arena.t["upot"][a:b].sum()
    return U, F, T


class NepEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, rot, quat, arena):
        q = rotate_quaternions(quat.to(torch.float64), rot.detach().to(torch.float64))
        U, F, T = evaluate(arena, pos, q)
        ctx.save_for_backward(F.clone(), T.clone())
        return U.to(pos.dtype)

    @staticmethod
    def backward(ctx, gU):
        F, T = ctx.saved_tensors
        return -gU * F, -gU * T, None, None


def nep_energy(pos, rot, quat, arena):
    return NepEnergy.apply(pos, rot, quat, arena)
