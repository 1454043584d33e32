"""Quaternion product (nep_simulator/MathUtils.py:3-24) on torch tensors, scalar first."""
import torch


def quaternion_multiplication(a, b):
    """(a.s b.s - a.v.b.v, a.s b.v + b.s a.v + a.v x b.v); not commutative."""
    s = a[:, 0] * b[:, 0] - torch.sum(a[:, 1:] * b[:, 1:], dim=1)
    v = a[:, :1] * b[:, 1:] + b[:, :1] * a[:, 1:] + torch.linalg.cross(a[:, 1:], b[:, 1:], dim=1)
    return torch.cat([s[:, None], v], dim=1)
