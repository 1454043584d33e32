"""Thin tensor-level wrappers of the pair kernels (torch tensors in, torch tensors out).

All tensors must already live on the GPU; the calls are asynchronous on the
current torch stream.  Pair inputs: qf (N,4) float32 orientations, pair_i /
pair_j (P,) int32, pair_t (P,4) float32 minimum-image X_j - X_i (w ignored).
"""
import ctypes as C

import torch

from . import _lib
from .model import EnergyNet, n_descriptor


def _count(P, device):
    return torch.tensor([P], dtype=torch.int32, device=device)


def _pairs_args(qf, pi, pj, pt):
    for t, dt in ((qf, torch.float32), (pi, torch.int32), (pj, torch.int32), (pt, torch.float32)):
        if t.dtype != dt or not t.is_cuda or not t.is_contiguous():
            raise TypeError("pair inputs must be contiguous CUDA tensors (qf/pt float32, pairs int32)")
    if pt.numel() != 4 * pi.numel() or pj.numel() != pi.numel():
        raise ValueError("pair_t must be (P,4) and pair_i/pair_j (P,)")
    return _lib.ptr(qf), _lib.ptr(pi), _lib.ptr(pj), _lib.ptr(pt)


def descriptor_forward(model, qf, pi, pj, pt, n_dev=None, stream=None):
    P = pi.numel()
    D = n_descriptor(model.nrad, model.nang, model.lmax)
    G = torch.empty((P, D), dtype=torch.float32, device=qf.device)
    if P == 0:
        return G
    n_dev = _count(P, qf.device) if n_dev is None else n_dev
    a = _pairs_args(qf, pi, pj, pt)
    _lib.check(_lib.lib.nepx_descriptor_fwd(C.byref(model), *a, _lib.ptr(n_dev), P, _lib.ptr(G),
                                            _lib.stream_handle(stream)), "nepx_descriptor_fwd")
    return G


def ffnn_forward_backward(net: EnergyNet, G, n_dev=None, stream=None):
    """E (P,) = en_min + (en_max-en_min) NN(G) and dE/dG (P, D)."""
    P, D = G.shape
    if D != net.d_in:
        raise ValueError("descriptor width does not match the network")
    E = torch.empty(P, dtype=torch.float32, device=G.device)
    dG = torch.empty_like(G)
    if P == 0:
        return E, dG
    n_dev = _count(P, G.device) if n_dev is None else n_dev
    cn = net.cstruct()
    _lib.check(_lib.lib.nepx_ffnn_fwd_bwd(C.byref(cn), _lib.ptr(G.contiguous()), _lib.ptr(n_dev), P, _lib.ptr(E),
                                          _lib.ptr(dG), _lib.stream_handle(stream)), "nepx_ffnn_fwd_bwd")
    return E, dG


def descriptor_adjoint(model, qf, pi, pj, pt, dEdG, E=None, clip=0.0, n_dev=None, stream=None):
    """Per-pair (P,12) = [F_i, E | tau_i, 0 | tau_j, 0]."""
    P = pi.numel()
    out = torch.empty((P, 12), dtype=torch.float32, device=qf.device)
    if P == 0:
        return out
    n_dev = _count(P, qf.device) if n_dev is None else n_dev
    a = _pairs_args(qf, pi, pj, pt)
    _lib.check(_lib.lib.nepx_descriptor_adjoint(C.byref(model), *a, _lib.ptr(n_dev), P, _lib.ptr(dEdG.contiguous()),
                                                _lib.ptr(E), float(clip), _lib.ptr(out),
                                                _lib.stream_handle(stream)), "nepx_descriptor_adjoint")
    return out


def pair_fused(model, net: EnergyNet, qf, pi, pj, pt, clip=0.0, n_dev=None, stream=None):
    P = pi.numel()
    out = torch.empty((P, 12), dtype=torch.float32, device=qf.device)
    if P == 0:
        return out
    n_dev = _count(P, qf.device) if n_dev is None else n_dev
    a = _pairs_args(qf, pi, pj, pt)
    cn = net.cstruct()
    _lib.check(_lib.lib.nepx_pair_fused(C.byref(model), C.byref(cn), *a, _lib.ptr(n_dev), P, float(clip),
                                        _lib.ptr(out), _lib.stream_handle(stream)), "nepx_pair_fused")
    return out


def cloud_descriptor(clouds, nrad, nang, lmax, r_nep, weights=None, stream=None):
    """Descriptors (P, D) float32 of P site clouds (P, Nd, 3) float32 on the GPU, each taken about its
    own centroid, with optional per-site weights (Nd,) (nepx_cloud_descriptor)."""
    if clouds.dtype != torch.float32 or not clouds.is_cuda or clouds.dim() != 3 or clouds.shape[2] != 3:
        raise TypeError("clouds must be a (P, Nd, 3) float32 CUDA tensor")
    clouds = clouds.contiguous()
    P, nd, _ = clouds.shape
    D = n_descriptor(nrad, nang, lmax)
    G = torch.empty((P, D), dtype=torch.float32, device=clouds.device)
    if P == 0:
        return G
    w = None
    if weights is not None:
        w = torch.as_tensor(weights, dtype=torch.float32, device=clouds.device).contiguous()
        if w.numel() != nd:
            raise ValueError(f"{w.numel()} site weights for clouds of {nd} sites")
    _lib.check(_lib.lib.nepx_cloud_descriptor(_lib.ptr(clouds), P, nd, _lib.ptr(w) if w is not None else None,
                                              int(nrad), int(nang), int(lmax), float(r_nep), _lib.ptr(G),
                                              _lib.stream_handle(stream)), "nepx_cloud_descriptor")
    return G
