"""Pair model (shape + NEP hypers) and the energy network as device data."""
import ctypes as C
import math

import numpy as np
import torch

from . import _lib
from .shapes import Shape


def n_descriptor(nrad, nang, lmax):
    """D = (nrad+1) + (nang+1)*lmax (nep_simulator/DescriptorGenerator.py:37-42)."""
    return (int(nrad) + 1) + (int(nang) + 1) * int(lmax)


def make_pair_model(shape: Shape, hypers, r_nep=None):
    nrad, nang, lmax = (int(h) for h in hypers[:3])
    if nang > nrad:
        raise ValueError("nang must not exceed nrad")
    s = np.asarray(shape.sites, dtype=np.float32)
    if len(s) > _lib.MAX_SITES:
        raise ValueError(f"at most {_lib.MAX_SITES} sites per body")
    m = _lib.PairModel()
    m.n_sites = len(s)
    m.nrad, m.nang, m.lmax = nrad, nang, lmax
    m.r_nep = float(shape.nep_cutoff if r_nep is None else r_nep)
    flat = np.zeros(3 * _lib.MAX_SITES, dtype=np.float32)
    flat[: s.size] = s.reshape(-1)
    m.sites[:] = flat.tolist()
    return m


class EnergyNet:
    """TheEnergyNet (nep_simulator/ForceEvaluator.py:6-26) held as MFMA-packed device weights.

    layers: list of (W (out, in), b (out,)) float arrays, input -> width ->
    (width)*n_hidden -> 1, activation after every layer but the last.
    """

    def __init__(self, layers, act="tanh", en_min=0.0, en_max=1.0, device="cuda", precision="auto"):
        self.layers = [(np.asarray(W, dtype=np.float32), np.asarray(b, dtype=np.float32).reshape(-1))
                       for W, b in layers]
        self.d_in = self.layers[0][0].shape[1]
        self.width = self.layers[0][0].shape[0]
        self.n_hidden = len(self.layers) - 2
        for W, _ in self.layers[1:-1]:
            if W.shape != (self.width, self.width):
                raise ValueError("hidden layers must be width x width")
        if self.layers[-1][0].shape != (1, self.width):
            raise ValueError("last layer must map width -> 1")
        self.act = act
        self.en_min = float(en_min)
        self.en_max = float(en_max)
        self.device = device
        self.packed = self._pack()
        if precision == "auto":
            # the fp16 split needs bounded hidden activations (tanh); ReLU networks run in FP32
            precision = "f16x3" if act == "tanh" and self._lowp_size("f16x3") > 0 else "f32"
        if precision not in _lib.NET_PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(_lib.NET_PRECISIONS)} or 'auto'")
        if precision in ("bf16", "bf16x3") and act != "tanh":
            # bf16 rounding moves pre-activations across zero, where ReLU's derivative jumps: dE/dG
            # errors of 15-30 % (bf16) and 4-6 % (bf16x3) measured on 40->64->64 and 150-wide networks
            raise ValueError("the bf16 forms are supported for tanh networks only; use f16x3 or f32 for ReLU")
        self.precision = precision
        self.packed_half = self._pack_lowp(precision) if precision != "f32" else None
        if precision != "f32" and self.packed_half is None:
            raise ValueError(f"the {precision} path needs a width whose 16-padding is a multiple of 32")
        self.w_dev = torch.from_numpy(self.packed).to(device) if device is not None else None
        self.wh_dev = (torch.from_numpy(self.packed_half.view(np.int16)).to(device)
                       if device is not None and self.packed_half is not None else None)

    @classmethod
    def random(cls, d_in, widths, seed, act="tanh", en_min=0.0, en_max=1.0, device="cuda", precision="auto"):
        """Seeded N(0, 1/fan_in) weights and zero biases."""
        rng = np.random.default_rng(seed)
        dims = [d_in] + list(widths) + [1]
        layers = [(rng.standard_normal((b, a)) / math.sqrt(a), np.zeros(b)) for a, b in zip(dims[:-1], dims[1:])]
        return cls(layers, act, en_min, en_max, device, precision)

    @classmethod
    def from_state_dict(cls, sd, act="relu", en_min=0.0, en_max=1.0, device="cuda", precision="auto"):
        """TheEnergyNet state_dict: layer_first, layer_hidden.k, layer_last."""
        def g(k):
            v = sd[k]
            return v.detach().cpu().numpy() if hasattr(v, "detach") else np.asarray(v)
        layers = [(g("layer_first.weight"), g("layer_first.bias"))]
        k = 0
        while f"layer_hidden.{k}.weight" in sd:
            layers.append((g(f"layer_hidden.{k}.weight"), g(f"layer_hidden.{k}.bias")))
            k += 1
        layers.append((g("layer_last.weight"), g("layer_last.bias")))
        return cls(layers, act, en_min, en_max, device, precision)

    @classmethod
    def load(cls, path, **kw):
        """Read a .pth state_dict with the safe loader (no pickled code)."""
        sd = torch.load(path, map_location="cpu", weights_only=True)
        return cls.from_state_dict(sd, **kw)

    def _pack(self):
        nl = len(self.layers)
        n = _lib.lib.nepx_net_packed_floats(self.d_in, self.width, self.n_hidden)
        if n <= 0:
            raise _lib.NepxError("bad network dimensions")
        out = np.zeros(n, dtype=np.float32)
        ws = [np.ascontiguousarray(W, dtype=np.float32) for W, _ in self.layers]
        bs = [np.ascontiguousarray(b, dtype=np.float32) for _, b in self.layers]
        WP = (_lib.c_f32p * nl)(*[w.ctypes.data_as(_lib.c_f32p) for w in ws])
        BP = (_lib.c_f32p * nl)(*[b.ctypes.data_as(_lib.c_f32p) for b in bs])
        _lib.check(_lib.lib.nepx_net_pack(self.d_in, self.width, self.n_hidden, WP, BP,
                                          out.ctypes.data_as(_lib.c_f32p)), "nepx_net_pack")
        return out

    def _lowp_size(self, precision):
        return _lib.lib.nepx_net_packed_lowp(_lib.NET_PRECISIONS[precision], self.d_in, self.width, self.n_hidden)

    def _pack_lowp(self, precision):
        """16-bit MFMA operands (fp16 or bf16, hi or hi/lo) in the kernels' operand order."""
        n = self._lowp_size(precision)
        if n <= 0:
            return None
        out = np.zeros(n, dtype=np.uint16)
        ws = [np.ascontiguousarray(W, dtype=np.float32) for W, _ in self.layers]
        WP = (_lib.c_f32p * len(ws))(*[w.ctypes.data_as(_lib.c_f32p) for w in ws])
        _lib.check(_lib.lib.nepx_net_pack_lowp(_lib.NET_PRECISIONS[precision], self.d_in, self.width, self.n_hidden,
                                               WP, out.ctypes.data_as(C.c_void_p)), "nepx_net_pack_lowp")
        return out

    def with_precision(self, precision):
        """A copy of this network run at another MFMA precision ('f32', 'f16x3', 'bf16', 'bf16x3')."""
        return EnergyNet(self.layers, self.act, self.en_min, self.en_max, self.device, precision)

    def cstruct(self):
        n = _lib.Net()
        n.d_in, n.width, n.n_hidden = self.d_in, self.width, self.n_hidden
        n.act = _lib.ACT_TANH if self.act == "tanh" else _lib.ACT_RELU
        n.en_min, n.en_max = self.en_min, self.en_max
        n.w_packed = self.w_dev.data_ptr()
        n.w_half = self.wh_dev.data_ptr() if self.wh_dev is not None else None
        n.precision = _lib.NET_PRECISIONS[self.precision]
        return n

    def layers_f64(self):
        return [(W.astype(np.float64), b.astype(np.float64)) for W, b in self.layers]
