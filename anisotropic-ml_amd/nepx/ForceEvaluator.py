"""ForceEvaluator with the reference call surface (nep_simulator/ForceEvaluator.py:6-146).

evaluate_interactions(G, pp, N_pair) -> (forces, torks), (N, 3) FP64 tensors.
  * analytic path (G has N_pair rows): MFMA network forward + backward to the
    descriptor (nepx_ffnn_fwd_bwd), then the hand-written descriptor adjoint
    (nepx_descriptor_adjoint) on the pair geometry kept by the attached
    DescriptorGenerator, clipped at +-100 per component like the reference;
  * reference FD path (G has 10 * N_pair stencil rows, DescriptorGenerator
    fd_compat): energies of every row, forward differences and clip
    (nepx_fd_forces), exactly the reference algorithm;
  * either way the per-pair results are summed per particle by the
    deterministic gather (nepx_forces_gather) instead of index_add_.
"""
import numpy as np
import torch
import torch.nn as nn

from . import _lib, ops
from .model import EnergyNet


class TheEnergyNet(nn.Module):
    """Same module layout as the reference network, so its state_dicts load unchanged."""

    def __init__(self, nn_width, num_hidden, input_size):
        super().__init__()
        self.layer_first = nn.Linear(input_size, nn_width)
        self.layer_hidden = nn.ModuleList([nn.Linear(nn_width, nn_width) for _ in range(num_hidden)])
        self.layer_last = nn.Linear(nn_width, 1)

    def forward(self, x):
        act = nn.ReLU()
        u = act(self.layer_first(x))
        for hidden in self.layer_hidden:
            u = act(hidden(u))
        return self.layer_last(u)


class ForceEvaluator:
    def __init__(self):
        self.act = "relu"
        self.clip = 100.0
        self.descriptor_generator = None
        self.is_sync = 0
        self.torch_device = torch.device("cuda:0")

    def setDevice(self, gpu_id):
        self.gpu_id = int(gpu_id)
        self.torch_device = torch.device(f"cuda:{self.gpu_id}")

    def setNN_dimensions(self, w, d):
        self.width = int(w)
        self.depth = int(d)

    def set_dx_dteta(self, dx, dteta):
        self.dx = float(dx)
        self.dteta = float(dteta)

    def setNparticles(self, Nparticles):
        self.Nparticles = int(Nparticles)

    def setNd(self, Nd):
        self.Nd = int(Nd)

    def setSync(self, is_sync):
        self.is_sync = is_sync

    def setActivation(self, act):
        """'relu' (the reference network) or 'tanh' (benchmark networks)."""
        self.act = act

    def read_model(self, model_path):
        """Load a TheEnergyNet state_dict (safe loader) and pack it for the MFMA kernels."""
        sd = torch.load(model_path, map_location="cpu", weights_only=True)
        net = TheEnergyNet(self.width, self.depth, self.Nd)
        net.load_state_dict(sd)   # validates the shapes against setNN_dimensions / setNd
        self.energy_net = net.eval()
        self._net_from_state(sd)

    def _net_from_state(self, sd):
        self.net = EnergyNet.from_state_dict(sd, act=self.act, device=self.torch_device)

    def set_model(self, net: EnergyNet):
        self.net = net

    def set_energy_range(self, en_min, en_max):
        self.en_min = float(en_min)
        self.en_max = float(en_max)
        if hasattr(self, "net"):
            self.net.en_min, self.net.en_max = self.en_min, self.en_max

    def move_data_to_device(self):
        self.net.en_min, self.net.en_max = self.en_min, self.en_max

    def attach(self, descriptor_generator):
        """Pair geometry source for the analytic adjoint (set by Sim.setShape)."""
        self.descriptor_generator = descriptor_generator

    # ------------------------------------------------------------------
    def evaluate_interactions(self, g_all_cupy, pp, N_pair):
        self.evaluate_gradients(g_all_cupy, N_pair)
        self.evaluate_forces_torques(pp)
        return self.forces, self.torks

    def evaluate_gradients(self, g_all_cupy, N_pair):
        dg = self.descriptor_generator
        if dg is None:
            raise _lib.NepxError("ForceEvaluator needs the DescriptorGenerator that produced G (attach())")
        G = torch.as_tensor(g_all_cupy)
        P = int(N_pair)
        E, dEdG = ops.ffnn_forward_backward(self.net, G.contiguous())
        if G.shape[0] == P:
            self.pair_out = ops.descriptor_adjoint(dg.model, dg._ctx.t["qf"].view(-1, 4), dg.pair_i.contiguous(),
                                                   dg.pair_j.contiguous(), dg.pair_t.contiguous(), dEdG, E=E,
                                                   clip=self.clip)
        elif G.shape[0] == 10 * P:
            self.pair_out = torch.empty((P, 12), dtype=torch.float32, device=G.device)
            _lib.check(_lib.lib.nepx_fd_forces(_lib.ptr(E), P, dg.dx, dg.dteta, self.clip, _lib.ptr(self.pair_out),
                                               _lib.stream_handle()), "nepx_fd_forces")
        else:
            raise ValueError(f"G has {G.shape[0]} rows; expected N_pair ({P}) or 10 * N_pair")
        self.e0 = self.pair_out[:, 3:4]
        self.gradient_dx = self.pair_out[:, 0:3]
        self.gradient_dtetax = self.pair_out[:, 4:7]
        self.gradient_dtetax2 = self.pair_out[:, 8:11]

    def evaluate_forces_torques(self, pp):
        dg = self.descriptor_generator
        ctx = dg._ctx
        if pp.shape[0] != self.pair_out.shape[0]:
            raise ValueError("pp does not match the pairs of the last descriptor call")
        # the gather reads per-pair rows by compacted index, as the filter laid them out
        if ctx.pair_out.numel() < self.pair_out.numel():
            ctx.pair_out = torch.zeros(self.pair_out.numel(), dtype=torch.float32, device=self.pair_out.device)
        ctx.pair_out[: self.pair_out.numel()].copy_(self.pair_out.reshape(-1))
        ctx.gather()
        F, T = ctx.forces(0)
        self.forces = F
        self.torks = T

