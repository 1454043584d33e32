"""Device-resident state of one or more rigid-body systems and the MD step.

Layout in HBM (all structure-of-arrays, FP64 unless noted):
  x y z | vx vy vz | q0..q3 | p0..p3 | mass | fx fy fz tx ty tz upot | xb yb zb
  qf (float4 copy of q for the pair kernels), sys (int32 system id, -1 = padding)
Each system occupies a block-aligned particle range (256 particles per
integrator block), so per-system reductions never straddle a block.
The neighbour list, compacted pair arrays and per-pair outputs live beside it.
"""
import ctypes as C
import math

import numpy as np
import torch

from . import _lib
from .model import EnergyNet, make_pair_model, n_descriptor
from .shapes import Shape

BLOCK = 256
REF_THERMO_INIT = dict(xi=-0.0317711, eta=0.108609, xi_r=0.00814227, eta_r=-0.014238)  # Sim.py:99-111


def _align(n, b=BLOCK):
    return (n + b - 1) // b * b


class SystemSpec:
    """Host description of one system (FP64 numpy arrays)."""

    def __init__(self, x, v, q, p, mass, L, kT, dt, moi, rc, skin=0.5, tau=None, thermo=None,
                 ndof=None, rndof=None):
        self.x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        n = len(self.x)
        self.v = np.ascontiguousarray(v, dtype=np.float64).reshape(n, 3)
        self.q = np.ascontiguousarray(q, dtype=np.float64).reshape(n, 4)
        self.p = np.ascontiguousarray(p, dtype=np.float64).reshape(n, 4)
        self.mass = np.broadcast_to(np.asarray(mass, dtype=np.float64), (n,)).copy()
        self.L = float(L)
        self.kT = float(kT)
        self.dt = float(dt)
        self.tau = 100.0 * self.dt if tau is None else float(tau)   # Sim.setdt (Sim.py:113-115)
        self.moi = tuple(float(m) for m in moi)
        self.rc = float(rc)
        self.skin = float(skin)
        self.thermo = dict(REF_THERMO_INIT if thermo is None else thermo)
        self.ndof = 3.0 * (n - 1) if ndof is None else float(ndof)   # Sim.py:35-36
        self.rndof = 3.0 * n if rndof is None else float(rndof)

    @property
    def n(self):
        return len(self.x)


def _nullctx():
    import contextlib
    return contextlib.nullcontext()


class Arena:
    """One or more systems sharing a pair model and energy network, resident on one GPU."""

    def __init__(self, specs, shape: Shape, hypers, net: EnergyNet, device="cuda", clip=100.0,
                 list_cap=None, r_nep=None, potential=None, reverse_records=False):
        if not torch.cuda.is_available():
            raise _lib.NepxError("nepx needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.device = torch.device(device)
        self.specs = list(specs)
        self.shape = shape
        self.hypers = tuple(int(h) for h in hypers[:3])
        self.D = n_descriptor(*self.hypers)
        if net is not None and net.d_in != self.D:
            raise ValueError(f"network input {net.d_in} != descriptor size {self.D}")
        self.net = net
        # a pair potential in place of the NEP model (label.BeadPotential); None = descriptor + network
        self.potential = potential
        self.model = make_pair_model(shape, self.hypers, r_nep)
        self.cnet = net.cstruct() if net is not None else None
        self.clip = float(clip)
        # reverse-order pair records: the fused pair kernel also writes each pair's j side at its
        # reverse-row entry, so the gather reads it contiguously (a third less gather traffic), at the
        # price of scattered writes in the pair kernel; measured slower per step, so off by default
        self.reverse_records = bool(reverse_records)
        self._layout()
        self._alloc_state()
        self._alloc_list(list_cap)
        self.steps = 0

    # ---------------------------------------------------------------- layout
    def _layout(self):
        starts, off = [], 0
        for sp in self.specs:
            starts.append(off)
            off += _align(sp.n)
        self.n = max(off, BLOCK)
        self.starts = starts
        rec = np.zeros(len(self.specs), dtype=_lib.SYSPARAM_DTYPE)
        cell_off = 0
        for k, sp in enumerate(self.specs):
            r_list = sp.rc + sp.skin
            if sp.L < 2.0 * sp.rc:
                raise ValueError(f"box {sp.L} smaller than twice the cutoff {sp.rc}")
            nc = int(math.floor(sp.L / r_list))
            nc = nc if nc >= 3 else 0
            r = rec[k]
            r["L"], r["kT"], r["dt"], r["tau"] = sp.L, sp.kT, sp.dt, sp.tau
            r["ndof"], r["rndof"] = sp.ndof, sp.rndof
            r["moi"] = sp.moi
            r["r_list"], r["rc"], r["skin"] = r_list, sp.rc, sp.skin
            r["p_start"], r["p_end"] = starts[k], starts[k] + sp.n
            r["blk_start"], r["blk_end"] = starts[k] // BLOCK, (starts[k] + sp.n + BLOCK - 1) // BLOCK
            r["nc"], r["cell_off"] = nc, cell_off
            r["active"] = 1
            cell_off += nc ** 3
        self.sysrec = rec
        self.n_cells = max(cell_off, 2 * self.n)

    def _alloc_state(self):
        dev, n = self.device, self.n
        f64 = lambda: torch.zeros(n, dtype=torch.float64, device=dev)
        self.t = {k: f64() for k in ["x", "y", "z", "vx", "vy", "vz", "q0", "q1", "q2", "q3", "p0", "p1", "p2",
                                     "p3", "mass", "fx", "fy", "fz", "tx", "ty", "tz", "upot", "xb", "yb",
                                     "zb"]}
        self.t["qf"] = torch.zeros(n * 4, dtype=torch.float32, device=dev)
        host = {k: np.zeros(n) for k in ["x", "y", "z", "vx", "vy", "vz", "q0", "q1", "q2", "q3", "p0", "p1",
                                         "p2", "p3"]}
        host["q0"][:] = 1.0
        mass = np.ones(n)
        sysid = np.full(n, -1, dtype=np.int32)
        for k, sp in enumerate(self.specs):
            a, b = self.starts[k], self.starts[k] + sp.n
            for c, key in enumerate("xyz"):
                host[key][a:b] = sp.x[:, c]
                host["v" + key][a:b] = sp.v[:, c]
            for c in range(4):
                host[f"q{c}"][a:b] = sp.q[:, c]
                host[f"p{c}"][a:b] = sp.p[:, c]
            mass[a:b] = sp.mass
            sysid[a:b] = k
        for k, v in host.items():
            self.t[k].copy_(torch.from_numpy(v))
        self.t["mass"].copy_(torch.from_numpy(mass))
        self.t["sys"] = torch.from_numpy(sysid).to(dev)
        self.t["sp"] = torch.from_numpy(self.sysrec.view(np.uint8).copy()).to(dev)
        th = np.zeros(len(self.sysrec), dtype=_lib.THERMO_DTYPE)
        for k, sp in enumerate(self.specs):
            for key, val in sp.thermo.items():
                th[k][key] = val
            th[k]["theta"] = 1.0
            th[k]["e_r"] = 1.0
        self.t["th"] = torch.from_numpy(th.view(np.uint8).copy()).to(dev)
        self.t["partials"] = torch.zeros(4 * (n // BLOCK + 1), dtype=torch.float64, device=dev)
        self.t["dp"] = torch.zeros(4 * n, dtype=torch.float64, device=dev)
        st = _lib.State()
        st.n, st.n_sys = n, len(self.sysrec)
        for k in _lib._STATE_PTRS:
            setattr(st, k, self.t[k].data_ptr())
        self.st = st

    def _estimate_cap(self):
        tot = 0.0
        for sp in self.specs:
            rho = sp.n / sp.L ** 3
            tot += 0.5 * sp.n * min(sp.n - 1, rho * 4.0 / 3.0 * math.pi * (sp.rc + sp.skin) ** 3)
        return int(2.0 * tot) + 4096    # structure can pack pairs well above the uniform-density count

    def _alloc_list(self, cap=None):
        dev, n = self.device, self.n
        cap = int(cap or self._estimate_cap())
        builds = int(self.lt["flags"][_lib.FLAG_BUILDS].item()) if getattr(self, "lt", None) else 0
        i32 = lambda m: torch.zeros(m, dtype=torch.int32, device=dev)
        L = {
            "cell_count": i32(self.n_cells + 1), "cell_start": i32(self.n_cells + 1),
            "cell_fill": i32(self.n_cells + 1), "cell_list": i32(n), "cell_of": i32(n),
            "cell_pos": torch.zeros(4 * n, dtype=torch.float64, device=dev),
            "row_count": i32(n + 1), "row_start": i32(n + 1), "rrow_count": i32(n + 1),
            "rrow_start": i32(n + 1), "nbr": i32(cap), "slot_row": i32(cap), "rnbr": i32(cap),
            "cidx": i32(cap), "blk_count": i32(cap // 1024 + 4), "flags": i32(_lib.N_FLAGS),
            "pair_i": i32(cap), "pair_j": i32(cap),
            "pair_t": torch.zeros(cap * 4, dtype=torch.float32, device=dev),
        }
        if self.reverse_records:
            # j side of each pair in reverse-row order, written by the fused pair kernel; zero tags
            # (never a pass tag) until then
            L.update({"rpos": i32(cap), "pair_e": i32(cap),
                      "rrec": torch.zeros(cap * 8, dtype=torch.float32, device=dev)})
        L["flags"][_lib.FLAG_BUILDS] = builds   # the build count carries over a regrown list
        self.lt = L
        self.pair_out = torch.zeros(cap * 12, dtype=torch.float32, device=dev)
        nl = _lib.NList()
        nl.cap = cap
        nl.n_cells_total = self.n_cells
        for k in _lib._NL_PTRS:
            setattr(nl, k, L[k].data_ptr() if k in L else None)
        self.nl = nl
        self.cap = cap

    # ---------------------------------------------------------------- steps
    def _s(self, stream=None):
        return _lib.stream_handle(stream)

    def build_list(self, force=True, stream=None):
        _lib.check(_lib.lib.nepx_nlist_build(C.byref(self.st), C.byref(self.nl), int(force), self._s(stream)),
                   "nepx_nlist_build")

    def flags(self):
        return self.lt["flags"].cpu().numpy()

    def ensure_capacity(self):
        """Synchronising check after a build: grow the list if it overflowed or is nearly full."""
        f = self.flags()
        pl = int(f[_lib.FLAG_PL])
        if f[_lib.FLAG_MAXROW] > _lib.MAX_ROW:
            sysid = int(f[_lib.FLAG_LONGSYS]) - 1
            raise _lib.RowLimitError(f"a particle of system {sysid} has {int(f[_lib.FLAG_MAXROW])} neighbours in "
                                     f"the list radius (maximum {_lib.MAX_ROW})", sysid)
        if f[_lib.FLAG_OVERFLOW] or pl > 0.9 * self.cap:
            self._alloc_list(int(1.5 * max(pl, self.cap)) + 4096)
            self.build_list(True)
            f = self.flags()
            if f[_lib.FLAG_OVERFLOW]:
                raise _lib.NepxError(f"neighbour list overflow persists (flags {f.tolist()})")
        return int(self.flags()[_lib.FLAG_PL])

    def init(self, stream=None, update_thermostat=True):
        """First list build (with capacity check), initial kinetic energies and, unless the
        thermostat state was restored from a saved run, the first thermostat update."""
        self.build_list(True, stream)
        self.ensure_capacity()
        self.nvt_init(stream, update_thermostat)

    def step(self, stream=None):
        """One MD step, device only: [list rebuild if the skin is used up] -> filter -> pair kernels ->
        gather -> NVT (Sim.step, Sim.py:257-285)."""
        self.build_list(False, stream)
        self.filter_pairs(stream)
        self.pair_fused(stream)
        self.gather(stream)
        self.nvt_step(stream)
        self.steps += 1

    def recover_overflow(self, stream=None):
        """Synchronising: if the list overflowed (the integrator then skips its steps, so the state
        is still that of the last good step), grow the list and rebuild it.  Returns True when an
        overflow was handled; the caller recaptures any graph of the step (the buffers moved)."""
        if stream is not None:
            stream.synchronize()
        f = self.flags()
        code = int(f[_lib.FLAG_OVERFLOW])
        if code == 0:
            return False
        if code != 1:
            raise _lib.NepxError(f"inconsistent neighbour list (code {code}, flags {f.tolist()})")
        need = max(int(f[_lib.FLAG_PL]), self.cap)
        self._alloc_list(int(1.5 * need) + 4096)
        self.build_list(True)
        self.ensure_capacity()
        return True

    def device_steps(self, k=0):
        """Steps system k has completed on the device (skipped steps are not counted)."""
        return int(self.thermo()[k]["steps"])

    def faults(self):
        """Systems whose state went non-finite; the device stopped each at that step."""
        if not self.specs:
            return []
        th = self.thermo()
        return [k for k in range(len(self.specs)) if th[k]["fault"] != 0]

    def check(self):
        """Raise on a list overflow or on a system stopped by a non-finite state; returns the flags."""
        f = self.flags()
        if f[_lib.FLAG_OVERFLOW]:
            raise _lib.NepxError(f"neighbour list overflow (flags {f.tolist()}); results since the last "
                                 "check are invalid")
        bad = self.faults()
        if bad:
            th = self.thermo()
            raise _lib.NepxError("non-finite state: system(s) " +
                                 ", ".join(f"{k} (stopped after {int(th[k]['steps'])} steps)" for k in bad))
        return f

    # ----------------------------------------------------------- components
    def filter_pairs(self, stream=None):
        _lib.check(_lib.lib.nepx_pairs_filter(C.byref(self.st), C.byref(self.nl), self._s(stream)),
                   "nepx_pairs_filter")

    def n_pairs(self):
        return int(self.lt["flags"][_lib.FLAG_P].item())

    def pairs(self):
        P = self.n_pairs()
        return torch.stack([self.lt["pair_i"][:P], self.lt["pair_j"][:P]], dim=1)

    def gather(self, stream=None):
        _lib.check(_lib.lib.nepx_forces_gather(C.byref(self.st), C.byref(self.nl), _lib.ptr(self.pair_out),
                                               self._s(stream)), "nepx_forces_gather")

    def pair_fused(self, stream=None):
        """Per-pair forces/torques/energies of the compacted pairs.

        One fused kernel when an instance exists for this (shape, hypers,
        network); otherwise the three HIP kernels descriptor -> network ->
        adjoint with G and dE/dG in HBM.
        """
        np_dev = C.c_void_p(self.lt["flags"].data_ptr() + 4 * _lib.FLAG_P)
        args = (_lib.ptr(self.t["qf"]), _lib.ptr(self.lt["pair_i"]), _lib.ptr(self.lt["pair_j"]),
                _lib.ptr(self.lt["pair_t"]), np_dev, self.cap)
        if self.potential is not None:
            self.potential.pair_forces(*args, _lib.ptr(self.pair_out), self.clip, stream)
            return
        self.prepare_pair()
        if not self.split:
            _lib.check(_lib.lib.nepx_pair_fused_list(C.byref(self.model), C.byref(self.cnet), C.byref(self.st),
                                                     C.byref(self.nl), self.clip, _lib.ptr(self.pair_out),
                                                     self._s(stream)), "nepx_pair_fused_list")
            return
        s = self._s(stream)
        _lib.check(_lib.lib.nepx_descriptor_fwd(C.byref(self.model), *args, _lib.ptr(self._G), s),
                   "nepx_descriptor_fwd")
        _lib.check(_lib.lib.nepx_ffnn_fwd_bwd(C.byref(self.cnet), _lib.ptr(self._G), np_dev, self.cap,
                                              _lib.ptr(self._E), _lib.ptr(self._dG), s), "nepx_ffnn_fwd_bwd")
        _lib.check(_lib.lib.nepx_descriptor_adjoint(C.byref(self.model), *args, _lib.ptr(self._dG),
                                                    _lib.ptr(self._E), self.clip, _lib.ptr(self.pair_out), s),
                   "nepx_descriptor_adjoint")

    def prepare_pair(self):
        """Choose fused or split pair kernels and allocate the split path's buffers (no launches)."""
        if self.potential is not None:
            return
        if getattr(self, "split", None) is None:
            self.split = not _lib.lib.nepx_pair_fused_supported(C.byref(self.model), C.byref(self.cnet))
        if self.split and (getattr(self, "_G", None) is None or self._G.numel() < self.cap * self.D):
            self._G = torch.empty(self.cap * self.D, dtype=torch.float32, device=self.device)
            self._dG = torch.empty(self.cap * self.D, dtype=torch.float32, device=self.device)
            self._E = torch.empty(self.cap, dtype=torch.float32, device=self.device)

    def nvt_step(self, stream=None):
        _lib.check(_lib.lib.nepx_nvt_step(C.byref(self.st), C.byref(self.nl), self._s(stream)), "nepx_nvt_step")

    def nvt_part(self, part, stream=None):
        _lib.check(_lib.lib.nepx_nvt_part(C.byref(self.st), C.byref(self.nl), int(part), self._s(stream)),
                   "nepx_nvt_part")

    def nvt_init(self, stream=None, update_thermostat=True):
        _lib.check(_lib.lib.nepx_nvt_init(C.byref(self.st), C.byref(self.nl), int(bool(update_thermostat)),
                                          self._s(stream)), "nepx_nvt_init")

    def load_pairs(self, pairs, stream=None):
        """Use an explicit (M, 2) candidate list (any order) as the list rows."""
        pairs = pairs.to(device=self.device, dtype=torch.int64).contiguous()
        m = pairs.shape[0]
        if m + 1024 > self.cap:
            self._alloc_list(m + 4096)
        _lib.check(_lib.lib.nepx_nlist_from_pairs(C.byref(self.st), C.byref(self.nl), _lib.ptr(pairs), m,
                                                  self._s(stream)), "nepx_nlist_from_pairs")
        f = self.flags()
        if f[_lib.FLAG_OVERFLOW]:
            raise _lib.NepxError(f"bad candidate pair list (code {int(f[_lib.FLAG_OVERFLOW])}): indices out of "
                                 "range, i == j, or a row longer than 2048")

    def set_configuration(self, x=None, q=None, k=0):
        """Overwrite positions (N,3) and/or orientations (N,4) of system k from tensors/arrays."""
        a, b = self.starts[k], self.starts[k] + self.specs[k].n
        dev = self.device
        if x is not None:
            x = torch.as_tensor(x).to(device=dev, dtype=torch.float64)
            for c, key in enumerate("xyz"):
                self.t[key][a:b].copy_(x[:, c])
        if q is not None:
            q = torch.as_tensor(q).to(device=dev, dtype=torch.float64)
            for c in range(4):
                self.t[f"q{c}"][a:b].copy_(q[:, c])
            self.t["qf"].view(-1, 4)[a:b].copy_(q.to(torch.float32))

    def set_forces(self, F, T, k=0):
        a, b = self.starts[k], self.starts[k] + self.specs[k].n
        F = torch.as_tensor(F).to(device=self.device, dtype=torch.float64)
        T = torch.as_tensor(T).to(device=self.device, dtype=torch.float64)
        for c, key in enumerate("xyz"):
            self.t["f" + key][a:b].copy_(F[:, c])
            self.t["t" + key][a:b].copy_(T[:, c])

    def forces(self, k=0):
        a, b = self.starts[k], self.starts[k] + self.specs[k].n
        F = torch.stack([self.t["fx"][a:b], self.t["fy"][a:b], self.t["fz"][a:b]], 1)
        T = torch.stack([self.t["tx"][a:b], self.t["ty"][a:b], self.t["tz"][a:b]], 1)
        return F, T

    def tensor(self, keys, k=0):
        a, b = self.starts[k], self.starts[k] + self.specs[k].n
        return torch.stack([self.t[key][a:b] for key in keys], 1)

    def set_early_rebuild(self, fraction, stream=None):
        """Request list rebuilds at the next build point once a particle has moved `fraction` of half
        the skin (0 turns this off; the list is rebuilt when it goes stale).  For multi-step graphs,
        whose only build point is their first step: with an early request the list rarely goes stale
        inside a graph, where the remaining steps would be frozen."""
        rec = self.t["sp"].cpu().numpy().view(_lib.SYSPARAM_DTYPE).copy()
        rec["pad"][:, 0] = float(fraction)
        with torch.cuda.stream(stream) if stream is not None else _nullctx():
            self.t["sp"].copy_(torch.from_numpy(rec.view(np.uint8).copy()))

    # --------------------------------------------------------------- access
    def thermo(self):
        return self.t["th"].cpu().numpy().view(_lib.THERMO_DTYPE).copy()

    def system_arrays(self, k=0):
        """FP64 host copies of system k: x (N,3), v, q (N,4), p, F, T, upot."""
        a, b = self.starts[k], self.starts[k] + self.specs[k].n
        g = lambda key: self.t[key][a:b].cpu().numpy()
        return dict(
            x=np.stack([g("x"), g("y"), g("z")], 1), v=np.stack([g("vx"), g("vy"), g("vz")], 1),
            q=np.stack([g("q0"), g("q1"), g("q2"), g("q3")], 1),
            p=np.stack([g("p0"), g("p1"), g("p2"), g("p3")], 1),
            F=np.stack([g("fx"), g("fy"), g("fz")], 1), T=np.stack([g("tx"), g("ty"), g("tz")], 1),
            upot=g("upot"), mass=g("mass"))

    def pair_outputs(self):
        P = self.n_pairs()
        return self.pair_out[: 12 * P].view(P, 12)
