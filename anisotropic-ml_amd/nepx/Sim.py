"""Sim with the reference call surface (nep_simulator/Sim.py:12-285).

The state lives on the GPU in FP64 structure-of-arrays (system.Arena); the
step runs as HIP kernels with no host synchronisation.  Setters, `run`,
`step`, `dumpConfig` and the timers keep the reference's names and order of
use (sim_input.py:50-61).  Differences, by design:
  * the neighbour list is a GPU cell list rebuilt when any particle has moved
    more than half the skin since the last build (instead of a host KD-tree
    every 100 steps); `refreshNList()` forces a rebuild;
  * forces come from the hand-written adjoint (fast path) instead of forward
    differences; set `fd_compat = True` before `run` to evaluate the
    reference's FD stencil through DescriptorGenerator/ForceEvaluator;
  * the kinetic energy used by the first thermostat update is the sum over
    particles (the reference reads particle 0's only, Sim.py:157);
  * dumps also store the thermostat state so a run resumes exactly.
"""
import time

import numpy as np
import torch

from . import _lib, gsd
from .DescriptorGenerator import DescriptorGenerator
from .ForceEvaluator import ForceEvaluator
from .GsdHandler import GsdHandler
from .engine import StepRunner
from .system import Arena, SystemSpec
from . import MathUtils as mu


class Sim:
    def __init__(self):
        print("A new simulation instance created.")
        self.fd_compat = False
        self.component_path = False   # drive DescriptorGenerator/ForceEvaluator/NVT methods each step
        self.use_graph = True
        self._nlist_cache = (None, None)
        self.Nlist_skin = 0.5
        self.nn_width, self.nn_depth = 150, 3      # Sim.py:60
        self.activation = "relu"
        self.arena = None
        self.runner = None
        self.gsdDumpPeriod = 0
        self.gsdFilename = None
        self.thermo_init = None
        self.set_timers()

    def setDevice(self, gpu_id):
        self.gpu_id = int(gpu_id)
        self.device = torch.device(f"cuda:{self.gpu_id}")
        torch.cuda.set_device(self.device)

    def initializeSystem(self, gsd_file):
        """Read the last frame of a GSD file (GsdHandler.py:9-69); COMs are typeid 0."""
        print("Simulation will be initialized from %s" % gsd_file)
        h = GsdHandler(gsd_file)
        self.initializeFromArrays(h.get_central_pos(), h.getOrientations(), h.get_COM_velocities(), h.getAngmom(),
                                  h.getMass(), h.getMoi(), h.getLx())
        self.thermo_init = h.thermostat

    def initializeFromArrays(self, pos, orientations, velocities, angmom, mass, moi, Lx):
        self.central_pos_h = np.asarray(pos, dtype=np.float64)
        self.Nparticles = len(self.central_pos_h)
        self.N_dof = (self.Nparticles - 1) * 3
        self.RN_dof = self.Nparticles * 3
        self.velocities_h = np.asarray(velocities, dtype=np.float64)
        self.mass_h = np.asarray(mass, dtype=np.float64)
        self.orientations_h = np.asarray(orientations, dtype=np.float64)
        self.angmom_h = np.asarray(angmom, dtype=np.float64)
        self.moi_to_dump = np.asarray(moi, dtype=np.float32)
        self.moi = np.asarray(moi, dtype=np.float64).reshape(-1, 3)[0]
        self.Lx = float(Lx)
        self.dx = 0.00001
        self.dteta = 0.00001

    def setEvaluator(self, model_path=None, net=None):
        self.evaluator = ForceEvaluator()
        self.evaluator.setDevice(self.gpu_id)
        self.evaluator.setNN_dimensions(self.nn_width, self.nn_depth)
        self.evaluator.setNd(self.N_descriptors)
        self.evaluator.setActivation(self.activation)
        if net is not None:
            self.evaluator.set_model(net)
        else:
            self.evaluator.read_model(model_path)
        self.evaluator.set_dx_dteta(self.dx, self.dteta)
        self.evaluator.setNparticles(self.Nparticles)

    def setNepDescriptors(self, hypers):
        self.descriptor_generator = DescriptorGenerator()
        self.descriptor_generator.setDevice(self.gpu_id)
        self.descriptor_generator.setHyperParameters(hypers)
        self.descriptor_generator.setBoxSize(self.Lx)
        self.N_descriptors = self.descriptor_generator.getN_descriptor()
        self.descriptor_generator.set_dx_dteta(self.dx, self.dteta)

    def setShape(self, shape_str):
        self.descriptor_generator.setShape(shape_str)
        en_min, en_max = self.descriptor_generator.get_en_min_max()
        self.evaluator.set_energy_range(en_min, en_max)
        self.cutoff = self.descriptor_generator.cutoff
        self.evaluator.move_data_to_device()
        self.descriptor_generator.move_data_to_device()
        self.evaluator.attach(self.descriptor_generator)

    def setNeighborList(self):
        """List radius = COM cutoff + skin; rebuilt on the GPU when the skin is used up."""
        self.Nlist_cutoff = self.cutoff + self.Nlist_skin
        self.Nlist_freq = 100   # kept for reference scripts; rebuilds are displacement-triggered

    def setGsdDump(self, dump_period, filename):
        self.gsdDumpPeriod = int(dump_period)
        self.gsdFilename = filename

    def setkT(self, kT):
        """Thermostat initial values as in the reference (Sim.py:99-111) unless the init file carried them."""
        self.kT = float(kT)
        self.integrator_vars0 = -0.0317711
        self.integrator_vars1 = 0.108609
        self.integrator_vars2 = 0.00814227
        self.integrator_vars3 = -0.014238
        self.m_exp_thermo_fac = 1.00011

    def setdt(self, dt):
        self.dt = float(dt)
        self.tau = 100.0 * self.dt

    # ------------------------------------------------------------------ state
    def move_data_to_device(self):
        if self.arena is not None:
            return
        thermo = self.thermo_init or dict(xi=self.integrator_vars0, eta=self.integrator_vars1,
                                          xi_r=self.integrator_vars2, eta_r=self.integrator_vars3)
        spec = SystemSpec(self.central_pos_h, self.velocities_h, self.orientations_h, self.angmom_h, self.mass_h,
                          self.Lx, self.kT, self.dt, tuple(self.moi), self.cutoff, skin=self.Nlist_skin,
                          tau=self.tau, thermo=thermo, ndof=self.N_dof, rndof=self.RN_dof)
        dg = self.descriptor_generator
        self.arena = Arena([spec], dg.shape, (dg.nrad, dg.nang, dg.lmax), self.evaluator.net, device=self.device)
        self.runner = StepRunner(self.arena, use_graph=self.use_graph and not (self.fd_compat or self.component_path))
        self.runner.init(update_thermostat=self.thermo_init is None)
        self.rot_kin_en = self.getRotKin()

    @property
    def central_pos(self):
        return self.arena.tensor(["x", "y", "z"])

    @property
    def orientations(self):
        return self.arena.tensor(["q0", "q1", "q2", "q3"])

    @property
    def velocities(self):
        return self.arena.tensor(["vx", "vy", "vz"])

    @property
    def angmom(self):
        return self.arena.tensor(["p0", "p1", "p2", "p3"])

    @property
    def mass(self):
        return self.arena.tensor(["mass"])[:, 0]

    @property
    def Nlist(self):
        """Current list rows as an (M, 2) int64 tensor (i < j), like the reference's KD-tree pairs.
        The same tensor is returned until the list is rebuilt."""
        f = self.arena.flags()
        builds, cached = self._nlist_cache
        if builds != int(f[_lib.FLAG_BUILDS]) or cached is None:
            PL = int(f[_lib.FLAG_PL])
            cached = torch.stack([self.arena.lt["slot_row"][:PL], self.arena.lt["nbr"][:PL]], 1).to(torch.int64)
            self._nlist_cache = (int(f[_lib.FLAG_BUILDS]), cached)
        return cached

    def getRotKin(self):
        q = self.orientations
        conj_q = q.clone()
        conj_q[:, 1:] = -conj_q[:, 1:]
        s = 0.5 * mu.quaternion_multiplication(conj_q, self.angmom)
        moi = torch.as_tensor(self.moi, dtype=s.dtype, device=s.device)
        return float((0.5 * torch.sum(s[:, 1:] ** 2 / moi, dim=1)).sum())

    def thermostat(self):
        th = self.arena.thermo()[0]
        return {k: float(th[k]) for k in ("xi", "eta", "xi_r", "eta_r", "theta", "e_r", "Kt", "Kr", "U")}

    def dumpConfig(self, ts):
        """Append one GSD frame (Sim.py:117-144) plus the thermostat state."""
        a = self.arena.system_arrays(0)
        th = self.thermostat()
        N = self.Nparticles
        frame = {
            "configuration/step": np.array([ts], dtype=np.uint64),
            "configuration/box": np.array([self.Lx, self.Lx, self.Lx, 0, 0, 0], dtype=np.float32),
            "particles/N": np.array([N], dtype=np.uint32),
            "particles/types": np.array([[65, 0], [66, 0]], dtype=np.uint8),
            "particles/typeid": np.zeros(N, dtype=np.uint32),
            "particles/position": a["x"].astype(np.float32),
            "particles/orientation": a["q"].astype(np.float32),
            "particles/velocity": a["v"].astype(np.float32),
            "particles/angmom": a["p"].astype(np.float32),
            "particles/mass": a["mass"].astype(np.float32),
            "particles/moment_inertia": np.broadcast_to(self.moi_to_dump.reshape(-1, 3), (N, 3)).astype(np.float32),
            "log/nepx/thermostat": np.array([th["xi"], th["eta"], th["xi_r"], th["eta_r"]], dtype=np.float64),
            "log/nepx/position_f64": a["x"],
            "log/nepx/orientation_f64": a["q"],
            "log/nepx/velocity_f64": a["v"],
            "log/nepx/angmom_f64": a["p"],
        }
        if self._writer is None:
            self._writer = gsd.GsdWriter(self.gsdFilename, "w" if ts == 0 else "a")
        self._writer.append(frame)

    def print_performance_info(self):
        if self.timestep > 0:
            total = time.time() - self.timer
            epoch = time.time() - self.epoch_timer
            print("TPS : %.2f   - last TPS : %.2f" % (self.timestep / total, self.gsdDumpPeriod / max(epoch, 1e-9)))
            self.epoch_timer = time.time()

    def refreshNList(self):
        self.arena.build_list(True, self.runner.stream)

    def set_timers(self):
        self.t_run = 0.0
        self.t_gen = 0.0
        self.t_eval = 0.0
        self.t_s1 = 0.0
        self.t_s2 = 0.0
        self.t_nlist = 0.0

    # ------------------------------------------------------------------ loop
    def run(self, Nsteps):
        self.is_sync = 0
        self.set_timers()
        self._writer = None
        self.move_data_to_device()
        self.timestep = 0
        self.timer = time.time()
        self.epoch_timer = time.time()
        check_every = 100
        base = self.arena.device_steps()
        dumping = bool(self.gsdDumpPeriod and self.gsdFilename)
        last_dump = -1
        while self.timestep < Nsteps:
            if dumping and self.timestep % self.gsdDumpPeriod == 0 and self.timestep != last_dump:
                # A list overflow (or a stale list inside a multi-step graph) freezes the state at its
                # last good step while the host keeps counting; settle it before writing so every
                # frame's label is the device's count.
                torch.cuda.synchronize(self.device)
                self._recover(base)
                if self.timestep % self.gsdDumpPeriod == 0 and self.timestep != last_dump:
                    self.print_performance_info()
                    self.dumpConfig(self.timestep)
                    last_dump = self.timestep
                continue
            # steps up to the next dump point, check point or the end, launched together
            nxt = min(Nsteps, (self.timestep // check_every + 1) * check_every)
            if dumping:
                nxt = min(nxt, (self.timestep // self.gsdDumpPeriod + 1) * self.gsdDumpPeriod)
            self._steps(nxt - self.timestep)
            if self.timestep % check_every == 0 or self.timestep == Nsteps:
                self._recover(base)
        torch.cuda.synchronize(self.device)
        self.arena.check()
        self.t_run = time.time() - self.timer
        print("Timers:")
        if self.fd_compat or self.component_path:
            for k in ("t_nlist", "t_gen", "t_eval", "t_s1", "t_s2"):
                print("%s : %.2f" % (k, getattr(self, k)))
        else:
            # the fast path is one device-only captured step: list test, filter, pair kernel, gather
            # and NVT are not separated by host synchronisations, so only the whole run is timed
            print("t_run : %.2f (%d steps, captured device-only steps)" % (self.t_run, Nsteps))
        print("Done")

    def _recover(self, base):
        """Synchronising: a list overflow makes the device skip steps (grow the list), a stale list
        inside a multi-step graph freezes them; resume from the device's step count."""
        if self.runner is not None:
            self.runner.recover()
        else:
            self.arena.recover_overflow()
        self.timestep = self.arena.device_steps() - base
        self.arena.check()

    def _steps(self, m):
        """Launch m steps (the host count advances by m; _recover reconciles it with the device)."""
        if self.fd_compat or self.component_path:
            for _ in range(m):
                self.step()
            return
        self.runner.run(m)
        self.timestep += m

    def step(self):
        self.descriptor_generator.fd_compat = self.fd_compat
        if not (self.fd_compat or self.component_path):
            self.runner.step()
            self.timestep += 1
            return
        # reference structure: descriptors -> evaluator -> step two -> step one
        self.arena.build_list(False)   # runs only when the skin is used up
        t0 = time.time()
        g, pp, n_pair = self.descriptor_generator.generate_nep_descriptors(self.central_pos, self.orientations,
                                                                           self.Nlist)
        t1 = time.time()
        self.forces, self.torks = self.evaluator.evaluate_interactions(g, pp, n_pair)
        self.arena.set_forces(self.forces, self.torks)
        t2 = time.time()
        self.integrate_step_two()
        t3 = time.time()
        self.integrate_step_one()
        t4 = time.time()
        self.t_gen += t1 - t0
        self.t_eval += t2 - t1
        self.t_s1 += t3 - t2
        self.t_s2 += t4 - t3
