"""Continuous batching of MD jobs on one GPU.

SlotArena: a fixed-capacity particle arena whose systems come and go.  Every
capacity (particles, system slots, list slots, cells) is fixed at creation,
so the hipGraphs captured for the step stay valid while jobs are admitted
and retired: the kernels read each system's parameters, activity flag and
step budget from device records.  A job is admitted into a free
block-aligned particle range when its memory footprint fits; its first
thermostat update runs on the device for that system only; it advances with
every other resident system of the same pair model, deactivates itself when
its step budget is reached (device side, no host round trip) and is
harvested between chunks of steps.

Server: per pair-model class one SlotArena on its own stream; jobs arrive
over time (e.g. Poisson), wait in a FIFO queue until they fit, and the
classes advance concurrently in chunks of hipGraph-replayed steps.  Per-step
latency comes from events recorded at every step boundary.

run_plain: the same jobs one at a time with the single-system step, for the
served-versus-plain comparison.
"""
import collections
import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .engine import StepRunner
from .model import EnergyNet
from .shapes import Shape
from .system import BLOCK, Arena, SystemSpec, _align


@dataclass
class Job:
    jid: int
    cls: str               # pair-model class (shape name)
    spec: SystemSpec
    steps: int
    arrival: float = 0.0   # seconds after the start of the run
    # filled by the server
    slot: int = -1
    admitted_at: float = -1.0
    done_at: float = -1.0
    steps_done: int = 0
    result: dict = field(default_factory=dict)
    cancelled: bool = False
    error: str = ""        # set when the server stops the job (e.g. its system collapsed)

    @property
    def n(self):
        return self.spec.n


class LazySpec:
    """A seeded system spec whose particle arrays are built on first use.  A long job trace
    holds thousands of jobs, of which only those that reach the head of a queue need their
    box and arrays; `release` drops the arrays again (they are rebuilt, identically, from the
    seed if asked for later)."""

    def __init__(self, shape_name, n, phi, **kw):
        self._args = (shape_name, n, phi)
        self._kw = kw
        self._spec = None
        self.n = n

    def build(self):
        from .configs import make_system
        if self._spec is None:
            self._spec = make_system(*self._args, **self._kw)
        return self._spec

    def release(self):
        self._spec = None

    def __getattr__(self, name):  # attributes other than n come from the built spec
        if name.startswith("_"):
            raise AttributeError(name)
        return getattr(self.build(), name)


def _spec_of(spec):
    return spec.build() if isinstance(spec, LazySpec) else spec


class _Intervals:
    """First-fit allocator over [0, size) in integer units."""

    def __init__(self, size):
        self.free = [(0, size)]

    def alloc(self, m):
        for k, (a, b) in enumerate(self.free):
            if b - a >= m:
                if b - a == m:
                    self.free.pop(k)
                else:
                    self.free[k] = (a + m, b)
                return a
        return None

    def release(self, a, m):
        self.free.append((a, a + m))
        self.free.sort()
        merged = []
        for x, y in self.free:
            if merged and merged[-1][1] == x:
                merged[-1] = (merged[-1][0], y)
            else:
                merged.append((x, y))
        self.free = merged


class SlotArena(Arena):
    def __init__(self, shape: Shape, hypers, net: EnergyNet, capacity, max_systems, list_cap=None,
                 cell_cap=1 << 20, device="cuda", clip=100.0):
        self._capacity = _align(capacity)
        self._max_systems = int(max_systems)
        self._cell_cap = int(cell_cap)
        self.free_blocks = _Intervals(self._capacity // BLOCK)
        self.free_cells = _Intervals(self._cell_cap)
        self.free_slots = list(range(self._max_systems))
        self.jobs = {}        # slot -> Job
        self.pair_est = {}    # slot -> estimated list pairs
        if list_cap is None:
            list_cap = 200 * self._capacity   # half-list slots per particle at the densest supported packing
        super().__init__([], shape, hypers, net, device=device, clip=clip, list_cap=list_cap)
        # device bytes per particle, per list slot and per cell of this arena's buffers
        n, cap = self.n, self.cap
        self.particle_bytes = (sum(v.numel() * v.element_size() for k, v in self.t.items()
                                   if k not in ("sp", "th", "partials")) +
                               sum(self.lt[k].numel() * self.lt[k].element_size()
                                   for k in ("cell_list", "cell_of", "cell_pos", "row_count", "row_start",
                                             "rrow_count", "rrow_start"))) / n
        self.slot_bytes = (sum(self.lt[k].numel() * self.lt[k].element_size()
                               for k in ("nbr", "slot_row", "rnbr", "cidx", "pair_i", "pair_j", "pair_t", "rpos",
                                         "pair_e", "rrec") if k in self.lt) +
                           self.pair_out.numel() * self.pair_out.element_size()) / cap
        self.cell_bytes = sum(self.lt[k].numel() * self.lt[k].element_size()
                              for k in ("cell_count", "cell_start", "cell_fill")) / (self.n_cells + 1)

    def device_bytes(self):
        """Bytes of device memory the arena holds (state, list, records, per-pair outputs)."""
        return (sum(v.numel() * v.element_size() for v in self.t.values()) +
                sum(v.numel() * v.element_size() for v in self.lt.values()) +
                self.pair_out.numel() * self.pair_out.element_size())

    def job_bytes(self, spec):
        """Device bytes a job occupies while resident: its block-aligned particle range, its share of
        list slots (estimated pairs within the list radius) and its cells."""
        nblk, pairs = self.footprint(spec)
        r_list = spec.rc + spec.skin
        nc = int(math.floor(spec.L / r_list))
        ncell = nc ** 3 if nc >= 3 else 0
        return int(nblk * BLOCK * self.particle_bytes + pairs * self.slot_bytes + ncell * self.cell_bytes)

    def _layout(self):
        self.n = self._capacity
        self.starts = []
        self.sysrec = np.zeros(self._max_systems, dtype=_lib.SYSPARAM_DTYPE)
        self.n_cells = max(self._cell_cap, 2 * self.n)

    # ------------------------------------------------------------ admission
    def footprint(self, spec):
        r_list = spec.rc + spec.skin
        rho = spec.n / spec.L ** 3
        pairs = 0.5 * spec.n * min(spec.n - 1, rho * 4.0 / 3.0 * math.pi * r_list ** 3) * 1.3 + 256
        return _align(spec.n) // BLOCK, pairs

    def fits_ever(self, job: Job):
        """Whether the job could be placed in the empty arena (else it is rejected, not queued)."""
        nblk, pairs = self.footprint(job.spec)
        r_list = job.spec.rc + job.spec.skin
        nc = int(math.floor(job.spec.L / r_list))
        return nblk * BLOCK <= self._capacity and pairs <= 0.9 * self.cap and (nc < 3 or nc ** 3 <= self._cell_cap)

    def admit(self, job: Job):
        """Place the job if it fits; returns True on success. Call between steps (stream idle)."""
        spec = job.spec
        if not self.free_slots:
            return False
        nblk, pairs = self.footprint(spec)
        if sum(self.pair_est.values()) + pairs > 0.9 * self.cap:
            return False
        r_list = spec.rc + spec.skin
        nc = int(math.floor(spec.L / r_list))
        nc = nc if nc >= 3 else 0
        b0 = self.free_blocks.alloc(nblk)
        if b0 is None:
            return False
        c0 = self.free_cells.alloc(nc ** 3) if nc else 0
        if c0 is None:
            self.free_blocks.release(b0, nblk)
            return False
        slot = self.free_slots.pop(0)
        a = b0 * BLOCK
        b = a + spec.n
        dev = self.device
        for c, key in enumerate("xyz"):
            self.t[key][a:b].copy_(torch.from_numpy(spec.x[:, c]))
            self.t["v" + key][a:b].copy_(torch.from_numpy(spec.v[:, c]))
        for c in range(4):
            self.t[f"q{c}"][a:b].copy_(torch.from_numpy(spec.q[:, c]))
            self.t[f"p{c}"][a:b].copy_(torch.from_numpy(spec.p[:, c]))
        self.t["mass"][a:b].copy_(torch.from_numpy(spec.mass))
        self.t["sys"][a:b].fill_(slot)
        r = np.zeros(1, dtype=_lib.SYSPARAM_DTYPE)[0]
        r["L"], r["kT"], r["dt"], r["tau"] = spec.L, spec.kT, spec.dt, spec.tau
        r["ndof"], r["rndof"] = spec.ndof, spec.rndof
        r["moi"] = spec.moi
        r["r_list"], r["rc"], r["skin"] = r_list, spec.rc, spec.skin
        r["p_start"], r["p_end"] = a, b
        r["blk_start"], r["blk_end"] = b0, b0 + nblk
        r["nc"], r["cell_off"] = nc, c0
        r["active"] = 1
        th = np.zeros(1, dtype=_lib.THERMO_DTYPE)[0]
        for key, val in spec.thermo.items():
            th[key] = val
        th["theta"] = th["e_r"] = 1.0
        th["budget"] = float(job.steps)
        th["pending"] = 1.0
        self._put_record("sp", slot, r)
        self._put_record("th", slot, th)
        self.jobs[slot] = job
        self.pair_est[slot] = pairs
        self.blk_of = getattr(self, "blk_of", {})
        self.blk_of[slot] = (b0, nblk, c0, nc ** 3)
        job.slot = slot
        if isinstance(spec, LazySpec):
            spec.release()  # the state now lives on the device
        return True

    def _put_record(self, key, slot, rec):
        raw = np.frombuffer(rec.tobytes(), dtype=np.uint8)
        sz = len(raw)
        self.t[key][slot * sz:(slot + 1) * sz].copy_(torch.from_numpy(raw.copy()))

    def commit_admissions(self, stream=None):
        """Initial kinetic energies + first thermostat update of the newly admitted systems, then a
        forced list rebuild at the next step."""
        _lib.check(_lib.lib.nepx_nvt_init(C_byref(self.st), C_byref(self.nl), 2, self._s(stream)), "nepx_nvt_init")
        self.lt["flags"][_lib.FLAG_REBUILD:_lib.FLAG_REBUILD + 1].fill_(1)

    # ------------------------------------------------------------ retirement
    def records(self):
        sp = self.t["sp"].cpu().numpy().view(_lib.SYSPARAM_DTYPE)
        th = self.t["th"].cpu().numpy().view(_lib.THERMO_DTYPE)
        return sp, th

    def harvest(self, keep_state=True):
        """Retire systems that finished their budget; returns their jobs (with results)."""
        if not self.jobs:
            return []
        sp, th = self.records()
        done = []
        for slot, job in list(self.jobs.items()):
            job.steps_done = int(th[slot]["steps"])
            if sp[slot]["active"] == 0:
                a, b = int(sp[slot]["p_start"]), int(sp[slot]["p_end"])
                if th[slot]["fault"] != 0:
                    job.result = {"error": f"non-finite state after {job.steps_done} steps; the job was "
                                           "stopped at that step", "fault": True}
                elif job.error:
                    job.result = {"error": job.error}
                elif job.cancelled:
                    job.result = {"error": f"cancelled after {job.steps_done} steps"}
                elif keep_state:
                    job.result = dict(self.slot_state(sp[slot]),
                                      thermo={k: float(th[slot][k]) for k in ("xi", "eta", "xi_r", "eta_r",
                                                                              "Kt", "Kr", "U", "steps")})
                self.t["sys"][a:b].fill_(-1)
                self._put_record("sp", slot, np.zeros(1, dtype=_lib.SYSPARAM_DTYPE)[0])
                b0, nblk, c0, ncell = self.blk_of.pop(slot)
                self.free_blocks.release(b0, nblk)
                if ncell:
                    self.free_cells.release(c0, ncell)
                self.free_slots.append(slot)
                self.free_slots.sort()
                del self.jobs[slot]
                del self.pair_est[slot]
                done.append(job)
        if done:
            self.lt["flags"][_lib.FLAG_REBUILD:_lib.FLAG_REBUILD + 1].fill_(1)
        return done

    def slot_state(self, rec):
        """Host copies (FP64) of the positions, velocities, orientations and angular momenta of the
        system whose record is `rec`."""
        a, b = int(rec["p_start"]), int(rec["p_end"])
        g = lambda key: self.t[key][a:b].cpu().numpy()
        return dict(x=np.stack([g("x"), g("y"), g("z")], 1), v=np.stack([g("vx"), g("vy"), g("vz")], 1),
                    q=np.stack([g("q0"), g("q1"), g("q2"), g("q3")], 1),
                    p=np.stack([g("p0"), g("p1"), g("p2"), g("p3")], 1))

    def _deactivate(self, job: Job):
        sp, _ = self.records()
        rec = sp[job.slot].copy()
        rec["active"] = 0
        self._put_record("sp", job.slot, rec)

    def cancel(self, job: Job):
        """Stop a resident job at the current step (call between steps, stream idle); the next
        harvest retires it with a 'cancelled' result."""
        self._deactivate(job)
        job.cancelled = True

    def stop(self, job: Job, reason):
        """Stop a resident job with an error (call between steps, stream idle); the next harvest
        retires it with that error."""
        self._deactivate(job)
        job.error = reason

    def particle_steps(self):
        """Particle-steps done so far by the resident systems."""
        if not self.jobs:
            return 0
        _, th = self.records()
        return sum(job.n * int(th[slot]["steps"]) for slot, job in self.jobs.items())


def C_byref(x):
    import ctypes
    return ctypes.byref(x)


class _ClassRunner:
    def __init__(self, arena: SlotArena):
        self.arena = arena
        self.runner = StepRunner(arena, use_graph=True)
        self.runner.stream.synchronize()
        with torch.cuda.stream(self.runner.stream):
            arena.build_list(True, self.runner.stream)   # an empty list; real builds follow admissions
        self.runner.stream.synchronize()
        self.runner._capture()
        self.step_times = []

    def run_chunk(self, k):
        st = self.runner.stream
        evs = [torch.cuda.Event(enable_timing=True) for _ in range(k + 1)]
        evs[0].record(st)
        for i in range(k):
            self.runner.step()
            evs[i + 1].record(st)
        return evs


class Server:
    """Continuous batching: one SlotArena per pair-model class, classes on separate streams.

    Admission is by memory: a job is admitted only while the device bytes of all resident jobs
    (SlotArena.job_bytes: particle range, share of list slots, cells) plus its own stay within
    `mem_budget`, and it finds a free particle range, list share and cell range in its class's
    arena.  The arenas are allocated up front (arena_bytes, reported); their capacities are what keeps
    the GPU busy, since more resident particles than that only lengthen every step."""

    def __init__(self, classes, capacity, max_systems, chunk=20, device="cuda", mem_budget=200 * 2 ** 30):
        # classes: name -> (Shape, hypers, EnergyNet)
        self.device = torch.device(device)
        self.classes = {}
        for name, (shape, hyp, net) in classes.items():
            ar = SlotArena(shape, hyp, net, capacity, max_systems, device=self.device)
            self.classes[name] = _ClassRunner(ar)
        self.mem_budget = int(mem_budget)
        self.arena_bytes = sum(cr.arena.device_bytes() for cr in self.classes.values())
        self.resident_bytes = 0
        self.peak_resident_bytes = 0
        self.job_bytes = {}
        self.chunk = chunk
        self.recoveries = 0
        self._cancel = set()

    def cancel(self, jid):
        """Cancel a job by id: a waiting job is dropped, a resident one stops at the next chunk
        boundary; either way it retires with an error result."""
        self._cancel.add(jid)

    def _fits_budget(self, cr, job):
        b = cr.arena.job_bytes(job.spec)
        return self.resident_bytes + b <= self.mem_budget, b

    def _stop_collapsed(self, cr, err):
        """A job whose system collapsed (a particle with more neighbours than the list build
        sorts) is stopped with an error and the list rebuilt without it; the steps the overflow
        skipped were skipped by every job of the class alike, so the others are unaffected."""
        job = cr.arena.jobs.get(err.system)
        if job is None:
            raise err
        st = cr.runner.stream
        with torch.cuda.stream(st):
            cr.arena.stop(job, f"the system collapsed: {err}")
            cr.arena.build_list(True, st)
        st.synchronize()
        cr.arena.ensure_capacity()
        cr.runner._capture()

    def _apply_cancellations(self, queue, waiting, rejected):
        if not self._cancel:
            return
        for dq in [queue] + list(waiting.values()):
            for j in [j for j in dq if j.jid in self._cancel]:
                dq.remove(j)
                j.cancelled = True
                j.result = {"error": "cancelled before admission"}
                rejected.append(j)
                self._cancel.discard(j.jid)
        for cr in self.classes.values():
            hit = [j for j in cr.arena.jobs.values() if j.jid in self._cancel]
            if hit:
                with torch.cuda.stream(cr.runner.stream):
                    for j in hit:
                        cr.arena.cancel(j)
                        self._cancel.discard(j.jid)

    def run(self, jobs, time_limit=None, log=None, on_chunk=None):
        """Serve the jobs (arrival times relative to the start) until all are done or the time
        limit passes.  on_chunk(server, seconds), if given, is called after every round of chunks
        (streams idle): the place to cancel jobs or inspect the arenas.  A job whose state goes
        non-finite is stopped by the device at that step and retired with an error result;
        the other jobs are unaffected."""
        queue = collections.deque(sorted(jobs, key=lambda j: j.arrival))
        waiting = {name: collections.deque() for name in self.classes}
        finished = []
        failed = []       # stopped on a non-finite state, or cancelled while resident
        rejected = []
        lat = []
        t0 = time.perf_counter()
        retired_ps = 0
        last_log = 0.0
        while True:
            now = time.perf_counter() - t0
            if time_limit is not None and now >= time_limit:
                break
            while queue and queue[0].arrival <= now:
                j = queue.popleft()
                waiting[j.cls].append(j)
            self._apply_cancellations(queue, waiting, rejected)
            # admission (FIFO per class; a job that does not fit blocks the ones behind it); all device
            # work of a class, copies included, stays on that class's stream
            for name, cr in self.classes.items():
                while waiting[name] and (not cr.arena.fits_ever(waiting[name][0]) or
                                         cr.arena.job_bytes(waiting[name][0].spec) > self.mem_budget):
                    j = waiting[name].popleft()
                    j.result = {"error": "job exceeds the arena capacity or the memory budget"}
                    rejected.append(j)
                with torch.cuda.stream(cr.runner.stream):
                    admitted = False
                    while waiting[name]:
                        ok, b = self._fits_budget(cr, waiting[name][0])
                        if not ok or not cr.arena.admit(waiting[name][0]):
                            break
                        j = waiting[name].popleft()
                        j.admitted_at = now
                        self.job_bytes[j.jid] = b
                        self.resident_bytes += b
                        self.peak_resident_bytes = max(self.peak_resident_bytes, self.resident_bytes)
                        admitted = True
                    if admitted:
                        cr.arena.commit_admissions(cr.runner.stream)
            busy = [cr for cr in self.classes.values() if cr.arena.jobs]
            if not busy:
                if not queue and not any(waiting.values()):
                    break
                time.sleep(0.001)
                continue
            evs = [(cr, cr.run_chunk(self.chunk)) for cr in busy]
            for cr, ev in evs:
                cr.runner.stream.synchronize()
                try:
                    regrown = cr.runner.recover()   # list overflow: those steps were skipped on the device
                except _lib.RowLimitError as e:
                    self._stop_collapsed(cr, e)
                    regrown = True
                if regrown:
                    self.recoveries += 1
                    continue
                lat.extend(ev[i].elapsed_time(ev[i + 1]) for i in range(len(ev) - 1))
                with torch.cuda.stream(cr.runner.stream):
                    retired = cr.arena.harvest()
                for j in retired:
                    j.done_at = time.perf_counter() - t0
                    retired_ps += j.n * j.steps_done
                    self.resident_bytes -= self.job_bytes.pop(j.jid)
                    (failed if "error" in j.result else finished).append(j)
            if on_chunk is not None:
                on_chunk(self, now)
            if log and now - last_log > 5.0:
                last_log = now
                log(now, {n: len(c.arena.jobs) for n, c in self.classes.items()}, len(finished))
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        resident_ps = sum(cr.arena.particle_steps() for cr in self.classes.values())
        resident = []
        for cr in self.classes.values():
            if cr.arena.jobs:
                sp, th = cr.arena.records()
                for slot, j in cr.arena.jobs.items():
                    j.steps_done = int(th[slot]["steps"])
                    j.result = {"resident": True, **cr.arena.slot_state(sp[slot])}
                    resident.append(j)
        self.last_resident = resident
        for cr in self.classes.values():
            cr.arena.check()
        self.last_finished = finished
        self.last_failed = failed
        self.last_rejected = rejected
        st = _stats("served", finished, retired_ps + resident_ps, wall, lat)
        st["jobs_rejected"] = len(rejected)
        st["jobs_failed"] = len(failed)
        st["list_regrowths"] = self.recoveries
        st["mem_budget_bytes"] = self.mem_budget
        st["arena_bytes"] = self.arena_bytes
        st["peak_resident_job_bytes"] = self.peak_resident_bytes
        return st


def _stats(mode, finished, particle_steps, wall, lat):
    lat = np.asarray(lat) if len(lat) else np.zeros(1)
    return {"mode": mode, "particle_steps": int(particle_steps), "wall_s": wall,
            "particle_steps_per_s": particle_steps / wall, "jobs_finished": len(finished),
            "step_ms_p50": float(np.percentile(lat, 50)), "step_ms_p99": float(np.percentile(lat, 99)),
            "steps_timed": int(lat.size)}


def run_plain(jobs, classes, time_limit=None, chunk=20, device="cuda", log=None):
    """The same jobs one at a time, each with its own single-system arena and captured step."""
    dev = torch.device(device)
    t0 = time.perf_counter()
    last_log = 0.0
    done_ps = 0
    finished = []
    lat = []
    regrowths = 0
    for j in sorted(jobs, key=lambda j: j.arrival):
        now = time.perf_counter() - t0
        if time_limit is not None and now >= time_limit:
            break
        if j.arrival > now:
            time.sleep(j.arrival - now)
        shape, hyp, net = classes[j.cls]
        ar = Arena([_spec_of(j.spec)], shape, hyp, net, device=dev)
        runner = StepRunner(ar, use_graph=True)      # multi-step graphs for small systems
        runner.init()
        done_dev = 0
        left = j.steps
        while left > 0:
            if time_limit is not None and time.perf_counter() - t0 >= time_limit:
                break
            k = min(chunk, left)
            st = runner.stream
            ea, eb = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            ea.record(st)
            runner.run(k)
            eb.record(st)
            st.synchronize()
            if runner.recover():          # list overflow: the device skipped steps
                regrowths += 1
            now = ar.device_steps()       # frozen or skipped steps are not counted
            kd = now - done_dev
            done_dev = now
            if kd > 0:
                lat.extend([ea.elapsed_time(eb) / kd] * kd)
            left -= kd
            done_ps += j.n * kd
            if log and time.perf_counter() - t0 - last_log > 5.0:
                last_log = time.perf_counter() - t0
                log(last_log, {"job": j.jid, "n": j.n, "left": left}, len(finished))
        j.steps_done = j.steps - left
        if left == 0:
            j.done_at = time.perf_counter() - t0
            a = ar.system_arrays(0)
            j.result = dict(x=a["x"], v=a["v"], q=a["q"], p=a["p"],
                            thermo={k: float(ar.thermo()[0][k]) for k in ("xi", "eta", "xi_r", "eta_r", "Kt", "Kr",
                                                                          "U", "steps")})
            finished.append(j)
        ar.check()
    torch.cuda.synchronize()
    st = _stats("plain", finished, done_ps, time.perf_counter() - t0, lat)
    st["list_regrowths"] = regrowths
    return st


def _reached_spec(job):
    """The job's system at the state it reached (its result), else None."""
    r = job.result
    if not all(k in r for k in "xvqp"):
        return None
    s = _spec_of(job.spec)
    return SystemSpec(r["x"], r["v"], r["q"], r["p"], s.mass, s.L, s.kT, s.dt, s.moi, s.rc, skin=s.skin, tau=s.tau)


def plain_step_costs(jobs, classes, device="cuda", steps=20, warmup=5, log=None, reached=False):
    """Per job: ms per step of the plain single-system captured step on that job's system
    (same N, box, kT), timed with hipEvents after `warmup` steps on a copy of the system: its
    initial state, or with `reached` the state the job reached (its result; jobs without one are
    timed at the initial state).  Pair counts, and so step costs, change as a system evolves."""
    import copy
    dev = torch.device(device)
    costs = {}
    t0 = time.perf_counter()
    for k, j in enumerate(jobs):
        shape, hyp, net = classes[j.cls]
        spec = (_reached_spec(j) if reached else None) or copy.deepcopy(_spec_of(j.spec))
        ar = Arena([spec], shape, hyp, net, device=dev)
        r = StepRunner(ar, use_graph=True)
        r.init()
        r.run_exact(warmup)
        for _ in range(3):
            d0 = ar.device_steps()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(r.stream)
            r.run(steps)
            b.record(r.stream)
            r.stream.synchronize()
            if not r.recover():
                break
        ar.check()
        costs[j.jid] = a.elapsed_time(b) / max(ar.device_steps() - d0, 1)
        del r, ar
        if log and (k + 1) % 16 == 0:
            log(time.perf_counter() - t0, {"plain_costs": k + 1}, len(jobs))
    return costs


def same_work_plain(jobs, costs, sample_scale=1.0):
    """Time the plain one-system-at-a-time step needs for the work the jobs record in steps_done:
    sum_j steps_done_j * ms_per_step_j (over the jobs that have a cost, times sample_scale)."""
    ms = sum(j.steps_done * costs[j.jid] for j in jobs if j.jid in costs) * sample_scale
    work = sum(j.n * j.steps_done for j in jobs)
    timed = [j for j in jobs if j.jid in costs and j.steps_done > 0]
    out = {"particle_steps": int(work), "seconds": ms * 1e-3,
           "particle_steps_per_s": work / (ms * 1e-3) if ms > 0 else float("nan"),
           "jobs": len(jobs), "jobs_timed": len(timed)}
    if timed:   # per-step latency of the plain step, each job weighted by its steps
        lat = np.array([costs[j.jid] for j in timed])
        w = np.array([j.steps_done for j in timed], dtype=np.float64)
        order = np.argsort(lat)
        cw = np.cumsum(w[order]) / w.sum()
        out["step_ms_p50"] = float(lat[order][np.searchsorted(cw, 0.5)])
        out["step_ms_p99"] = float(lat[order][min(np.searchsorted(cw, 0.99), len(lat) - 1)])
    return out


def poisson_jobs(rate, duration, seed, shapes=("cube8", "tet10"), n_range=(128, 65536),
                 steps_range=(1000, 100000), kT_range=(0.5, 2.0), phi_range=(0.2, 0.45), max_jobs=None):
    """Seeded job trace: Poisson arrivals, N and step budget log-uniform, kT and packing uniform.
    Each job's particle arrays are built when it is first looked at (LazySpec)."""
    from .shapes import get_shape
    rng = np.random.default_rng(seed)
    t, jobs = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration or (max_jobs is not None and len(jobs) >= max_jobs):
            break
        shape = shapes[rng.integers(len(shapes))]
        n = int(round(math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1])))))
        steps = int(round(math.exp(rng.uniform(math.log(steps_range[0]), math.log(steps_range[1])))))
        kT = float(rng.uniform(*kT_range))
        phi = float(rng.uniform(*phi_range))
        sh = get_shape(shape)
        # the minimum image needs a box of at least 2 rc: small systems get a lower packing fraction
        phi = min(phi, n * sh.volume / (2.05 * sh.cutoff) ** 3)
        spec = LazySpec(shape, n, phi, kT=kT, seed=int(rng.integers(1 << 30)), non_overlap=False)
        jobs.append(Job(len(jobs), shape, spec, steps, t))
    return jobs
