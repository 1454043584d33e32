"""Benchmark of the nepx MD step on one MI355X.

Workload (default): 32,768 rigid tetrahedra (edge 1, 4 vertex + 6 mid-edge
sites), packing fraction 0.45, rc = r_nep = 2.5, hypers (11, 6, 4) -> D = 40,
FFNN 40 -> 64 -> 64 -> 1 tanh with seeded N(0, 1/fan_in) weights, FP32 pair
math + FP64 integrator, NVT at kT = 1, dt = 0.001.  A "step" is one full MD
step: list rebuild when the skin is exceeded, COM filter, fused descriptor /
MFMA network / adjoint kernel, deterministic gather, FP64 NVT update.

Prints one JSON line (see the "value"/"roofline"/"cpu_baseline" keys).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "anisotropic-ml_amd")]

PEAK_FP32_MFMA_TFLOPS = 157.3      # MI355X FP32 (vector and matrix)
NET_DESC = {"f32": "MFMA f32",
            "f16x3": "fp16 hi/lo split MFMA (3 products per FP32 product, FP32 accumulate)",
            "bf16x3": "bf16 hi/lo split MFMA (3 products per FP32 product, FP32 accumulate)",
            "bf16": "bf16 MFMA (one product, FP32 accumulate)"}
PEAK_FP16_MFMA_TFLOPS = 2500.0
PEAK_HBM_GBS = 8000.0


def profiled(workload):
    """Per-launch figures of the fused pair kernel from the committed rocprofv3 summaries:
    HBM traffic (FETCH_SIZE and WRITE_SIZE runs scaled by a 1 GiB copy), VALU issue utilisation and
    MFMA busy fraction (SQ counters).  Missing summaries give None."""
    import re
    out = {"traffic": None, "valu_issue_util": None, "mfma_busy": None, "source": []}
    for name, keys in ((f"k_fused_{workload}_hbm_counters.md", ("traffic_bytes_per_launch",)),
                       (f"k_fused_{workload}_sq_counters.md", ("valu_issue_util", "mfma_busy"))):
        path = os.path.join(ROOT, "profiles", name)
        if not os.path.exists(path):
            continue
        text = open(path).read()
        for k in keys:
            m = re.search(rf"^{k}:\s*([0-9.eE+-]+)", text, re.M)
            if m:
                out["traffic" if k.startswith("traffic") else k] = float(m.group(1))
        out["source"].append(os.path.relpath(path, ROOT))
    return out


def ffnn_flops_per_pair(D, widths):
    h1, h2 = widths
    return 4 * (D * h1 + h1 * h2 + h2)


def pair_flops(n_sites, hyp, D, widths):
    """Algorithmic FP32 FLOPs per pair of the fused kernel, counting its contractions: the angular
    accumulation A[n][lm] += f_n Z_lm (2 (nang+1) NLM per site), the adjoint's c_lm and d_lm
    (4 (nang+1) NLM per site) and the network forward + backward-to-input 4 (D H1 + H1 H2 + H2).
    Site geometry, radial basis and harmonics recurrences are not counted."""
    nrad, nang, lmax = hyp
    nlm = lmax * (lmax + 2)
    return 2 * n_sites * 6 * (nang + 1) * nlm + ffnn_flops_per_pair(D, widths)


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def _cpu_workers():
    """Worker processes for the CPU baseline: the cores this process may run on, at most 16 (the host
    share of one GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


class _InlinePool:
    """Stand-in with the one method cpu_baseline uses, running in this process."""

    def map(self, fn, jobs, timeout=None):
        return [fn(j) for j in jobs]


def _start_pool(workers, probe_job, timeout=60.0):
    """A process pool of `workers` spawned workers that have answered one job within `timeout`
    seconds, or None (the caller then runs in this process): a pool that cannot start must not
    hold up the benchmark."""
    import concurrent.futures as cf
    import multiprocessing as mp
    from oracle import reference_step as R
    try:
        pool = cf.ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn"))
    except (OSError, ValueError):
        return None
    try:
        list(pool.map(R.fd_chunk, [probe_job] * workers, timeout=timeout))
        return pool
    except Exception as e:                  # timeout, a broken pool, a worker that failed to import
        print(f"[bench] CPU baseline worker pool unavailable ({type(e).__name__}); using one process",
              file=sys.stderr, flush=True)
        pool.shutdown(wait=False, cancel_futures=True)
        for p in list(getattr(pool, "_processes", {}).values()):
            p.terminate()
        return None


def cpu_baseline(spec, shape_name, hyp, net, target_s=15.0):
    """The reference algorithm (KD-tree list every 100 steps, 10-row FD stencil, NEP descriptor, FFNN,
    clip, scatter, NVT; FP64 numpy) on the host cores, pair chunks spread over worker processes with
    one BLAS thread each.  Systems of <= 4096 bodies run whole steps of the workload itself; larger
    ones time a sample of the workload's own pairs plus its list build and integrator."""
    import numpy as np
    from oracle import nep_oracle as O, reference_step as R
    sh = O.shape_table()[shape_name]
    layers = net.layers_f64()
    workers = _cpu_workers()
    info = {"cores": workers, "os_cpu_count": os.cpu_count(), "cpu_model": _cpu_model(),
            "blas_threads_per_worker": 1, "kind": "port", "unit": "particle-steps/s"}
    state = dict(x=spec.x.copy(), v=spec.v.copy(), q=spec.q.copy(), p=spec.p.copy(), m=spec.mass.copy())
    thermo = O.thermo_update(dict(xi=-0.0317711, eta=0.108609, xi_r=0.00814227, eta_r=-0.014238),
                             0.5 * float(np.sum(spec.mass[:, None] * spec.v ** 2)),
                             float(np.sum(O.rot_ke(spec.q, spec.p, spec.moi))), spec.ndof, spec.rndof, spec.dt,
                             spec.tau, spec.kT)
    act = getattr(net, "act", "tanh")
    probe = (spec.x[:2], spec.q[:2], np.array([[0, 1]], np.int64), sh, spec.L, hyp, layers, act, 0.0, 1.0,
             1e-5, 1e-5)
    pool = _start_pool(workers, probe) if workers > 1 else None
    if pool is None:
        workers = 1
        info["cores"] = 1
    try:
        if spec.n <= 4096:
            run = lambda st, th, n: R.reference_run(st, th, sh, spec.L, hyp, layers, spec.dt, spec.kT, spec.tau,
                                                    spec.moi, spec.ndof, spec.rndof, n, act=act, pool=pool,
                                                    chunk=max(64, 5000 // workers))
            t0 = time.perf_counter()
            state, thermo, _ = run(state, thermo, 1)                     # one step sizes the sample
            steps = int(min(1000, max(2, target_s / (time.perf_counter() - t0))))
            t0 = time.perf_counter()
            state, thermo, tm = run(state, thermo, steps)
            wall = time.perf_counter() - t0
            builds = (steps + R.LIST_EVERY - 1) // R.LIST_EVERY
            # one list build per 100 steps, as in the reference, whatever the sample length
            per_step = (tm["force_s"] + tm["integrate_s"]) / steps + tm["list_s"] / builds / R.LIST_EVERY
            return {**info, "value": spec.n / per_step,
                    "sample": f"{steps} steps of the workload itself through the reference algorithm (KD-tree "
                              f"list within cutoff + 2 every 100 steps, {tm['list_s'] / builds * 1e3:.1f} ms per "
                              f"build; COM filter; 10-row FD stencil, NEP descriptor, FFNN, clip, scatter on "
                              f"{workers} worker processes; FP64 NVT) in {wall:.1f} s"}
        t0 = time.perf_counter()
        nlist = R.kdtree_list(spec.x, spec.L, sh["cutoff"] + R.LIST_SKIN)
        t_list = time.perf_counter() - t0
        pairs = R.com_filter(spec.x, spec.L, nlist, sh["cutoff"])
        sample = pairs[np.random.default_rng(0).permutation(len(pairs))]
        chunk = 500
        done = 0
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < target_s and done < len(sample):
            batch = sample[done:done + chunk * workers]
            R.reference_forces(spec.x, spec.q, sh, spec.L, hyp, layers, act, pairs=batch, pool=pool, chunk=chunk)
            done += len(batch)
        t_pairs = time.perf_counter() - t0
    finally:
        if pool is not None:
            pool.shutdown(wait=True, cancel_futures=True)
    F0 = np.zeros_like(spec.x)
    t1 = time.perf_counter()
    O.nvt_step(state, F0, F0, spec.dt, spec.kT, spec.tau, spec.L, spec.moi, thermo)
    t_int = time.perf_counter() - t1
    step_time = len(pairs) * t_pairs / done + t_int + t_list / R.LIST_EVERY
    return {**info, "value": spec.n / step_time,
            "sample": f"{done} of the workload's {len(pairs)} pairs through the reference algorithm (10-row "
                      f"FD stencil, NEP descriptor, FFNN, clip, scatter; FP64 numpy on {workers} worker "
                      f"processes) in {t_pairs:.1f} s, scaled to all pairs; KD-tree list of the whole system "
                      f"({t_list:.2f} s per build, one per 100 steps) and the FP64 NVT update timed on the full "
                      f"state"}


def serve_bench(a):
    """Served (continuous batching) vs plain (one system at a time) on the same seeded job trace."""
    import copy
    import torch
    import nepx
    import numpy as np
    from nepx.serve import Server, plain_step_costs, poisson_jobs, run_plain, same_work_plain
    dev = torch.device("cuda:0")
    nets = {40: nepx.EnergyNet.random(40, [64, 64], seed=7, device=dev, precision=a.precision),
            60: nepx.EnergyNet.random(60, [128, 128], seed=7, device=dev, precision=a.precision)}
    classes = {}
    for name in a.serve_shapes.split(","):
        sh = nepx.get_shape(name)
        classes[name] = (sh, sh.hypers, nets[nepx.n_descriptor(*sh.hypers)])
    jobs = poisson_jobs(rate=20.0, duration=a.serve_arrivals, seed=2025, shapes=tuple(classes),
                        n_range=(128 if a.serve_n_max > 128 else 16, a.serve_n_max))
    plain_jobs = copy.deepcopy(jobs)
    def log(t, resident, done):
        print(f"[serve] t={t:.0f}s resident={resident} finished={done}", file=sys.stderr, flush=True)
    srv = Server(classes, capacity=a.serve_capacity, max_systems=1024, chunk=a.serve_chunk, device=dev)
    served = srv.run(jobs, time_limit=a.serve_seconds, log=log)
    # plain on the same work: every job the server worked on, timed alone with the single-system step
    worked = [j for j in srv.last_finished + srv.last_resident if j.steps_done > 0]
    rng = np.random.default_rng(0)
    timed = (worked if len(worked) <= a.serve_plain_jobs else
             [worked[i] for i in rng.choice(len(worked), a.serve_plain_jobs, replace=False)])
    # each sampled job's plain step is timed at its initial state and at the state it reached in
    # the served run (pair counts change as a system evolves); the comparison charges each job the
    # mean of the two costs (trapezoid over its run) and reports both ends
    c0 = plain_step_costs(timed, classes, dev, log=log)
    c1 = plain_step_costs(timed, classes, dev, log=log, reached=True)
    costs = {k: 0.5 * (c0[k] + c1[k]) for k in c0}
    # uniform sample: scale the sampled plain time up to all jobs
    scale = len(worked) / max(len(timed), 1)
    plain_same = same_work_plain(worked, costs, sample_scale=scale)
    plain_initial = same_work_plain(worked, c0, sample_scale=scale)
    plain_reached = same_work_plain(worked, c1, sample_scale=scale)
    res = {
        "metric": "served-batch vs plain particle-steps/s (mixed cube/tet/cylinder job trace); p99 step latency",
        "value": served["particle_steps_per_s"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "served_over_plain": served["particle_steps_per_s"] / plain_same["particle_steps_per_s"],
        "served": served,
        "plain_same_work": plain_same,
        "plain_same_work_initial_states": plain_initial,
        "plain_same_work_reached_states": plain_reached,
        "config": {"workload": f"Poisson 20 jobs/s for {a.serve_arrivals} s of arrivals ({len(jobs)} jobs), "
                               f"{'/'.join(classes)}, N log-uniform [128, {a.serve_n_max}], steps log-uniform [1e3, 1e5], "
                               "kT U[0.5, 2], phi U[0.2, 0.45], FFNN D->64->64->1 (D=40) or D->128->128->1 "
                               "(D=60) tanh",
                   "window_s": a.serve_seconds, "capacity_per_class": a.serve_capacity, "chunk": a.serve_chunk,
                   "jobs_per_class": {c: sum(j.cls == c for j in jobs) for c in classes}},
        "dtype": f"f32 pair math ({NET_DESC[nets[40].precision]}), f64 integrator",
        "data": "synthetic",
    }
    if a.serve_plain_window:
        res["plain_window"] = run_plain(plain_jobs, classes, time_limit=a.serve_seconds, chunk=a.serve_chunk,
                                        device=dev, log=log)
    return res


def reference_run_bench(a):
    """The reference's documented run (sim_input.py:10: --hypers 10 10 10 4.5 --shape th_v1, a
    TheEnergyNet(150, 3, 121) ReLU checkpoint) through the drop-in surface: NVT() with the setter
    sequence of sim_input.py:50-59 on the reference's init/th_init_N200.gsd (200 tetrahedra, box 33,
    copied to tests/golden), kT 0.5, dt 0.005, a 121->150x4->1 ReLU network with seeded weights (the
    checkpoint itself is not shipped) on the pair kernels without a fused instance (descriptor ->
    network -> adjoint through HBM).  Sim.run(steps) is timed end to end."""
    import contextlib
    import torch
    import nepx
    from nepx.NVT import NVT
    dev = torch.device("cuda:0")
    net = nepx.EnergyNet.random(121, [150, 150, 150, 150], seed=21, act="relu", en_min=-6.7, en_max=15.0,
                                device=dev)
    sim = NVT()
    with contextlib.redirect_stdout(sys.stderr):   # Sim prints its timers; stdout carries the JSON line
        sim.setDevice(0)
        sim.initializeSystem(os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden",
                                          "th_init_N200.gsd"))
        sim.setNepDescriptors([10, 10, 10, 4.5])
        sim.setEvaluator(net=net)
        sim.setShape("th_v1")
        sim.setNeighborList()
        sim.setkT(0.5)
        sim.setdt(0.005)
        sim.run(a.warmup)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sim.run(a.steps)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
    ar = sim.arena
    n = ar.specs[0].n
    res = {
        "metric": "particle-steps/s, the reference's documented run (200 th_v1 tetrahedra, D 121)",
        "value": n * a.steps / wall, "unit": "particle-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": wall * 1e3 / a.steps, "higher_is_better": True,
        "dtype": f"f32 pair math ({NET_DESC[net.precision]}), f64 integrator",
        "data": "the reference's init/th_init_N200.gsd; network 121->150x4->1 ReLU, seeded weights",
        "config": {"workload": "200 th_v1 tetrahedra (th_init_N200.gsd, box 33) hypers=(10, 10, 10) r_nep 4.5 "
                               "rc 5.75 FFNN 121->150->150->150->150->1 relu, NVT kT 0.5 dt 0.005",
                   "pairs_per_step": ar.n_pairs(), "fused_kernel": bool(not ar.split),
                   "steps_per_graph": sim.runner.K if sim.runner is not None else 1},
    }
    if not a.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline(ar.specs[0], "th_v1", (10, 10, 10), net)
    return res


def ffnn_bench(a):
    """The energy network alone (nepx_ffnn_fwd_bwd: forward and backward to the input) on the
    descriptor rows of one 32,768-tetrahedron step: 4,052,090 rows of D = 40 through
    40 -> 64 -> 64 -> 1 tanh, inputs and outputs resident in HBM."""
    import ctypes
    import torch
    import nepx
    from nepx import _lib
    dev = torch.device("cuda:0")
    P, D, widths = 4052090, 40, [64, 64]
    gen = torch.Generator().manual_seed(11)
    G = (torch.rand((P, D), generator=gen) * 4.0).to(dev)        # descriptor-like magnitudes
    net = nepx.EnergyNet.random(D, widths, seed=7, device=dev, precision=a.precision)
    E = torch.empty(P, dtype=torch.float32, device=dev)
    dG = torch.empty_like(G)
    n_dev = torch.tensor([P], dtype=torch.int32, device=dev)
    cn = net.cstruct()
    stream = torch.cuda.current_stream()
    h = _lib.stream_handle(stream)

    def run():
        _lib.check(_lib.lib.nepx_ffnn_fwd_bwd(ctypes.byref(cn), _lib.ptr(G), _lib.ptr(n_dev), P, _lib.ptr(E),
                                              _lib.ptr(dG), h), "nepx_ffnn_fwd_bwd")
    for _ in range(a.warmup):
        run()
    torch.cuda.synchronize()
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record(stream)
    for _ in range(a.steps):
        run()
    t1.record(stream)
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / a.steps
    flops = ffnn_flops_per_pair(D, widths) * P                    # FP32-equivalent, unpadded
    half = net.precision == "f16x3"
    # fp16 MFMAs the split network issues per 16-row tile (3 per 32-deep k-step): forward
    # 64-padded input and 64-wide layer, backward through both (dE/dG to 48 padded inputs)
    hw_f16 = (P / 16) * 90 * 16 * 16 * 32 * 2 if half else None
    return {
        "metric": "energy-network rows/s (forward + backward to the descriptor, 32k-tetrahedron step)",
        "value": P / (ms * 1e-3), "unit": "rows/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": f"f32 ({NET_DESC[net.precision]})",
        "data": "synthetic (uniform [0, 4) descriptors, random-init FFNN)",
        "config": {"workload": f"{P} rows, FFNN {D}->{widths[0]}->{widths[1]}->1 tanh, nepx_ffnn_fwd_bwd"},
        "roofline": {"bound": "mfma", "achieved": flops / (ms * 1e-3) / 1e12, "peak": PEAK_FP32_MFMA_TFLOPS,
                     "unit": "TFLOP/s", "frac": flops / (ms * 1e-3) / 1e12 / PEAK_FP32_MFMA_TFLOPS,
                     "traffic": None, "kernel": "k_ffnn", "flops_per_launch": flops,
                     "fp16_mfma_tflops": None if hw_f16 is None else hw_f16 / (ms * 1e-3) / 1e12,
                     "fp16_peak_tflops": PEAK_FP16_MFMA_TFLOPS,
                     "algorithmic_bytes_per_launch": P * (2 * 4 * D + 4)},
    }


def replicas_bench(a):
    """256 independent cube systems advanced by the same launches (one arena, hipGraph steps) against
    the same systems run one at a time with the single-system step."""
    import torch
    import nepx
    from nepx.configs import replica_specs
    from nepx.engine import StepRunner
    from nepx.serve import Job, plain_step_costs, same_work_plain
    dev = torch.device("cuda:0")
    shape = nepx.get_shape("cube8")
    hyp = shape.hypers
    net = nepx.EnergyNet.random(40, [64, 64], seed=7, device=dev, precision=a.precision)
    specs = replica_specs(256, seed=2026)
    tuning = None
    if a.autotune:
        from nepx.tune import Autotuner
        tr = Autotuner().tune(specs, shape, hyp, net, dev, steps=20, warmup=5, use_cache=a.autotune == 1)
        for sp in specs:
            sp.skin = tr.best.skin
        tuning = tr.summary()
    ar = nepx.Arena(specs, shape, hyp, net, device=dev)
    r = StepRunner(ar, use_graph=True)
    r.init()
    overflows = r.run_exact(a.warmup)
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s0 = ar.device_steps()
    t0.record(r.stream)
    overflows += r.run_exact(a.steps)      # a list overflow inside is grown and its skipped steps rerun
    t1.record(r.stream)
    r.stream.synchronize()
    ar.check()
    ms = t0.elapsed_time(t1) / max(ar.device_steps() - s0, 1)
    n_tot = sum(sp.n for sp in specs)
    jobs = [Job(k, "cube8", sp, a.steps) for k, sp in enumerate(specs)]
    for j in jobs:
        j.steps_done = a.steps
    log = lambda t, what, n: print(f"[replicas] t={t:.0f}s {what} of {n}", file=sys.stderr, flush=True)
    costs = plain_step_costs(jobs, {"cube8": (shape, hyp, net)}, dev, log=log)
    plain = same_work_plain(jobs, costs)
    return {
        "metric": "batched replicas vs plain particle-steps/s (256 cube systems)",
        "list_overflows_recovered": overflows,
        "value": n_tot / (ms * 1e-3), "unit": "particle-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms, "higher_is_better": True,
        "batched_over_plain": (n_tot / (ms * 1e-3)) / plain["particle_steps_per_s"],
        "plain_same_work": plain,
        "plain_ms_per_step": {"min": min(costs.values()), "max": max(costs.values()),
                              "sum_over_systems": sum(costs.values())},
        "config": {"workload": "256 cube8 systems, N log-uniform [64, 4096], phi U[0.2, 0.5], kT U[0.5, 2], "
                               "rc 2.5, hypers (11, 6, 4), FFNN 40->64->64->1 tanh, hipGraph steps",
                   "particles": n_tot, "systems": len(specs)},
        "dtype": f"f32 pair math ({NET_DESC[net.precision]}), f64 integrator",
        "data": "synthetic", "tuning": tuning,
    }


def long_run_bench(a):
    """A long NVT run of the headline system (32,768 tetrahedra, phi 0.45, kT 1, dt 0.001; --long-steps
    steps, default 10,000) with production steps, checked every 100 steps: a neighbour-list overflow
    is recovered (list grown, skipped steps rerun) and counted; a non-finite thermostat / energy
    record or particle state exits non-zero.  Reports the translational and rotational
    temperatures 2 K_t / N_dof and 2 K_r / RN_dof (means and spread over the second half, against kT)
    and the Nose-Hoover extended energy K_t + K_r + U + N_dof kT (tau^2 xi^2/2 + eta)
    + RN_dof kT (tau^2 xi_r^2/2 + eta_r) along the run."""
    import numpy as np
    import torch
    import nepx
    from nepx import _lib
    from nepx.configs import WORKLOADS, make_system
    from nepx.engine import StepRunner
    w = WORKLOADS["tet32k"]
    shape = nepx.get_shape(w["shape"])
    hyp = shape.hypers
    spec = make_system(w["shape"], w["n"], w["phi"], kT=w["kT"], dt=w["dt"], seed=2024)
    dev = torch.device("cuda:0")
    net = nepx.EnergyNet.random(40, [64, 64], seed=7, device=dev, precision=a.precision)
    ar = nepx.Arena([spec], shape, hyp, net, device=dev)
    runner = StepRunner(ar, use_graph=True)
    runner.init()
    chunk, n_chunks = 100, max(1, a.long_steps // 100)
    rows = []
    t0 = time.perf_counter()
    overflows = 0
    stopped = None
    for c in range(n_chunks):
        try:
            overflows += run_recovering(runner, chunk)
        except _lib.RowLimitError as e:
            # the random-init network is attractive at short range, so the system condenses; the list
            # build reports a particle with more neighbours than it sorts and the run ends there
            stopped = {"after_step": int(ar.device_steps()), "reason": str(e)}
            print(f"[long] stopped: {e}", file=sys.stderr, flush=True)
            break
        th = ar.thermo()[0]
        rec = {k: float(th[k]) for k in ("xi", "eta", "xi_r", "eta_r", "Kt", "Kr", "U", "steps")}
        if not all(np.isfinite(v) for v in rec.values()):
            print(json.dumps({"error": f"non-finite state by step {(c + 1) * chunk}", "last": rec}))
            sys.exit(4)
        rows.append(rec)
        if (c + 1) % 10 == 0:
            print(f"[long] step {(c + 1) * chunk}: T_t {2 * rec['Kt'] / spec.ndof:.4f} "
                  f"T_r {2 * rec['Kr'] / spec.rndof:.4f} U {rec['U']:.1f}", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    if len(rows) < 2:
        print(json.dumps({"error": "the run stopped before two samples", "stopped": stopped}))
        sys.exit(4)
    x = ar.system_arrays(0)
    if not all(np.all(np.isfinite(x[k])) for k in ("x", "v", "q", "p")):
        print(json.dumps({"error": "non-finite particle state at the end"}))
        sys.exit(4)
    kT, tau = spec.kT, spec.tau
    Tt = np.array([2 * r["Kt"] / spec.ndof for r in rows])
    Tr = np.array([2 * r["Kr"] / spec.rndof for r in rows])
    H = np.array([r["Kt"] + r["Kr"] + r["U"] + spec.ndof * kT * (0.5 * tau * tau * r["xi"] ** 2 + r["eta"])
                  + spec.rndof * kT * (0.5 * tau * tau * r["xi_r"] ** 2 + r["eta_r"]) for r in rows])
    half = len(rows) // 2
    steps = np.array([r["steps"] for r in rows])
    slope = float(np.polyfit(steps, H, 1)[0]) if len(rows) > 1 else 0.0
    K0 = rows[0]["Kt"] + rows[0]["Kr"]
    return {
        "metric": "long NVT run: temperatures and extended-energy drift (32k tetrahedra)",
        "value": float(Tt[half:].mean() / kT), "unit": "<T_t>/kT over the second half",
        "n_gpus": 1, "steps": int(steps[-1]), "higher_is_better": None,
        "T_trans": {"mean_2nd_half": float(Tt[half:].mean()), "std_2nd_half": float(Tt[half:].std()),
                    "first": float(Tt[0]), "last": float(Tt[-1])},
        "T_rot": {"mean_2nd_half": float(Tr[half:].mean()), "std_2nd_half": float(Tr[half:].std()),
                  "first": float(Tr[0]), "last": float(Tr[-1])},
        "kT": kT,
        "extended_energy": {"first": float(H[0]), "last": float(H[-1]), "min": float(H.min()),
                            "max": float(H.max()), "slope_per_1000_steps": 1000.0 * slope,
                            "drift_over_kinetic": float((H[-1] - H[0]) / K0),
                            "note": "K at the half-step velocities of the record, U of the step's pairs; pair "
                                    "energies do not vanish at the COM cutoff, so pairs crossing it change U "
                                    "discontinuously"},
        "samples": len(rows), "sample_every": chunk, "wall_s": wall, "list_overflows_recovered": overflows,
        "stopped": stopped,
        "U": {"first": float(rows[0]["U"]), "last": float(rows[-1]["U"])},
        "config": {"workload": f"{w['n']} {w['shape']} phi={w['phi']} rc={shape.cutoff} hypers={hyp} D=40 "
                               f"FFNN 40->64->64->1 tanh ({net.precision}), NVT kT={kT} dt={w['dt']} tau={tau}",
                   "particles": spec.n, "list_builds": int(ar.flags()[_lib.FLAG_BUILDS])},
        "series": {"step": steps[:: max(1, len(rows) // 20)].tolist(),
                   "T_trans": Tt[:: max(1, len(rows) // 20)].tolist(),
                   "T_rot": Tr[:: max(1, len(rows) // 20)].tolist(),
                   "H": H[:: max(1, len(rows) // 20)].tolist(),
                   "U": [r["U"] for r in rows][:: max(1, len(rows) // 20)]},
        "data": "synthetic",
    }


def run_recovering(runner, n):
    """n steps that the device completes (StepRunner.run_exact): list overflows are grown and frozen
    or skipped steps are run again.  Returns the overflows handled."""
    return runner.run_exact(n)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--workload", default="tet32k")
    ap.add_argument("--autotune", type=int, default=1, choices=[0, 1, 2],
                    help="pick skin and launch shapes by timing candidates before the timed region "
                         "(1: use the cache when it has this workload, 2: always re-measure, 0: defaults)")
    ap.add_argument("--precision", default="auto", choices=["auto"] + sorted(NET_DESC),
                    help="network MFMA operands: " + "; ".join(f"{k}: {v}" for k, v in NET_DESC.items()))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--reverse-records", action="store_true",
                    help="the pair kernel also writes each pair's j side in reverse-row order for the gather "
                         "(less gather traffic, scattered writes in the pair kernel; off by default)")
    ap.add_argument("--long-steps", type=int, default=10000, help="nvt_long workload: steps of the run")
    ap.add_argument("--graph", type=int, default=1, help="capture the step in a hipGraph (1) or launch eagerly (0)")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="after the timed region, copy MIB MiB device-to-device once (a kernel of known traffic "
                         "for scaling profiler byte counters)")
    ap.add_argument("--serve-seconds", type=float, default=60.0, help="serve workload: measurement window per mode")
    ap.add_argument("--serve-arrivals", type=float, default=20.0, help="serve workload: seconds of Poisson arrivals")
    ap.add_argument("--serve-capacity", type=int, default=262144, help="serve workload: particles per class arena")
    ap.add_argument("--serve-chunk", type=int, default=20, help="serve workload: steps per scheduling chunk")
    ap.add_argument("--serve-shapes", default="cube8,tet10,cyl12", help="serve workload: job classes")
    ap.add_argument("--serve-n-max", type=int, default=65536, help="serve workload: largest job size")
    ap.add_argument("--serve-plain-jobs", type=int, default=96,
                    help="serve workload: jobs timed alone for the same-work plain comparison (uniform sample)")
    ap.add_argument("--serve-plain-window", type=int, default=0,
                    help="serve workload: also run the jobs one at a time for the same wall window")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "nepx runs on one GPU per process; multi-GPU runs are not part of this build "
                                   "(independent systems shard by placing whole systems per GPU)"}))
        sys.exit(2)

    if a.workload == "serve":
        print(json.dumps(serve_bench(a)))
        return
    if a.workload == "ffnn":
        print(json.dumps(ffnn_bench(a)))
        return
    if a.workload == "ref_th200":
        print(json.dumps(reference_run_bench(a)))
        return
    if a.workload == "replicas256":
        print(json.dumps(replicas_bench(a)))
        return
    if a.workload == "nvt_long":
        print(json.dumps(long_run_bench(a)))
        return

    import numpy as np
    import torch

    import nepx
    from nepx.configs import WORKLOADS, make_system
    from nepx.engine import StepRunner

    w = WORKLOADS[a.workload]
    shape = nepx.get_shape(w["shape"])
    hyp = shape.hypers
    D = nepx.n_descriptor(*hyp)
    widths = [64, 64] if D == 40 else [128, 128]
    spec = make_system(w["shape"], w["n"], w["phi"], kT=w["kT"], dt=w["dt"], seed=2024)
    dev = torch.device("cuda:0")
    net = nepx.EnergyNet.random(D, widths, seed=7, device=dev, precision=a.precision)
    tuning = None
    if a.autotune:
        from nepx.tune import Autotuner
        tr = Autotuner().tune([spec], shape, hyp, net, dev, steps=40, warmup=10, use_cache=a.autotune == 1)
        spec.skin = tr.best.skin
        tuning = tr.summary()
        print(f"[bench] autotune {tr.seconds:.1f} s: {tuning['best']}", file=sys.stderr, flush=True)
    ar = nepx.Arena([spec], shape, hyp, net, device=dev, reverse_records=a.reverse_records)
    runner = StepRunner(ar, use_graph=bool(a.graph))
    runner.init()
    overflows = run_recovering(runner, a.warmup)

    # timed region: K production steps (one captured graph each), an event between steps for the
    # per-step latency percentiles; a window in which the list overflowed (its steps were skipped
    # on the device) is discarded and run again after the list has grown
    stream = runner.stream
    from nepx import _lib
    # (with multi-step graphs an event per graph: each of its steps is charged the graph's mean)
    K = runner.K
    launches = [K] * (a.steps // K) + [1] * (a.steps % K)
    for attempt in range(3):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(len(launches) + 1)]
        builds0 = int(ar.flags()[_lib.FLAG_BUILDS])
        done0 = ar.device_steps()
        torch.cuda.synchronize()
        wall0 = time.perf_counter()
        ev[0].record(stream)
        for k, m in enumerate(launches):
            runner.run_graph_k() if m > 1 else runner.step()
            ev[k + 1].record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - wall0
        if not runner.recover():
            break
        overflows += 1
    flags = ar.check()
    done = ar.device_steps() - done0          # frozen steps (stale list inside a graph) are not done
    step_ms = np.concatenate([np.full(m, ev[k].elapsed_time(ev[k + 1]) / m) for k, m in enumerate(launches)])
    ms = ev[0].elapsed_time(ev[-1])
    ms_window = ms / max(done, 1)
    builds_window = int(flags[_lib.FLAG_BUILDS]) - builds0
    P = int(flags[3])
    th = ar.thermo()[0]

    # the pair kernel alone: the same steps with events around it (three graphs per step)
    runner.enable_kernel_events(a.steps)
    runner.run(a.steps)
    fused_ms = runner.kernel_event_ms()
    runner.disable_kernel_events()
    if runner.recover():
        overflows += 1

    # list rebuilds: cost of a rebuild over the no-op test every step runs, and how often the skin
    # runs out (untimed probe of production steps); a window without a rebuild is charged its
    # expected share, one with rebuilds its own
    def timed_build(force, n=3):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(n):
            ar.build_list(force, stream)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / n
    rebuild_ms = timed_build(True) - timed_build(False)
    probe = max(200, 4 * a.steps)
    b0 = int(ar.flags()[_lib.FLAG_BUILDS])
    overflows += run_recovering(runner, probe)
    rate = (int(ar.check()[_lib.FLAG_BUILDS]) - b0) / probe
    ms_step = ms_window + (rate * done - builds_window) * rebuild_ms / max(done, 1)

    import ctypes
    lanes = int(nepx._lib.lib.nepx_pair_fused_supported(ctypes.byref(ar.model), ctypes.byref(ar.cnet)))
    kernel = {4: "k_fused_quad (four lanes per pair: descriptor + FFNN fwd/bwd + adjoint)",
              1: "k_fused (one lane per pair: descriptor + FFNN fwd/bwd + adjoint)"}.get(lanes, "split pair kernels")
    compute = ("FP32 MFMA v_mfma_f32_4x4x1_16b_f32 (per-pair site contractions) + MFMA (network) + FP32 VALU "
               "(site geometry, radial functions, harmonics, gradient terms); the bound is FP32 VALU issue"
               if lanes == 4 else "FP32 VALU (descriptor, adjoint) + MFMA (network); the bound is FP32 compute")
    flops = pair_flops(shape.n_sites, hyp, D, widths) * P
    achieved = flops / (fused_ms * 1e-3) / 1e12
    peak = PEAK_FP32_MFMA_TFLOPS
    mfma_desc = NET_DESC[net.precision]
    prof = profiled(a.workload) if net.precision == "f16x3" else profiled(None)
    res = {
        "metric": ("particle-steps/s at 32k tetrahedra" if a.workload == "tet32k"
                   else f"particle-steps/s ({a.workload})"),
        "value": spec.n / (ms_step * 1e-3),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "ms_per_step_window": ms_window,
        "steps_per_graph": K, "steps_done_in_window": done,
        "step_latency_ms": {"p50": float(np.percentile(step_ms, 50)), "p99": float(np.percentile(step_ms, 99)),
                            "max": float(step_ms.max())},
        "list_rebuild": {"ms": rebuild_ms, "per_step": rate, "in_window": builds_window,
                         "overflows_recovered": overflows,
                         "note": "ms_per_step and value charge the window rebuild_ms x per_step x steps in place "
                                 "of the rebuilds that fell inside it"},
        "dtype": f"f32 pair math ({mfma_desc}), f64 integrator",
        "data": "synthetic (seeded lattice + random orientations, random-init FFNN)",
        "config": {"workload": f"{w['n']} {w['shape']} phi={w['phi']} rc={shape.cutoff} hypers={hyp} D={D} "
                               f"FFNN {D}->{widths[0]}->{widths[1]}->1 tanh ({net.precision}), NVT kT={w['kT']} "
                               f"dt={w['dt']}",
                   "particles": spec.n, "pairs_per_step": P, "hipgraph": bool(a.graph),
                   "list_builds": int(flags[4]), "wall_s": wall},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": prof["traffic"],
                     "kernel": kernel, "kernel_ms": fused_ms, "compute": compute,
                     "flops_per_launch": flops, "flops_per_pair": pair_flops(shape.n_sites, hyp, D, widths),
                     "ffnn_flops_per_launch": ffnn_flops_per_pair(D, widths) * P,
                     "algorithmic_bytes_per_launch": P * 72 + spec.n * 16,
                     "valu_issue_util": prof["valu_issue_util"], "mfma_busy": prof["mfma_busy"],
                     "profile_source": prof["source"]},
        "thermo": {"Kt_per_dof": float(th["Kt"] / spec.ndof * 2), "U": float(th["U"])},
        "tuning": tuning,
    }
    if a.copy_calib:
        n = a.copy_calib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        res["copy_calib_bytes"] = {"read": 4 * n, "write": 4 * n}
    if not a.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline(spec, w["shape"], hyp, net)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
