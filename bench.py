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

PEAK_FP32_MFMA_TFLOPS = 157.3
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


def cpu_baseline(spec, shape_name, phi, hyp, net, target_s=15.0):
    """Reference algorithm (10-row FD stencil, FP64 numpy, one thread) on a bounded pair sample."""
    import numpy as np
    from oracle import nep_oracle as O, reference_step as R
    sh = O.shape_table()[shape_name]
    layers = net.layers_f64()
    # sample: pairs of the workload itself when small, else of a 600-body box at its packing
    # fraction (same per-pair work)
    from nepx.configs import make_system
    small = spec if spec.n <= 4096 else make_system(shape_name, 600, phi, seed=123)
    pairs = O.brute_pairs(small.x, small.L, sh["cutoff"])
    from threadpoolctl import threadpool_limits
    rng = np.random.default_rng(0)
    done = 0
    chunk = 1000
    order = rng.permutation(len(pairs))
    with threadpool_limits(limits=1):          # the BLAS calls (network layers) on one thread too
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < target_s and done < len(pairs):
            pc = pairs[order[done:done + chunk]]
            clouds, t = O.pair_clouds(small.x, small.q, pc, sh["sites"], small.L)
            O.fd_pair_forces(clouds, len(sh["sites"]), t, sh["nep_cutoff"], hyp, layers, "tanh", 0.0, 1.0,
                             1e-5, 1e-5)
            done += len(pc)
        t_pairs = time.perf_counter() - t0
        # integrator per particle (vectorised FP64 numpy)
        th = dict(xi=0.0, eta=0.0, xi_r=0.0, eta_r=0.0, theta=1.0, e_r=1.0)
        st = dict(x=spec.x, v=spec.v, q=spec.q, p=spec.p, m=spec.mass)
        t1 = time.perf_counter()
        O.nvt_step(st, np.zeros_like(spec.x), np.zeros_like(spec.x), spec.dt, spec.kT, spec.tau, spec.L,
                   spec.moi, th)
        per_particle = (time.perf_counter() - t1) / spec.n
    per_pair = t_pairs / done
    # pairs per particle of the full workload (half list) at this density
    p_per_particle = len(pairs) / small.n
    step_time = spec.n * (p_per_particle * per_pair + per_particle)
    return {"value": spec.n / step_time, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": f"{done} pairs of a {small.n}-body {shape_name} box at the same density through the reference "
                      f"algorithm (10-row forward-FD stencil, NEP descriptor, FFNN, clip; FP64 numpy, 1 thread) "
                      f"in {t_pairs:.1f} s, scaled by pairs per particle ({p_per_particle:.1f}); FP64 NVT "
                      f"update timed on the full state"}


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
    costs = plain_step_costs(timed, classes, dev, log=log)
    # uniform sample: scale the sampled plain time up to all jobs
    plain_same = same_work_plain(worked, costs, sample_scale=len(worked) / max(len(timed), 1))
    res = {
        "metric": "served-batch vs plain particle-steps/s (mixed cube/tet/cylinder job trace); p99 step latency",
        "value": served["particle_steps_per_s"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "served_over_plain": served["particle_steps_per_s"] / plain_same["particle_steps_per_s"],
        "served": served,
        "plain_same_work": plain_same,
        "config": {"workload": f"Poisson 20 jobs/s for {a.serve_arrivals} s of arrivals ({len(jobs)} jobs), "
                               f"{'/'.join(classes)}, N log-uniform [128, {a.serve_n_max}], steps log-uniform [1e3, 1e5], "
                               "kT U[0.5, 2], phi U[0.2, 0.45], FFNN D->64->64->1 (D=40) or D->128->128->1 "
                               "(D=60) tanh",
                   "window_s": a.serve_seconds, "capacity_per_class": a.serve_capacity, "chunk": a.serve_chunk,
                   "jobs_per_class": {c: sum(j.cls == c for j in jobs) for c in classes}},
        "dtype": f"f32 pair math ({'fp16 hi/lo split MFMA' if nets[40].precision == 'f16x3' else 'MFMA f32'}), "
                 "f64 integrator",
        "data": "synthetic",
    }
    if a.serve_plain_window:
        res["plain_window"] = run_plain(plain_jobs, classes, time_limit=a.serve_seconds, chunk=a.serve_chunk,
                                        device=dev, log=log)
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
        "dtype": "f32 (fp16 hi/lo split MFMA, FP32 accumulate)" if half else "f32 (MFMA f32)",
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
    r.run(a.warmup)
    r.stream.synchronize()
    ar.check()
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record(r.stream)
    r.run(a.steps)
    t1.record(r.stream)
    r.stream.synchronize()
    ar.check()
    ms = t0.elapsed_time(t1) / a.steps
    n_tot = sum(sp.n for sp in specs)
    jobs = [Job(k, "cube8", sp, a.steps) for k, sp in enumerate(specs)]
    for j in jobs:
        j.steps_done = a.steps
    log = lambda t, what, n: print(f"[replicas] t={t:.0f}s {what} of {n}", file=sys.stderr, flush=True)
    costs = plain_step_costs(jobs, {"cube8": (shape, hyp, net)}, dev, log=log)
    plain = same_work_plain(jobs, costs)
    return {
        "metric": "batched replicas vs plain particle-steps/s (256 cube systems)",
        "value": n_tot / (ms * 1e-3), "unit": "particle-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms, "higher_is_better": True,
        "batched_over_plain": (n_tot / (ms * 1e-3)) / plain["particle_steps_per_s"],
        "plain_same_work": plain,
        "plain_ms_per_step": {"min": min(costs.values()), "max": max(costs.values()),
                              "sum_over_systems": sum(costs.values())},
        "config": {"workload": "256 cube8 systems, N log-uniform [64, 4096], phi U[0.2, 0.5], kT U[0.5, 2], "
                               "rc 2.5, hypers (11, 6, 4), FFNN 40->64->64->1 tanh, hipGraph steps",
                   "particles": n_tot, "systems": len(specs)},
        "dtype": f"f32 pair math ({'fp16 hi/lo split MFMA' if net.precision == 'f16x3' else 'MFMA f32'}), "
                 "f64 integrator",
        "data": "synthetic", "tuning": tuning,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--workload", default="tet32k")
    ap.add_argument("--autotune", type=int, default=1, choices=[0, 1, 2],
                    help="pick skin and launch shapes by timing candidates before the timed region "
                         "(1: use the cache when it has this workload, 2: always re-measure, 0: defaults)")
    ap.add_argument("--precision", default="auto", choices=["auto", "f32", "f16x3"],
                    help="network MFMA operands: FP32, or fp16 hi/lo split (3 products, FP32 accumulate)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
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
    if a.workload == "replicas256":
        print(json.dumps(replicas_bench(a)))
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
    ar = nepx.Arena([spec], shape, hyp, net, device=dev)
    runner = StepRunner(ar, use_graph=bool(a.graph))
    runner.init()
    runner.run(a.warmup)
    torch.cuda.synchronize()
    ar.check()

    # timed region: K steps, events around every fused pair kernel on the step stream
    stream = runner.stream
    fused_ev = runner.enable_kernel_events(a.steps)
    t_start = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    t_start.record(stream)
    runner.run(a.steps)
    t_end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    flags = ar.check()
    ms = t_start.elapsed_time(t_end)
    ms_step = ms / a.steps
    fused_ms = runner.kernel_event_ms()
    P = int(flags[3])
    th = ar.thermo()[0]

    import ctypes
    lanes = int(nepx._lib.lib.nepx_pair_fused_supported(ctypes.byref(ar.model), ctypes.byref(ar.cnet)))
    kernel = {4: "k_fused_quad (four lanes per pair: descriptor + FFNN fwd/bwd + adjoint)",
              1: "k_fused (one lane per pair: descriptor + FFNN fwd/bwd + adjoint)"}.get(lanes, "split pair kernels")
    compute = ("FP32 MFMA v_mfma_f32_4x4x1_16b_f32 (per-pair site contractions) + MFMA (network) + FP32 VALU "
               "(site geometry, radial functions, harmonics, gradient terms); the bound is FP32 VALU issue"
               if lanes == 4 else "FP32 VALU (descriptor, adjoint) + MFMA (network); the bound is FP32 compute")
    flops = pair_flops(shape.n_sites, hyp, D, widths) * P
    achieved = flops / (fused_ms * 1e-3) / 1e12
    half = net.precision == "f16x3"
    peak = PEAK_FP32_MFMA_TFLOPS
    mfma_desc = "fp16 hi/lo split MFMA (3 products per FP32 product, FP32 accumulate)" if half else "MFMA f32"
    prof = profiled(a.workload)
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
        "dtype": f"f32 pair math ({mfma_desc}), f64 integrator",
        "data": "synthetic (seeded lattice + random orientations, random-init FFNN)",
        "config": {"workload": f"{w['n']} {w['shape']} phi={w['phi']} rc={shape.cutoff} hypers={hyp} D={D} "
                               f"FFNN {D}->{widths[0]}->{widths[1]}->1 tanh, NVT kT={w['kT']} dt={w['dt']}",
                   "particles": spec.n, "pairs_per_step": P, "hipgraph": bool(a.graph),
                   "list_builds": int(flags[4]), "wall_s": wall},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
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
        res["cpu_baseline"] = cpu_baseline(spec, w["shape"], w["phi"], hyp, net)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
