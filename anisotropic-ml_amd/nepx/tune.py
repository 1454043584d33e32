"""Autotuner for the step's launch shapes and the Verlet-list skin.

What is tuned (none of it changes which pairs interact; only pair_lanes changes per-pair values,
at FP32 rounding level):
  skin                 Verlet skin: list length (filter cost every step) against rebuild frequency
  gather_lanes         lanes per particle in the deterministic force gather (4 .. 64)
  fused_blocks_per_cu  grid of the fused descriptor/network/adjoint kernel (persistent blocks per CU)
  pair_blocks_per_cu   grid cap of the split descriptor and adjoint kernels
  net_blocks_per_cu    grid cap of the split network kernel
  pair_lanes           lanes per pair in the fused kernel: 4 puts the per-pair site sums on the
                       matrix cores (default where the shape allows), 1 keeps one lane per pair;
                       the two differ in FP32 rounding only, so it is not searched by default

Each candidate is measured on the workload itself, on a scratch copy of the
systems: the launch parameters are set process-wide (nepx_set_tuning), the
step is captured as hipGraphs, and hipEvents time `steps` steps after
`warmup` untimed ones.  Knobs are searched one at a time (coordinate descent
from the defaults), and the result is cached in a JSON file keyed by device,
pair model, network and the size/density bucket of the workload.  Skin and
gather lanes change only the summation order of FP64 force sums (gather) or
nothing at all (skin: the list is complete for any skin), so tuned and
untuned runs agree to rounding.

There is no counterpart in the reference: its launch shapes are fixed
(nep_simulator/NVT.py:46-111, DescriptorGenerator.py:349-397) and its list
margin is a constant cutoff + 2.0 (Sim.py:85-92).
"""
import copy
import ctypes as C
import json
import math
import os
import time
from dataclasses import asdict, dataclass, field

import torch

from . import _lib

KNOBS = ("fused_blocks_per_cu", "pair_blocks_per_cu", "gather_lanes", "net_blocks_per_cu", "pair_lanes")


@dataclass
class Tuning:
    skin: float = 0.5
    fused_blocks_per_cu: int = 1
    pair_blocks_per_cu: int = 16
    gather_lanes: int = 16
    net_blocks_per_cu: int = 2
    pair_lanes: int = 4

    def launch(self):
        t = _lib.Tuning()
        for k in KNOBS:
            setattr(t, k, int(getattr(self, k)))
        return t


DEFAULT_SPACE = {
    "skin": (0.3, 0.4, 0.5, 0.7),
    "gather_lanes": (4, 8, 16, 32),
    "fused_blocks_per_cu": (1, 2, 4),
    "pair_blocks_per_cu": (4, 16, 64),
    "net_blocks_per_cu": (1, 2, 4),
}


def set_launch(t: Tuning):
    _lib.check(_lib.lib.nepx_set_tuning(C.byref(t.launch())), "nepx_set_tuning")


def get_launch():
    t = _lib.Tuning()
    _lib.check(_lib.lib.nepx_get_tuning(C.byref(t)), "nepx_get_tuning")
    return {k: int(getattr(t, k)) for k in KNOBS}


def workload_key(specs, shape, hyp, net, device):
    n = sum(s.n for s in specs)
    rho = sum(s.n for s in specs) / sum(s.L ** 3 for s in specs)
    phi = rho * shape.volume
    name = torch.cuda.get_device_name(device) if torch.cuda.is_available() else "cpu"
    return (f"{name}|{shape.name}|hyp={tuple(hyp)}|net={net.d_in}x{net.width}x{net.n_hidden + 1}:{net.act}:"
            f"{net.precision}|N~2^{round(math.log2(max(n, 1)))}|phi~{round(phi * 20) / 20:.2f}|"
            f"systems~2^{round(math.log2(len(specs)))}")


@dataclass
class Trial:
    tuning: dict
    ms_per_step: float


@dataclass
class Result:
    key: str
    best: Tuning
    trials: list = field(default_factory=list)
    seconds: float = 0.0
    cached: bool = False

    def summary(self):
        return {"key": self.key, "best": asdict(self.best), "cached": self.cached, "seconds": self.seconds,
                "trials": [{"tuning": t.tuning, "ms_per_step": t.ms_per_step} for t in self.trials]}


def coordinate_descent(measure, space, start: Tuning, order=None):
    """Minimise measure(Tuning) one knob at a time, in `order`, keeping each knob's best value.

    measure returns ms/step (or inf when a candidate cannot run).  Returns (best, trials).
    """
    order = order or list(space)
    best = copy.deepcopy(start)
    trials = []
    cache = {}

    def run(t):
        k = json.dumps(asdict(t), sort_keys=True)
        if k not in cache:
            cache[k] = float(measure(t))
            trials.append(Trial(asdict(t), cache[k]))
        return cache[k]

    best_ms = run(best)
    for knob in order:
        for v in space[knob]:
            cand = copy.deepcopy(best)
            setattr(cand, knob, v)
            ms = run(cand)
            if ms < best_ms:
                best, best_ms = cand, ms
    return best, trials


class Autotuner:
    def __init__(self, cache_path=None):
        self.cache_path = cache_path or os.environ.get(
            "NEPX_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "nepx", "tune.json"))

    # cache ------------------------------------------------------------------
    def _load(self):
        try:
            with open(self.cache_path) as f:
                return json.load(f)
        except (OSError, ValueError):
            return {}

    def _store(self, key, best: Tuning):
        db = self._load()
        db[key] = asdict(best)
        os.makedirs(os.path.dirname(self.cache_path) or ".", exist_ok=True)
        tmp = self.cache_path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(db, f, indent=1, sort_keys=True)
        os.replace(tmp, self.cache_path)

    # measurement ------------------------------------------------------------
    @staticmethod
    def _measure(specs, shape, hyp, net, device, t: Tuning, steps, warmup):
        from .engine import StepRunner
        from .system import Arena
        set_launch(t)
        work = [copy.deepcopy(s) for s in specs]
        for s in work:
            s.skin = float(t.skin)
        ar = Arena(work, shape, hyp, net, device=device)
        r = StepRunner(ar, use_graph=True)
        r.init()
        r.run(warmup)
        r.stream.synchronize()
        r.recover()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(r.stream)
        r.run(steps)
        b.record(r.stream)
        r.stream.synchronize()
        if r.recover():      # a list regrowth inside the timed steps: time them again
            a.record(r.stream)
            r.run(steps)
            b.record(r.stream)
            r.stream.synchronize()
        ar.check()
        ms = a.elapsed_time(b) / steps
        del r, ar
        return ms

    def tune(self, specs, shape, hyp, net, device="cuda", steps=60, warmup=10, space=None, use_cache=True):
        """Pick the launch shapes and skin for these systems; leaves the winner set process-wide.

        The caller applies best.skin to its own systems before building its arena.
        """
        device = torch.device(device)
        key = workload_key(specs, shape, hyp, net, device)
        if use_cache:
            hit = self._load().get(key)
            if hit:
                best = Tuning(**hit)
                set_launch(best)
                return Result(key, best, cached=True)
        space = dict(DEFAULT_SPACE if space is None else space)
        from .model import make_pair_model
        fused = bool(_lib.lib.nepx_pair_fused_supported(C.byref(make_pair_model(shape, hyp)),
                                                         C.byref(net.cstruct())))
        order = ["skin", "gather_lanes"] + (["fused_blocks_per_cu"] if fused else
                                            ["pair_blocks_per_cu", "net_blocks_per_cu"])
        t0 = time.perf_counter()
        best, trials = coordinate_descent(
            lambda t: self._measure(specs, shape, hyp, net, device, t, steps, warmup), space, Tuning(), order)
        set_launch(best)
        self._store(key, best)
        return Result(key, best, trials, time.perf_counter() - t0)
