"""Step runner: a dedicated HIP stream, hipGraph capture, optional kernel timing events.

A production step is one captured graph replayed per step:
  conditional list rebuild + COM filter/compaction
  -> fused descriptor / MFMA network / adjoint kernel
  -> deterministic gather + FP64 NVT update + thermostat/rebuild test.
Small single systems, whose step is bound by launch latency, capture K steps per
graph with the conditional rebuild only at the first: the other steps launch
four kernels each.  Rebuilds are requested early (once a particle has moved
EARLY_REBUILD of half the skin), so the list rarely goes stale inside a graph; if
it does, the filter freezes the remaining steps of that graph (nothing moves,
the device step count stops) and the next graph rebuilds and continues, so the
trajectory is that of single-step graphs.  Callers that need an exact step count
read it from the device (Arena.device_steps) at their synchronisation points.
For timing the dominant kernel with events on the stream it runs on, the step
is recaptured as three graphs (pre / pair / post) while events are enabled;
the kernels and their arguments are the same.  No segment contains
hipMemsetAsync (all clears are fill kernels).  Graphs are captured through
torch.cuda.CUDAGraph, which drives hipStreamBeginCapture.
"""
import torch

from .system import Arena

EARLY_REBUILD = 0.75          # fraction of half the skin at which multi-step graphs request a rebuild
SMALL_SYSTEM = 16384          # particles up to which a single system gets multi-step graphs
STEPS_PER_GRAPH = 16


def auto_steps_per_graph(arena: Arena):
    """Steps per captured graph: several for one small system (launch-latency bound), else one
    (a stale list freezes every system of the arena, so batches of systems keep one step)."""
    if len(arena.specs) == 1 and arena.specs[0].n <= SMALL_SYSTEM:
        return STEPS_PER_GRAPH
    return 1


class StepRunner:
    def __init__(self, arena: Arena, use_graph=True, stream=None, steps_per_graph="auto"):
        self.ar = arena
        self.use_graph = use_graph
        self.stream = stream or torch.cuda.Stream(device=arena.device)
        self.K = auto_steps_per_graph(arena) if steps_per_graph == "auto" else max(1, int(steps_per_graph))
        if not use_graph:
            self.K = 1
        self.graphs = None     # one-step graph(s)
        self.graph_k = None    # K-step graph (K > 1)
        self.split = False      # three graphs per step (pre / pair / post) while kernel events are on
        self._ev = None
        self._ev_i = 0

    def init(self, update_thermostat=True):
        with torch.cuda.stream(self.stream):
            self.ar.init(self.stream, update_thermostat)
            if self.K > 1:
                self.ar.set_early_rebuild(EARLY_REBUILD, self.stream)
        self.stream.synchronize()
        if self.use_graph:
            self._capture()

    # segments ---------------------------------------------------------------
    def _pre(self):
        self.ar.build_list(False, self.stream)
        self.ar.filter_pairs(self.stream)

    def _pair(self):
        self.ar.pair_fused(self.stream)

    def _post(self):
        self.ar.gather(self.stream)
        self.ar.nvt_step(self.stream)

    def _whole(self):
        self._pre()
        self._pair()
        self._post()

    def _k_steps(self):
        self._whole()
        for _ in range(self.K - 1):
            self.ar.filter_pairs(self.stream)
            self._pair()
            self._post()

    def _capture(self):
        # buffers and kernel choice are settled before capture; capture launches nothing
        self.ar.prepare_pair()
        self.stream.synchronize()
        self.graphs = []
        for seg in ((self._pre, self._pair, self._post) if self.split else (self._whole,)):
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=self.stream):
                seg()
            self.graphs.append(g)
        self.graph_k = None
        if self.K > 1 and not self.split:
            self.graph_k = torch.cuda.CUDAGraph()
            with torch.cuda.graph(self.graph_k, stream=self.stream):
                self._k_steps()
        self.stream.synchronize()

    def _eager_step(self):
        with torch.cuda.stream(self.stream):
            self._pre()
            self._timed(self._pair)
            self._post()
        self.ar.steps += 1

    def _timed(self, fn):
        if self._ev is not None and self._ev_i < len(self._ev):
            a, b = self._ev[self._ev_i]
            a.record(self.stream)
            fn()
            b.record(self.stream)
            self._ev_i += 1
        else:
            fn()

    def step(self):
        """One step (a one-step graph)."""
        if not self.graphs:
            self._eager_step()
            return
        with torch.cuda.stream(self.stream):
            if self.split:
                self.graphs[0].replay()
                self._timed(self.graphs[1].replay)
                self.graphs[2].replay()
            else:
                self.graphs[0].replay()
        self.ar.steps += 1

    def run_graph_k(self):
        """K steps in one graph launch (K > 1 only)."""
        with torch.cuda.stream(self.stream):
            self.graph_k.replay()
        self.ar.steps += self.K

    def run(self, n):
        """Launch n steps: K-step graphs, then single steps for the remainder.  A step frozen by a
        stale list inside a K-step graph is not done on the device (see Arena.device_steps)."""
        if self.graph_k is not None:
            for _ in range(n // self.K):
                self.run_graph_k()
            n %= self.K
        for _ in range(n):
            self.step()

    def run_exact(self, n):
        """n steps done on the device: launch, synchronise, grow the list after an overflow, and
        launch again until the device step count has advanced by n.  Returns the overflows handled."""
        target = self.ar.device_steps() + n
        handled = 0
        left = n
        while left > 0:
            self.run(left)
            self.stream.synchronize()
            if self.recover():
                handled += 1
            self.ar.check()
            left = target - self.ar.device_steps()
        return handled

    def recover(self):
        """Handle a list overflow seen at a synchronisation point: grow and rebuild the list and
        recapture the step.  Steps launched while the list was overflowed did nothing; callers
        count progress with Arena.device_steps().  Returns True when an overflow was handled."""
        if not self.ar.recover_overflow(self.stream):
            return False
        if self.use_graph:
            self._capture()
        return True

    # timing -----------------------------------------------------------------
    def enable_kernel_events(self, n):
        """Record events around the pair kernel of the next n steps (the step is recaptured as three
        graphs while they are on)."""
        if self.use_graph and not self.split:
            self.split = True
            self._capture()
        self._ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(n)]
        self._ev_i = 0
        return self._ev

    def disable_kernel_events(self):
        """Back to the production graphs."""
        self._ev = None
        if self.use_graph and self.split:
            self.split = False
            self._capture()

    def kernel_event_ms(self):
        """Average duration of the timed segment over the recorded steps (synchronises)."""
        self.stream.synchronize()
        ts = [a.elapsed_time(b) for a, b in self._ev[: self._ev_i]]
        return sum(ts) / max(len(ts), 1)
