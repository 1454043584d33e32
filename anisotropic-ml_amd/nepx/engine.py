"""Step runner: a dedicated HIP stream, hipGraph capture, optional kernel timing events.

A production step is one captured graph replayed per step:
  conditional list rebuild + COM filter/compaction
  -> fused descriptor / MFMA network / adjoint kernel
  -> deterministic gather + FP64 NVT update + thermostat/rebuild test.
For timing the dominant kernel with events on the stream it runs on, the step
is recaptured as three graphs (pre / pair / post) while events are enabled;
the kernels and their arguments are the same.  No segment contains
hipMemsetAsync (all clears are fill kernels).  Graphs are captured through
torch.cuda.CUDAGraph, which drives hipStreamBeginCapture.
"""
import torch

from .system import Arena


class StepRunner:
    def __init__(self, arena: Arena, use_graph=True, stream=None):
        self.ar = arena
        self.use_graph = use_graph
        self.stream = stream or torch.cuda.Stream(device=arena.device)
        self.graphs = None
        self.split = False      # three graphs per step (pre / pair / post) while kernel events are on
        self._ev = None
        self._ev_i = 0

    def init(self, update_thermostat=True):
        with torch.cuda.stream(self.stream):
            self.ar.init(self.stream, update_thermostat)
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

    def run(self, n):
        for _ in range(n):
            self.step()

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
        """Back to one graph per step."""
        self._ev = None
        if self.use_graph and self.split:
            self.split = False
            self._capture()

    def kernel_event_ms(self):
        """Average duration of the timed segment over the recorded steps (synchronises)."""
        self.stream.synchronize()
        ts = [a.elapsed_time(b) for a, b in self._ev[: self._ev_i]]
        return sum(ts) / max(len(ts), 1)
