"""Captured launch sequences (tg_graph_* in libtopohip): one Krylov iteration
of the state solve -- the V-cycle and the PCG updates, about 50 launches --
captured once as two sequences and replayed with two calls per iteration.

The reference's PCG loop (``_cupy_pcg``, src/gpu_fem/solver_v2.py:552) issues
every kernel from the host on each iteration. The host cost of a launch
(argument packing plus the runtime's dispatch, several microseconds) is as
large as the coarse-level kernels themselves (about 3.5 us each at
160x80x80), so a host-driven V-cycle leaves the GPU idle between them after
every host read of the residual; a replayed graph does not (9.9 against
10.5 ms per SIMP iteration there). HipStateSolver replays by default
(graphs=True) and queues the next iteration's first part (the level-0
pre-smoothing and the restriction, which touch only the hierarchy's own
buffers) before it waits for r.r.

A graph keeps the arguments its launches had at capture (pointers and
scalars such as the Chebyshev coefficients), so every graph carries a key of
what it baked in and is captured again when the key changes (about 1.3 ms
per capture, after a lambda re-estimation).
"""

from __future__ import annotations

import ctypes as C
from collections import defaultdict

import torch

from . import lib as L


class LaunchGraph:
    """An instantiated capture: replay with launch(stream)."""

    def __init__(self, key, handle: int, counts: dict, nbytes: dict, replay: dict | None = None):
        self.key = key
        self._h = handle
        self.counts = dict(counts)      # (kernel, epilogue) -> accounted launches per replay
        self.nbytes = dict(nbytes)
        self.replay = dict(replay or {})   # (kernel, epilogue) -> (direct launch, bytes) for timing

    def launch(self, stream) -> None:
        L.call("tg_graph_launch", self._h, stream)

    def close(self) -> None:
        if self._h:
            h, self._h = self._h, None
            L.call("tg_graph_destroy", h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _Recorder:
    """Stand-in for a KernelTimer while capturing: counts the level-0
    launches of one replay and issues them (into the capture)."""

    events = False

    def __init__(self):
        self.counts = defaultdict(int)
        self.bytes = defaultdict(float)
        self.replay = {}

    def launch(self, name, nbytes, fn, mode=-1):
        self.counts[(name, mode)] += 1
        self.bytes[(name, mode)] += float(nbytes)
        self.replay[(name, mode)] = (fn, float(nbytes))
        fn()


def capture(key, issue, owner, device) -> LaunchGraph:
    """Capture the launches issue(stream) makes on a fresh side stream.

    owner.timer (a KernelTimer or None) is swapped for a recorder while
    capturing, so the graph knows the accounted launches of one replay."""
    side = torch.cuda.Stream(device=device)
    st = side.cuda_stream
    rec = _Recorder()
    saved = owner.timer
    owner.timer = rec
    L.call("tg_graph_begin", st)
    h = C.c_void_p()
    try:
        issue(st)
    except BaseException:
        # end the capture so the stream is usable again, then drop it
        L.load().tg_graph_end(st, C.byref(h))
        if h.value:
            L.load().tg_graph_destroy(h)
        raise
    finally:
        owner.timer = saved
    L.call("tg_graph_end", st, C.byref(h))
    return LaunchGraph(key, h.value, rec.counts, rec.bytes, rec.replay)


def account(timer, graph: LaunchGraph) -> None:
    """Add one replay's accounted launches to a KernelTimer's counters (as a
    pending replay when the timer keeps them, see KernelTimer.flush)."""
    if timer is None:
        return
    pending = getattr(timer, "graphs", None)
    if pending is not None:
        n = pending.get(graph)
        if n is None:
            # the graph's launches, with this replay's operands, are the
            # ones a later timing of each key relaunches
            timer.replay.update(graph.replay)
        pending[graph] = 1 if n is None else n + 1
        return
    for k, n in graph.counts.items():
        timer.counts[k] += n
        timer.bytes[k] += graph.nbytes.get(k, 0.0)
    if hasattr(timer, "replay"):
        timer.replay.update(graph.replay)
