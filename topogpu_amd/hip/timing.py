"""Optional accounting of selected kernel launches.

``KernelTimer(events=True)`` brackets every launch with HIP events recorded
on the stream the kernel is launched on (torch's current stream, which every
libtopohip call here receives), so the measured durations are those of the
kernels themselves. Events cost host time per launch, so a timed benchmark
region uses ``events=False``: launches and algorithmic bytes are counted per
(kernel, epilogue mode) and nothing else is done. The last launch of each
key is kept as a closure (``replay``; called with no argument it launches on
its original stream, called with a stream handle on that stream), so that its
duration can be measured afterwards by launching it again back to back.
A replayed launch graph is counted as one increment (``graphs``); its
launches join ``counts`` and ``bytes`` when ``flush`` is called, so the
accounting costs the timed loop one dictionary update per replay.
"""

from __future__ import annotations

from collections import defaultdict

import torch


class KernelTimer:
    def __init__(self, events: bool = True):
        self.events = events
        self.records = defaultdict(list)          # name -> [(start, end, bytes)]
        self.counts = defaultdict(int)            # (name, mode) -> launches
        self.bytes = defaultdict(float)           # (name, mode) -> bytes
        self.replay = {}                          # (name, mode) -> (fn, bytes of one launch)
        self.graphs = {}                          # LaunchGraph -> replays not yet in counts

    def flush(self) -> None:
        """Move the pending graph replays into counts and bytes."""
        for g, reps in self.graphs.items():
            for k, n in g.counts.items():
                self.counts[k] += n * reps
                self.bytes[k] += g.nbytes.get(k, 0.0) * reps
        self.graphs.clear()

    def launch(self, name: str, nbytes: float, fn, mode: int = -1):
        self.counts[(name, mode)] += 1
        self.bytes[(name, mode)] += float(nbytes)
        self.replay[(name, mode)] = (fn, float(nbytes))
        if not self.events:
            fn()
            return
        s = torch.cuda.Event(enable_timing=True)
        e = torch.cuda.Event(enable_timing=True)
        s.record()
        fn()
        e.record()
        self.records[name].append((s, e, float(nbytes)))

    def summary(self) -> dict:
        """Per kernel name: launches, and with events the average duration
        and the achieved algorithmic bandwidth."""
        out = {}
        self.flush()
        for (name, _), n in self.counts.items():
            out.setdefault(name, {"launches": 0})["launches"] += n
        if self.events:
            torch.cuda.synchronize()
            for name, recs in self.records.items():
                ms = [s.elapsed_time(e) for s, e, _ in recs]
                tot_ms = sum(ms)
                tot_b = sum(b for _, _, b in recs)
                out[name].update({"avg_us": 1e3 * tot_ms / max(len(recs), 1),
                                  "bytes_per_launch": tot_b / max(len(recs), 1),
                                  "gbps": tot_b / max(tot_ms * 1e-3, 1e-30) / 1e9})
        return out

    def reset(self):
        self.records.clear()
        self.counts.clear()
        self.bytes.clear()
        self.replay.clear()
        self.graphs.clear()
