"""Optional per-launch timing of selected kernels with HIP events.

Events are recorded on the stream the kernel is launched on (torch's current
stream, which is the stream every libtopohip call here receives), so the
measured durations are those of the kernels themselves.
"""

from __future__ import annotations

from collections import defaultdict

import torch


class KernelTimer:
    def __init__(self):
        self.records = defaultdict(list)   # name -> [(start, end, bytes)]

    def launch(self, name: str, nbytes: float, fn):
        s = torch.cuda.Event(enable_timing=True)
        e = torch.cuda.Event(enable_timing=True)
        s.record()
        fn()
        e.record()
        self.records[name].append((s, e, float(nbytes)))

    def summary(self) -> dict:
        torch.cuda.synchronize()
        out = {}
        for name, recs in self.records.items():
            ms = [s.elapsed_time(e) for s, e, _ in recs]
            tot_ms = sum(ms)
            tot_b = sum(b for _, _, b in recs)
            out[name] = {"launches": len(recs), "avg_us": 1e3 * tot_ms / max(len(recs), 1),
                         "bytes_per_launch": tot_b / max(len(recs), 1),
                         "gbps": tot_b / max(tot_ms * 1e-3, 1e-30) / 1e9}
        return out

    def reset(self):
        self.records.clear()
