"""Reads of a few device scalars by the host through pinned memory.

At a host decision (has the Krylov loop converged, has the OC bisection
stopped) the GPU idles from the moment the value is ready until the host has
queued the next launch. A blocking wait (torch's ``.item()`` / ``.cpu()``,
``Event.synchronize``) wakes the host 25-45 us after the GPU reached the
event on MI355X, so ``wait`` polls the event first and blocks only when the
value is far off. ``start`` queues the copies into pinned memory and may run
well before the value is needed (the host keeps queueing work meanwhile).
"""

from __future__ import annotations

import ctypes as C
import time

import numpy as np
import torch

from . import lib as L

# a wait longer than this is not on a latency-critical path: block
POLL_S = 2e-3


def _poll(event: torch.cuda.Event) -> None:
    deadline = time.perf_counter() + POLL_S
    while not event.query():
        if time.perf_counter() > deadline:
            event.synchronize()
            return


class Readback:
    def __init__(self, n: int, dtype=torch.float64):
        self.host = torch.zeros(int(n), dtype=dtype, pin_memory=True)
        self.event = torch.cuda.Event()
        self._k = 0

    def start(self, *srcs: torch.Tensor) -> "Readback":
        """Queue copies of the device tensors srcs (of the buffer's dtype, at
        most n elements together), in order, on torch's current stream."""
        k = 0
        for src in srcs:
            m = src.numel()
            if k + m > self.host.numel():
                raise ValueError(f"Readback of {k + m} values into a buffer of {self.host.numel()}")
            self.host[k:k + m].copy_(src.reshape(-1), non_blocking=True)
            k += m
        self.event.record()
        self._k = k
        return self

    def wait(self) -> np.ndarray:
        """The values of the last ``start``, once its copies have finished."""
        _poll(self.event)
        return self.host[:self._k].numpy().copy()

    def read(self, *srcs: torch.Tensor) -> np.ndarray:
        return self.start(*srcs).wait()


class Mapped:
    """n values that a kernel writes straight into pinned host memory
    through its device address (``device_ptr``, an output argument of the
    kernel's entry point). ``start`` records an event behind that kernel and
    ``wait`` returns the values once the event has passed: no copy launch
    sits between the kernel and the host's read."""

    def __init__(self, n: int, dtype=torch.float64):
        self.host = torch.zeros(int(n), dtype=dtype, pin_memory=True)
        self.event = torch.cuda.Event()
        dev = C.c_void_p()
        L.call("tg_mapped_pointer", self.host.data_ptr(), C.addressof(dev))
        self.device_ptr: int = dev.value

    def start(self) -> "Mapped":
        """Record the event on torch's current stream, behind the kernel
        that writes the values."""
        self.event.record()
        return self

    def wait(self) -> np.ndarray:
        _poll(self.event)
        return self.host.numpy().copy()
