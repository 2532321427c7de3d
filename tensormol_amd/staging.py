"""Pinned host staging for host<->device copies that must not block the host.

A copy from pageable numpy memory makes the host wait until the stream has
drained up to the copy; staging the data in pinned memory first lets the copy
queue behind running kernels while the host prepares the next batch.
"""
import numpy as np
import torch


class PinnedStager:
    """Host->device uploads through a ring of pinned slots.

    begin() opens the next slot (waiting, if needed, until the copies queued
    from it last time have completed), put() stages an array and queues its
    copy on the current stream, end() records the slot's completion event.
    """

    def __init__(self, device, slot_bytes=16 << 20, slots=2):
        self.device = device
        self.size = int(slot_bytes)
        self.bufs = [torch.empty(self.size, dtype=torch.uint8, pin_memory=True) for _ in range(slots)]
        self.events = [None] * slots
        self.cur = 0
        self.off = 0
        self.open = False

    def begin(self):
        self.cur = (self.cur + 1) % len(self.bufs)
        ev = self.events[self.cur]
        if ev is not None:
            ev.synchronize()
        self.off = 0
        self.open = True

    def end(self):
        ev = torch.cuda.Event()
        ev.record(torch.cuda.current_stream(self.device))
        self.events[self.cur] = ev
        self.open = False

    def put(self, arr, out=None):
        """Queue a copy of `arr` into `out` (or a new device tensor); returns it."""
        arr = np.ascontiguousarray(arr)
        src = torch.from_numpy(arr)
        dst = out if out is not None else torch.empty(arr.shape, dtype=src.dtype, device=self.device)
        n = arr.nbytes
        if n == 0:
            return dst
        off = (self.off + 255) & ~255
        if not self.open or off + n > self.size:
            dst.copy_(src)   # outside a slot, or too large: a plain copy
            return dst
        host = self.bufs[self.cur][off:off + n].view(src.dtype).view(arr.shape)
        host.copy_(src)
        dst.copy_(host, non_blocking=True)
        self.off = off + n
        return dst


class PinnedOutputs:
    """Named pinned host arrays that device results are copied into without
    blocking; read them after the event recorded by mark()."""

    def __init__(self, specs):
        self.arrs = {k: torch.empty(shape, dtype=dt, pin_memory=True) for k, (shape, dt) in specs.items()}
        self.event = None

    def fetch(self, name, src):
        """Queue a copy of the device tensor src into the first rows of `name`."""
        n = src.shape[0]
        if n:
            self.arrs[name][:n].copy_(src, non_blocking=True)

    def mark(self, device):
        self.event = torch.cuda.Event()
        self.event.record(torch.cuda.current_stream(device))

    def wait(self):
        if self.event is not None:
            self.event.synchronize()

    def host(self, name, n):
        return self.arrs[name][:n].numpy()
