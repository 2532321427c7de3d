"""Pinned host staging for host<->device copies that must not block the host.

A copy from pageable numpy memory makes the host wait until the stream has
drained up to the copy; staging the data in pinned memory first lets the copy
queue behind running kernels while the host prepares the next batch.
"""
import numpy as np
import torch


class PinnedStager:
    """Host->device uploads of one batch through a ring of pinned slots and a
    single device arena, in one copy.

    begin() opens the next slot (waiting, if needed, until the copy queued from
    it last time has completed).  put() stages an array in the slot and returns
    the device view where it will land (or, with `out`, a copy into `out` is
    queued after the upload).  flush() queues the one host->device copy of the
    slot, then those copies; device work that reads a put() result must be
    queued after flush().  end() flushes if needed and records the slot's
    completion event.  Everything runs on the current stream, so the arena is
    rewritten only after the kernels queued before the next flush() have read it.
    """

    def __init__(self, device, slot_bytes=16 << 20, slots=2):
        self.device = device
        self.size = int(slot_bytes)
        self.bufs = [torch.empty(self.size, dtype=torch.uint8, pin_memory=True) for _ in range(slots)]
        self.np_bufs = [b.numpy() for b in self.bufs]
        self.arena = torch.empty(self.size, dtype=torch.uint8, device=device)
        self.events = [None] * slots
        self.cur = 0
        self.off = 0
        self.open = False
        self.flushed = True
        self.pending = []

    def begin(self):
        self.cur = (self.cur + 1) % len(self.bufs)
        ev = self.events[self.cur]
        if ev is not None:
            ev.synchronize()
        self.off = 0
        self.open = True
        self.flushed = False
        self.pending = []

    def flush(self):
        if self.flushed:
            return
        if self.off:
            self.arena[:self.off].copy_(self.bufs[self.cur][:self.off], non_blocking=True)
        for dst, src in self.pending:
            dst.copy_(src)
        self.pending = []
        self.flushed = True

    def end(self):
        self.flush()
        ev = torch.cuda.Event()
        ev.record(torch.cuda.current_stream(self.device))
        self.events[self.cur] = ev
        self.open = False

    def put(self, arr, out=None):
        """Stage `arr`; returns its device copy (valid for work queued after
        flush()), or `out`, into which it is copied at flush()."""
        arr = np.ascontiguousarray(arr)
        n = arr.nbytes
        off = (self.off + 255) & ~255
        if not self.open or self.flushed or off + n > self.size:
            # outside an open batch, or too large: a plain (blocking) copy
            src = torch.from_numpy(arr)
            dst = out if out is not None else torch.empty(arr.shape, dtype=src.dtype, device=self.device)
            if n:
                dst.copy_(src)
            return dst
        if n >= (64 << 10):   # large: torch's multi-threaded host copy
            self.bufs[self.cur][off:off + n].copy_(torch.from_numpy(arr.reshape(-1).view(np.uint8)))
        else:                 # small: numpy's lower per-call cost
            self.np_bufs[self.cur][off:off + n] = arr.reshape(-1).view(np.uint8)
        self.off = off + n
        dev = self.arena[off:off + n].view(torch.from_numpy(arr[:0].reshape(-1)).dtype).view(arr.shape)
        if out is None:
            return dev
        self.pending.append((out, dev))
        return out


class PackedOutputs:
    """Device results of one step gathered into one device buffer by a single
    tm_pack_copy launch and brought to pinned host memory by one copy.

    begin() starts a new set; add() appends a contiguous device array under a
    name (8-byte aligned, in call order); mark() queues the pack launch, the
    copy of the used bytes and an event on the current stream; after wait(),
    host(name) is a numpy view of that array's values.
    """

    def __init__(self, device, nbytes, max_parts=16):
        from . import _lib
        self._lib = _lib
        self.device = device
        self.size = int(nbytes)
        self.dev = torch.empty(self.size, dtype=torch.uint8, device=device)
        self.buf = torch.empty(self.size, dtype=torch.uint8, pin_memory=True)
        self.np_buf = self.buf.numpy()
        self.parts = (_lib.PackPart * max_parts)()
        self.views = {}
        self.n = 0
        self.off = 0
        self.event = None

    def begin(self):
        self.views = {}
        self.n = 0
        self.off = 0

    def add(self, name, src):
        """Queue the device tensor src (contiguous) for the next mark()."""
        assert src.is_contiguous() and src.device.type == "cuda"
        nb = src.numel() * src.element_size()
        if self.off + nb > self.size or self.n == len(self.parts):
            raise ValueError(f"PackedOutputs: {name} ({nb} bytes) exceeds the buffer")
        dt = torch.empty(0, dtype=src.dtype).numpy().dtype
        self.views[name] = (self.off, nb, dt, tuple(src.shape))
        if nb:
            p = self.parts[self.n]
            p.src, p.bytes, p.offset = src.data_ptr(), nb, self.off
            self.n += 1
        self.off = (self.off + nb + 7) & ~7

    def mark(self):
        st = torch.cuda.current_stream(self.device)
        if self.n:
            self._lib.lib().tm_pack_copy(self.n, self.parts, self._lib.ptr(self.dev),
                                         self._lib.ctypes.c_void_p(st.cuda_stream))
            self.buf[:self.off].copy_(self.dev[:self.off], non_blocking=True)
        self.event = torch.cuda.Event()
        self.event.record(st)

    def wait(self):
        if self.event is not None:
            self.event.synchronize()

    def host(self, name):
        off, nb, dt, shape = self.views[name]
        return self.np_buf[off:off + nb].view(dt).reshape(shape)
