"""Host-side caches of the launch path (no GPU): the engine's per-workspace
launch argument lists and the ctypes call wrappers of the C-ABI binding."""
import pytest
import torch

from tensormol_amd.engine import BPEEEngine, Workspace


@pytest.fixture(scope="module")
def lib():
    import os
    from tensormol_amd import build, _lib
    if not os.path.exists(_lib.LIB_PATH):
        build.build()
    return _lib.lib()


class _RecordingLib:
    def __init__(self):
        self.calls = []

    def __getattr__(self, name):
        return lambda *args: self.calls.append((name, args))


def _engine():
    eng = BPEEEngine.__new__(BPEEEngine)   # only _call is exercised: no device state
    eng.lib = _RecordingLib()
    return eng


def test_launch_arguments_reused_until_layout_or_key_changes():
    eng = _engine()
    ws = Workspace()
    ws.epoch, ws.argc = 0, {}
    built = []

    def args(tag):
        built.append(tag)
        return (tag, 7)

    eng._call(ws, "tm_stage#energy", (1, 2), lambda: args("a"))
    eng._call(ws, "tm_stage#energy", (1, 2), lambda: args("b"))
    assert built == ["a"]
    assert eng.lib.calls == [("tm_stage", ("a", 7)), ("tm_stage", ("a", 7))]
    eng._call(ws, "tm_stage#energy", (1, 3), lambda: args("c"))     # positions or stream moved
    assert built == ["a", "c"] and eng.lib.calls[-1] == ("tm_stage", ("c", 7))
    ws.epoch += 1                                                   # a new layout
    eng._call(ws, "tm_stage#energy", (1, 3), lambda: args("d"))
    assert built == ["a", "c", "d"]
    # entries are per name: another user of the same entry point keeps its own list
    eng._call(ws, "tm_stage#grad", (1, 3), lambda: args("e"))
    assert built[-1] == "e" and eng.lib.calls[-1] == ("tm_stage", ("e", 7))


def test_set_topology_invalidates_cached_arguments():
    """Every layout change goes through set_topology, which bumps the layout
    version and drops the cached lists."""
    import inspect
    src = inspect.getsource(BPEEEngine.set_topology)
    assert "ws.epoch" in src and "ws.argc = {}" in src


def test_ctypes_wrappers_cached(lib):
    first = lib.tm_version
    assert lib.tm_version is first          # kept on the instance after the first lookup
    assert isinstance(lib.tm_version(), int)
    with pytest.raises(AttributeError):
        lib.tm_no_such_entry


def test_packed_outputs_layout_without_gpu(monkeypatch):
    """The pack records PackedOutputs builds: 8-byte aligned offsets in call
    order, whole byte counts, empty parts skipped (checked without launching)."""
    from tensormol_amd import staging
    from tensormol_amd import _lib as TL
    po = staging.PackedOutputs.__new__(staging.PackedOutputs)
    po.size, po.parts, po.views, po.n, po.off = 4096, (TL.PackPart * 8)(), {}, 0, 0

    class FakeDev:
        def __init__(self, n, dt, shape):
            self._t = torch.zeros(shape, dtype=dt)
            self.dtype, self.shape, self.device = dt, shape, torch.device("cuda", 0)

        def is_contiguous(self):
            return True

        def numel(self):
            return self._t.numel()

        def element_size(self):
            return self._t.element_size()

        def data_ptr(self):
            return 1 << 20

    po.add("e", FakeDev(3, torch.float64, (3,)))
    po.add("c", FakeDev(5, torch.int32, (5,)))
    po.add("z", FakeDev(0, torch.float64, (0, 3)))
    po.add("f", FakeDev(4, torch.float64, (2, 2)))
    assert po.n == 3
    assert [(p.bytes, p.offset) for p in po.parts[:3]] == [(24, 0), (20, 24), (32, 48)]
    assert po.views["c"][:2] == (24, 20) and po.views["z"][1] == 0
    assert po.off == 80
    with pytest.raises(ValueError):
        po.add("big", FakeDev(600, torch.float64, (600,)))
