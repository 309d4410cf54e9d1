"""The ctypes binding waits for torch's stream before the library sees a
tensor's address.

The library's streams are non-blocking, so a library call must not start
while work queued on torch's stream still reads or writes the memory it is
handed (torch's caching allocator gives freed memory to the next tensor at
once).  These checks run on the CPU with a stand-in stream.
"""
import ctypes

import torch

from cuvhs_amd import native as N


class _FakeStream:
    def __init__(self):
        self.syncs = 0

    def synchronize(self):
        self.syncs += 1


class _FakeCudaTensor:
    is_cuda = True
    device = "cuda:0"

    def data_ptr(self):
        return 0x1000


def test_ptr_synchronises_torch_stream_for_device_tensors(monkeypatch):
    s = _FakeStream()
    seen = []

    def current_stream(dev=None):
        seen.append(dev)
        return s

    monkeypatch.setattr(torch.cuda, "current_stream", current_stream)
    p = N.ptr(_FakeCudaTensor())
    assert isinstance(p, ctypes.c_void_p) and p.value == 0x1000
    assert s.syncs == 1 and seen == ["cuda:0"]


def test_ptr_leaves_host_tensors_and_none_alone(monkeypatch):
    def current_stream(dev=None):
        raise AssertionError("no stream for a host tensor")

    monkeypatch.setattr(torch.cuda, "current_stream", current_stream)
    t = torch.zeros(4)
    assert N.ptr(t).value == t.data_ptr()
    assert N.ptr(None) is None

