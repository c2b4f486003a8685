"""Host-side bookkeeping of Array, with the library calls replaced by a
recorder (no device): an upload from the host mirror that may still be
queued is waited for before the array (and with it the mirror's memory) goes
away, also after later operations gave the array a newer job."""
import gc

import numpy as np
import pytest

from vulkpy_amd import vkarray


class FakeJob:
    def __init__(self, log, ticket):
        self.log, self.ticket = log, ticket

    def wait(self, timeout_ns=None):
        self.log.append(("wait", self.ticket))


class FakeGPU:
    _handle = 1

    def __init__(self):
        self.log = []
        self.tickets = 0

    def _job(self):
        self.tickets += 1
        return FakeJob(self.log, self.tickets)


@pytest.fixture
def fake(monkeypatch):
    gpu = FakeGPU()

    def recorder(name, *args):
        gpu.log.append((name,))

    monkeypatch.setattr(vkarray, "call", recorder)
    return gpu


def make(gpu, n=1000):
    # below the pinning threshold: the mirror is ordinary memory, no pool call
    return vkarray.Array(gpu, data=np.arange(n, dtype=np.float32))


def test_upload_waited_for_at_delete(fake):
    x = make(fake)
    assert ("vkh_upload",) in fake.log and x._upload_pending
    fake.log.clear()
    del x
    gc.collect()
    assert fake.log == [("wait", 1)]


def test_inplace_op_keeps_upload_pending(fake):
    x = make(fake)
    x *= 2.0                                  # new ticket; the upload may still be queued
    assert ("vkh_binary_scalar",) in fake.log
    assert x._upload_pending
    fake.log.clear()
    del x
    gc.collect()
    # the newest ticket covers the earlier upload on the in-order stream
    assert fake.log == [("wait", 2)]


def test_wait_clears_pending(fake):
    x = make(fake)
    x += 1.0
    x.wait()
    assert not x._upload_pending
    fake.log.clear()
    del x
    gc.collect()
    assert fake.log == []                     # nothing left to wait for


def test_result_of_an_op_has_no_pending_upload(fake):
    x = make(fake)
    y = x + 1.0
    assert not y._upload_pending
    fake.log.clear()
    del y
    gc.collect()
    assert fake.log == []


def test_delete_during_capture_waits_after_it(fake):
    """A finalizer that runs while a graph is captured (garbage collection at
    any allocation) must not wait on the capturing stream: the wait is queued
    and runs, mirror still held, when the capture ends."""
    x = make(fake)
    host = x._host
    fake.log.clear()
    vkarray._capture_begin()
    try:
        del x
        gc.collect()
        assert fake.log == []                 # nothing waited for inside the capture
        assert len(vkarray._deferred) == 1
    finally:
        vkarray._capture_end()
    assert fake.log == [("wait", 1)]
    assert vkarray._deferred == []
    assert host is not None


def test_capture_end_runs_queued_work_once_outermost():
    ran = []
    vkarray._capture_begin()
    vkarray._capture_begin()
    assert vkarray._after_capture(lambda: ran.append(1))
    vkarray._capture_end()
    assert ran == []                          # an outer capture still runs
    vkarray._capture_end()
    assert ran == [1]
    assert not vkarray._after_capture(lambda: ran.append(2))
    assert ran == [1]
