"""Host logic of the serving loop (no GPU): HBM budget admission, factor cache
pinning, LRU eviction of idle entries only, one builder per key, failure
propagation."""
import threading
import time

import pytest

from bielcc_amd.serve import Budget, FactorCache


class FakeBuf:
    def __init__(self, log, name):
        self.log, self.name = log, name

    def free(self):
        self.log.append(self.name)


def _ready_entry(cache, key, nbytes, log):
    e, builder = cache.pin(key, nbytes)
    assert builder
    cache.budget.reserve(nbytes, cache.evict)
    e.bufs = {"Z": FakeBuf(log, key)}
    cache.finish(e)
    return e


def test_lru_eviction_skips_pinned_entries():
    log = []
    b = Budget(100)
    c = FactorCache(b)
    e1 = _ready_entry(c, "a", 40, log)
    e2 = _ready_entry(c, "b", 40, log)
    c.release(e1)                      # a idle, b still pinned
    b.reserve(30, c.evict)             # needs 10 more bytes: evicts a only
    assert log == ["a"] and "b" in c.entries and b.used == 70
    c.release(e2)
    b.release(30)
    b.reserve(90, c.evict)             # now b goes too
    assert log == ["a", "b"] and b.used == 90 and c.evictions == 2


def test_reservation_waits_for_release_and_rejects_oversize():
    b = Budget(100)
    c = FactorCache(b)
    b.reserve(80, c.evict)
    done = []

    def waiter():
        b.reserve(50, c.evict)
        done.append(time.monotonic())

    t = threading.Thread(target=waiter)
    t.start()
    time.sleep(0.2)
    assert not done
    b.release(80)
    t.join(timeout=5)
    assert done and b.used == 50
    with pytest.raises(MemoryError):
        b.reserve(101, c.evict)


def test_one_builder_per_key_and_waiters_see_the_result():
    b = Budget(1000)
    c = FactorCache(b)
    builds, got = [], []
    start = threading.Barrier(4)

    def worker(i):
        start.wait()
        e, builder = c.pin("k", 100)
        if builder:
            b.reserve(100, c.evict)
            time.sleep(0.2)
            e.bufs = {"Z": FakeBuf([], "k")}
            builds.append(i)
            c.finish(e)
        else:
            c.wait_ready(e)
        got.append(e.ready)
        c.release(e)

    ts = [threading.Thread(target=worker, args=(i,)) for i in range(4)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=10)
    assert len(builds) == 1 and got == [True] * 4 and c.entries["k"].refs == 0


def test_build_failure_reaches_waiters_and_drops_the_entry():
    b = Budget(1000)
    c = FactorCache(b)
    e, builder = c.pin("k", 100)
    e2, builder2 = c.pin("k", 100)
    assert builder and not builder2
    c.finish(e, RuntimeError("boom"))
    with pytest.raises(RuntimeError):
        c.wait_ready(e2)
    assert "k" not in c.entries
