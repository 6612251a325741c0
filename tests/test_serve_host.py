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


def test_evicted_buffers_are_freed_outside_the_lock():
    """A device free waits for the device: it must not run while the budget's
    lock is held (every other worker would stall in pin, release and reserve).
    The evicted bytes stay counted until the free is done."""
    b = Budget(100)
    c = FactorCache(b)
    seen = []

    class CheckBuf:
        def free(self):
            got = b.cv.acquire(blocking=False)   # the lock is not reentrant: fails if this thread holds it
            if got:
                b.cv.release()
            seen.append((got, b.used))

    e, _ = c.pin("a", 60)
    b.reserve(60, c.evict)
    e.bufs = {"Z": CheckBuf(), "piv": CheckBuf()}
    c.finish(e)
    c.release(e)
    b.reserve(70, c.evict)
    assert [g for g, _ in seen] == [True, True]
    assert all(u == 60 for _, u in seen)        # still charged while being freed
    assert b.used == 70 and "a" not in c.entries and c.evictions == 1


class _FakeDev:
    def __init__(self, log, nbytes):
        self.log, self.nbytes = log, nbytes

    def free(self):
        self.log.append(self.nbytes)


class _FakeCtx:
    """The context calls the server makes, with a chosen one failing."""

    def __init__(self, log, fail_lu64=False, fail_lu32=False):
        self.log, self.fail_lu64, self.fail_lu32 = log, fail_lu64, fail_lu32

    def empty(self, nbytes):
        return _FakeDev(self.log, nbytes)

    def copy(self, dst, src, nbytes):
        pass

    def lu_factor(self, A, n, lda, piv):
        if self.fail_lu64:
            from bielcc_amd.api import BielccError
            raise BielccError(5, "zero pivot")

    def lu32_factor(self, A, n, lda, F, ldf, piv):
        if self.fail_lu32:
            from bielcc_amd.api import BielccError
            raise BielccError(5, "zero FP32 pivot")

    def sync(self):
        pass


def _bare_server(ctx, form="mixed", total=10 ** 9):
    import collections
    from bielcc_amd.serve import Server
    s = Server.__new__(Server)
    s.form, s.budget = form, Budget(total)
    s.cache = FactorCache(s.budget)
    s.ctxs, s.lock, s.upgrades = [ctx], threading.Lock(), []
    s.np = None
    return s


def test_failed_upgrade_returns_its_bytes_and_buffers():
    from bielcc_amd.api import BielccError
    log = []
    s = _bare_server(_FakeCtx(log, fail_lu64=True))
    e, _ = s.cache.pin(("m", 1.0), 100)
    e.bufs = {"Z": _FakeDev([], 16 * 30 * 30)}
    before = s.budget.used
    with pytest.raises(BielccError):
        s._upgrade(0, e, "stalled")
    assert s.budget.used == before
    assert sorted(log) == sorted([16 * 30 * 30, 4 * 30])   # the FP64 factors and pivots were freed
    assert "LU64" not in e.bufs and e.nbytes == 100


def test_failed_build_returns_the_entrys_bytes():
    """A build whose FP32 LU meets a zero pivot gets FP64 factors (its entry
    grows); if the FP64 LU then fails too, every byte goes back."""
    from bielcc_amd.api import BielccError
    log = []
    s = _bare_server(_FakeCtx(log, fail_lu32=True, fail_lu64=True))

    class M:
        n = 15

        def assemble_into(self, *a, **k):
            pass
    key = ("m", 1.0)
    fbytes = s._factor_bytes(M.n)
    e, builder = s.cache.pin(key, fbytes)
    with pytest.raises(BielccError):
        s._build_entry(0, M(), 1.0, e, fbytes)
    assert s.budget.used == 0 and key not in s.cache.entries
    assert len(s.upgrades) == 1                 # the FP64 factors were tried


def test_charge_of_held_bytes_never_waits():
    """Bytes a worker already holds (grow-only library work buffers) are charged
    without waiting - the worker may hold pins others wait for - and idle entries
    are evicted to get back under the budget."""
    log = []
    b = Budget(100)
    c = FactorCache(b)
    e = _ready_entry(c, "a", 60, log)
    b.charge(30, c.evict)                 # fits: nothing evicted
    assert b.used == 90 and log == []
    t0 = time.monotonic()
    b.charge(30, c.evict)                 # over, but "a" is pinned: counted, no wait
    assert time.monotonic() - t0 < 0.5 and b.used == 120 and log == []
    c.release(e)
    b.charge(0, c.evict)                  # "a" idle now: evicted back under the budget
    assert log == ["a"] and b.used == 60


def test_reservation_beside_fixed_bytes_fails_at_once():
    """Bytes held until the server closes (worker pools, library scratch) can never
    be freed by an eviction: a reservation that cannot fit beside them raises at
    once instead of waiting out its timeout, also when they grow while it waits."""
    b = Budget(100)
    c = FactorCache(b)
    b.reserve(30, c.evict, fixed=True)
    b.charge(40, c.evict, fixed=True)
    assert b.fixed == 70 and b.used == 70
    t0 = time.monotonic()
    with pytest.raises(MemoryError):
        b.reserve(31, c.evict)
    assert time.monotonic() - t0 < 0.5
    b.reserve(20, c.evict)                # fits beside the fixed bytes
    err = []

    def waiter():
        try:
            b.reserve(25, c.evict, timeout=30.0)
        except MemoryError as x:
            err.append(x)

    t = threading.Thread(target=waiter)
    t.start()
    time.sleep(0.2)
    b.charge(10, c.evict, fixed=True)     # fixed 80: 25 more can never fit
    t.join(timeout=5)
    assert not t.is_alive() and err
    b.release(10, fixed=True)
    b.release(20)
    assert b.fixed == 70 and b.used == 70
