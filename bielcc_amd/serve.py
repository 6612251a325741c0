"""Serving loop for scattering queries on one GPU.

A query is (mesh, k, incidence directions with polarisations, observation
points); the answer is the scattered-plus-incident E and H at the points for
every incidence (one current set per direction).  The reference has no
solver or server; this is the request loop around the path:

- continuous batching: worker threads, each with its own library context
  (own HIP stream, scratch and BLAS handle), take queries from one FIFO queue
  as soon as they are free, so queries that arrive while others run are
  admitted as slots free up and their kernels overlap on the device;
- resident state: meshes stay uploaded per worker; the factors of the
  collocation matrix stay on the device per (mesh, k) in an LRU cache
  (FP64 LU in the "fp64" form; the FP64 matrix plus its FP32 LU in the
  "mixed" form, whose solves run FGMRES in FP64 preconditioned by the FP32
  factors, past `refine_tol` down to the FP64 rounding floor).  Systems too
  ill-conditioned for FP32 factors (where even FGMRES stalls above
  `refine_tol`, or the FP32 LU meets a zero pivot) get FP64 factors once,
  beside the FP32 ones, and are solved with them from then on;
  their fields are summed in FP64 too (the currents then carry a large
  non-radiating part whose exterior field cancels);
- admission by HBM budget: factors, every query's buffers (right-hand
  sides, currents, points, fields and the field driver's scratch) and the
  workers' grow-only buffer pools are reserved from one byte budget; a query
  waits while its bytes do not fit, and idle factors are evicted
  least-recently-used first.  Query buffers come from the worker's pool, so
  serving a query allocates and frees nothing on the device once the pool
  has grown (a device free would stall every other worker's stream);
- metrics: per-query latency (arrival to result), p50 / p99, solves/s.

The library calls release the GIL (ctypes), so workers drive their streams
concurrently.
"""
from __future__ import annotations

import collections
import dataclasses
import queue
import threading
import time
from typing import Callable, Optional

import numpy as np

from . import api


@dataclasses.dataclass
class Query:
    """One scattering query.  dirs [nrhs][3] unit propagation directions,
    pols [nrhs][3] polarisations (orthogonal to dirs), points [M][3]."""
    mesh: str
    k: float
    dirs: np.ndarray
    pols: np.ndarray
    points: np.ndarray
    qid: int = 0
    arrival: float = 0.0

    @property
    def nrhs(self) -> int:
        return int(np.asarray(self.dirs).reshape(-1, 3).shape[0])


@dataclasses.dataclass
class Result:
    qid: int
    E: np.ndarray          # [nrhs][M][3] complex, total field (incident included)
    H: np.ndarray          # [nrhs][M][3] complex, total field (incident included)
    arrival: float
    start: float
    end: float
    factored: bool         # this query built the factors it used
    refine_iters: int = 0
    refine_resid: float = 0.0
    fp64_fallback: bool = False   # mixed form: solved with FP64 factors
    phases: dict = dataclasses.field(default_factory=dict)   # host seconds: build, wait, solve, field, copy

    @property
    def latency(self) -> float:
        return self.end - self.arrival


class Budget:
    """Device bytes reserved against a fixed budget, with eviction of idle
    cache entries when a reservation does not fit."""

    def __init__(self, total: int):
        self.total = int(total)
        self.used = 0
        self.fixed = 0         # part of `used` held until the server closes (worker pools, library scratch)
        self.peak = 0          # highest `used` seen
        self.cv = threading.Condition()

    def reserve(self, nbytes: int, evict: Callable[[int], tuple], timeout: float = 600.0,
                fixed: bool = False) -> None:
        """Block until nbytes fit.  evict(need) runs under the lock: it takes
        idle cache entries out of the cache and returns (bytes, buffers).  The
        buffers are freed after the lock is released (a device free waits for
        the device, and other workers must not wait on the lock meanwhile); their
        bytes stay counted as used until they are freed, so the device never
        holds more than the budget.  fixed: the bytes stay held until the server
        closes.  Raises at once if nbytes cannot fit beside the fixed bytes (no
        eviction or finished query would ever make room)."""
        nbytes = int(nbytes)
        deadline = time.monotonic() + timeout
        while True:
            victims, vbytes = [], 0
            with self.cv:
                if self.fixed + nbytes > self.total:
                    raise MemoryError(f"request of {nbytes} B does not fit the HBM budget of {self.total} B "
                                      f"beside {self.fixed} B of worker buffers held for good")
                while self.used + nbytes > self.total:
                    freed, bufs = evict(self.used + nbytes - self.total)
                    if freed > 0:
                        victims, vbytes = bufs, freed
                        break
                    left = deadline - time.monotonic()
                    if left <= 0:
                        raise TimeoutError("HBM budget: reservation did not fit in time")
                    self.cv.wait(timeout=min(left, 1.0))
                    if self.fixed + nbytes > self.total:   # fixed bytes grew meanwhile
                        raise MemoryError(f"request of {nbytes} B does not fit the HBM budget of "
                                          f"{self.total} B beside {self.fixed} B of worker buffers held for good")
                else:
                    self.used += nbytes
                    if fixed:
                        self.fixed += nbytes
                    self.peak = max(self.peak, self.used)
                    return
            try:
                for b in victims:
                    b.free()
            finally:
                self.release(vbytes)

    def charge(self, nbytes: int, evict: Callable[[int], tuple], fixed: bool = False) -> None:
        """Count bytes that are already allocated (no waiting: the caller may hold
        pins that others wait for); idle entries are evicted to get back under
        the budget as far as possible.  fixed: held until the server closes."""
        with self.cv:
            self.used += int(nbytes)
            if fixed:
                self.fixed += int(nbytes)
            self.peak = max(self.peak, self.used)
            over = self.used - self.total
            freed, victims = evict(over) if over > 0 else (0, [])
        try:
            for b in victims:
                b.free()
        finally:
            if freed:
                self.release(freed)

    def release(self, nbytes: int, fixed: bool = False) -> None:
        with self.cv:
            self.used -= int(nbytes)
            if fixed:
                self.fixed -= int(nbytes)
            self.cv.notify_all()


class _Entry:
    def __init__(self, key, nbytes):
        self.key = key
        self.nbytes = nbytes
        self.ready = False
        self.failed: Optional[BaseException] = None
        self.refs = 0
        self.bufs: dict = {}
        self.owner = None        # worker context that allocated the buffers
        self.upgrade = threading.Lock()   # one FP64 fallback factorisation at a time


class FactorCache:
    """Device-resident factors per (mesh, k) with reference counts and LRU
    eviction of entries nobody uses.  Must be used under Budget.cv."""

    def __init__(self, budget: Budget):
        self.budget = budget
        self.entries: "collections.OrderedDict[tuple, _Entry]" = collections.OrderedDict()
        self.evictions = 0

    def evict(self, need: int) -> tuple:
        """Take idle entries (least recently used first) out of the cache until
        `need` bytes are covered or none is left; returns (bytes, buffers): the
        caller frees the buffers outside the lock.  Caller holds budget.cv."""
        freed, bufs = 0, []
        for key in list(self.entries.keys()):
            if freed >= need:
                break
            e = self.entries[key]
            if e.refs == 0 and e.ready:
                bufs.extend(e.bufs.values())
                e.bufs = {}
                del self.entries[key]
                freed += e.nbytes
                self.evictions += 1
        return freed, bufs

    def pin(self, key, nbytes: int) -> tuple:
        """Pin the entry of key (creating it, not yet built, if absent).
        Returns (entry, builder): builder is True for the caller that created
        it and must reserve its bytes and build it."""
        with self.budget.cv:
            e = self.entries.get(key)
            if e is not None:
                e.refs += 1
                self.entries.move_to_end(key)
                return e, False
            e = _Entry(key, nbytes)
            e.refs = 1
            self.entries[key] = e
            return e, True

    def wait_ready(self, e: _Entry, timeout: float = 1200.0) -> None:
        """Wait (holding no budget bytes) until another caller has built e.
        The caller keeps its pin either way and releases it."""
        cv = self.budget.cv
        deadline = time.monotonic() + timeout
        with cv:
            while not e.ready and e.failed is None:
                left = deadline - time.monotonic()
                if left <= 0:
                    raise TimeoutError(f"factors of {e.key} not ready in time")
                cv.wait(timeout=min(left, 1.0))
            if e.failed is not None:
                raise RuntimeError(f"factorisation of {e.key} failed") from e.failed

    def finish(self, e: _Entry, exc: Optional[BaseException] = None) -> None:
        """Builder: mark e ready, or failed (its buffers freed, the entry dropped;
        the builder returns the entry's bytes to the budget itself)."""
        cv = self.budget.cv
        bufs = []
        with cv:
            if exc is None:
                e.ready = True
            else:
                e.failed = exc
                bufs = list(e.bufs.values())
                e.bufs = {}
                if self.entries.get(e.key) is e:
                    del self.entries[e.key]
            cv.notify_all()
        for b in bufs:          # device frees outside the lock
            b.free()

    def release(self, e: _Entry) -> None:
        with self.budget.cv:
            e.refs -= 1
            self.budget.cv.notify_all()


def _field_scratch_bytes(M: int, N: int, nrhs: int) -> int:
    """Upper bound of the field driver's partial sums (bielcc_field_batch)."""
    bx = (M + 255) // 256
    nsplit = max(1, min((2048 + bx - 1) // bx, (N + 63) // 64))
    return nsplit * nrhs * 12 * M * 8 + 64 * M


class Server:
    """Continuous-batching server over a pool of meshes.

    meshes: name -> vertex array [n][3|4][3].  form: "fp64" (FP64 LU) or
    "mixed" (FP32 LU + FP64 refinement, FP32 far-pair field sums).

    refine_max_iter: corrections (FP32 solves after the first) a refinement may take.

    refine_ferr_exp = e > 0: a solve does not stop at refine_tol alone but goes on
    to the FP64 rounding floor (FGMRES; classic refinement, refine_mode 0:
    until its forward-error estimate, cond(Z) from the residual's contraction
    times the residual, is below 10^-e).  A residual target alone does not bound
    the error of an ill-conditioned system: refined to a 1e-10 residual, a
    low-frequency elongated ellipsoid (cond about 4e6) kept a 1.9e-5 field
    error.  0: residual only.
    """

    _upgrade_gate = threading.Lock()   # one FP64 upgrade at a time (bounds the budget overshoot)

    def __init__(self, meshes: dict, form: str = "mixed", workers: int = 3, hbm_budget: float = 200e9,
                 num_param=None, refine_tol: float = 1e-10, device: int = 0, refine_ferr_exp: int = 7,
                 refine_max_iter: int = 60):
        if form not in ("fp64", "mixed"):
            raise ValueError("form must be 'fp64' or 'mixed'")
        self.meshes = {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in meshes.items()}
        self.form = form
        self.np = num_param if num_param is not None else api.num_param()
        self.refine_tol = refine_tol
        self.refine_max_iter = int(refine_max_iter)
        self.budget = Budget(int(hbm_budget))
        self.cache = FactorCache(self.budget)
        self.q: "queue.Queue[Optional[Query]]" = queue.Queue()
        self.results: dict = {}
        self.errors: dict = {}
        self.upgrades: list = []          # (mesh, k, reason) of entries given FP64 factors
        self.lock = threading.Lock()
        self.done_cv = threading.Condition(self.lock)
        self.ctxs = [api.Context(device) for _ in range(workers)]
        for c in self.ctxs:
            c.set_tuning("refine_ferr_exp", int(refine_ferr_exp))
        self.dev_meshes = [dict() for _ in range(workers)]
        self.pools = [dict() for _ in range(workers)]   # name -> DeviceArray, grow-only
        self.scratch_hw = [0] * workers   # bytes charged for the context's own grow-only scratch
        self.threads = [threading.Thread(target=self._run, args=(w,), daemon=True) for w in range(workers)]
        for t in self.threads:
            t.start()

    # -- public
    def submit(self, q: Query) -> None:
        if q.arrival == 0.0:
            q.arrival = time.perf_counter()
        self.q.put(q)

    def wait(self, qids, timeout: float = 1200.0) -> dict:
        deadline = time.monotonic() + timeout
        with self.done_cv:
            while not all(i in self.results or i in self.errors for i in qids):
                left = deadline - time.monotonic()
                if left <= 0:
                    raise TimeoutError("queries did not finish in time")
                self.done_cv.wait(timeout=min(left, 1.0))
        bad = {i: self.errors[i] for i in qids if i in self.errors}
        if bad:
            raise RuntimeError(f"queries failed: {bad}")
        return {i: self.results[i] for i in qids}

    def close(self) -> None:
        for _ in self.threads:
            self.q.put(None)
        for t in self.threads:
            t.join(timeout=600)
        with self.budget.cv:
            for e in list(self.cache.entries.values()):
                for b in e.bufs.values():
                    b.free()
            self.cache.entries.clear()
        for m in self.dev_meshes:
            for dm in m.values():
                dm.close()
        for pool in self.pools:
            for d in pool.values():
                d.free()
        for c in self.ctxs:
            c.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- worker
    def _mesh(self, w: int, name: str) -> api.Mesh:
        m = self.dev_meshes[w].get(name)
        if m is None:
            m = self.ctxs[w].mesh(self.meshes[name])
            self.dev_meshes[w][name] = m
        return m

    def _buf(self, w: int, name: str, nbytes: int) -> api.DeviceArray:
        """Worker w's pooled buffer `name` of at least nbytes; growth is charged
        to the budget for good."""
        pool = self.pools[w]
        d = pool.get(name)
        if d is not None and d.nbytes >= nbytes:
            return d
        grow = int(nbytes * 1.25) + 4096
        self.budget.reserve(grow, self.cache.evict, fixed=True)
        if d is not None:
            d.free()
            self.budget.release(d.nbytes, fixed=True)
        d = self.ctxs[w].empty(grow)
        pool[name] = d
        return d

    def _charge_scratch(self, w: int, ctx: api.Context) -> None:
        """Charge what the context's grow-only work buffers (field partial sums,
        near lists, operator matrices, solve work) hold beyond what was charged."""
        actual = ctx.get_tuning("scratch_kb") << 10
        if actual > self.scratch_hw[w]:
            self.budget.charge(actual - self.scratch_hw[w], self.cache.evict, fixed=True)
            self.scratch_hw[w] = actual

    def _factor_bytes(self, n: int) -> int:
        n2 = 2 * n
        per = 16 + 8 if self.form == "mixed" else 16   # FP64 matrix (+ FP32 factors)
        return per * n2 * n2 + 4 * n2

    def _build(self, w: int, m: api.Mesh, k: float, e: _Entry) -> None:
        ctx = self.ctxs[w]
        n2 = 2 * m.n
        e.owner = ctx
        e.bufs["Z"] = ctx.empty(n2 * n2 * 16)
        e.bufs["piv"] = ctx.empty(n2 * 4)
        m.assemble_into(e.bufs["Z"], n2, self.np, k)
        if self.form == "mixed":
            e.bufs["F"] = ctx.empty(n2 * n2 * 8)
            try:
                ctx.lu32_factor(e.bufs["Z"], n2, n2, e.bufs["F"], n2, e.bufs["piv"])
            except api.BielccError as err:
                if err.code != 5:          # only a zero FP32 pivot is handled here
                    raise
                ctx.sync()
                self._upgrade(w, e, str(err))
        else:
            ctx.lu_factor(e.bufs["Z"], n2, n2, e.bufs["piv"])
        ctx.sync()

    def _build_entry(self, w: int, m: api.Mesh, k: float, e: _Entry, fbytes: int) -> None:
        """Builder of entry e: reserve its bytes (waiting callers of the same key
        hold none), build it and mark it ready; on failure drop it and return
        every byte it held (e.nbytes includes FP64 factors the build added)."""
        try:
            self.budget.reserve(fbytes, self.cache.evict)
        except BaseException as exc:
            self.cache.finish(e, exc)
            raise
        try:
            self._build(w, m, k, e)
        except BaseException as exc:
            nb = e.nbytes
            self.cache.finish(e, exc)
            self.budget.release(nb)
            raise
        self.cache.finish(e)

    def _upgrade(self, w: int, e: _Entry, reason: str = "") -> None:
        """FP64 factors beside the FP32 ones (mixed form), for a system FP32
        factors cannot refine.  Done once per entry."""
        with e.upgrade:
            if "LU64" in e.bufs:
                return
            with self.lock:
                self.upgrades.append({"mesh": e.key[0], "k": e.key[1], "reason": reason})
            ctx = self.ctxs[w]
            n2 = int(round(np.sqrt(e.bufs["Z"].nbytes / 16)))
            extra = 16 * n2 * n2 + 4 * n2
            # the caller holds a pin on e, so it must not wait for budget bytes (pinned
            # workers waiting on one another could deadlock): the bytes are charged at
            # once, idle entries evicted, and one upgrade at a time bounds the overshoot
            # to one FP64 factorisation
            Server._upgrade_gate.acquire()
            try:
                self.budget.charge(extra, self.cache.evict)
            except BaseException:
                Server._upgrade_gate.release()
                raise
            lu = piv = None
            try:
                lu, piv = ctx.empty(16 * n2 * n2), ctx.empty(4 * n2)
                ctx.copy(lu, e.bufs["Z"], 16 * n2 * n2)
                ctx.lu_factor(lu, n2, n2, piv)
                ctx.sync()
            except BaseException:
                for b in (lu, piv):
                    if b is not None:
                        b.free()
                self.budget.release(extra)
                raise
            finally:
                Server._upgrade_gate.release()
            with self.budget.cv:
                e.bufs["LU64"], e.bufs["piv64"] = lu, piv
                e.nbytes += extra

    def _run(self, w: int) -> None:
        ctx = self.ctxs[w]
        while True:
            q = self.q.get()
            if q is None:
                return
            try:
                res = self._serve(w, ctx, q)
                with self.done_cv:
                    self.results[q.qid] = res
                    self.done_cv.notify_all()
            except BaseException as exc:   # report, keep serving
                with self.done_cv:
                    self.errors[q.qid] = repr(exc)
                    self.done_cv.notify_all()

    def _serve(self, w: int, ctx: api.Context, q: Query) -> Result:
        start = time.perf_counter()
        m = self._mesh(w, q.mesh)
        n, n2 = m.n, 2 * m.n
        dirs = np.asarray(q.dirs, dtype=np.float64).reshape(-1, 3)
        pols = np.asarray(q.pols, dtype=np.float64).reshape(-1, 3)
        pts = np.ascontiguousarray(q.points, dtype=np.float64).reshape(-1, 3)
        nrhs, M = dirs.shape[0], pts.shape[0]
        # the field driver's scratch is context-owned and grow-only: charge its high-water mark
        grow = max(0, _field_scratch_bytes(M, n, nrhs) - self.scratch_hw[w])
        if grow:
            self.budget.reserve(grow, self.cache.evict, fixed=True)
            self.scratch_hw[w] += grow
        # the query's buffers grow (and are charged) before any cache pin is taken: a
        # worker that waits for budget bytes must hold no pin, or workers that each
        # pin an entry and wait could block one another until nothing is evictable
        dB = self._buf(w, "B", n2 * nrhs * 16)
        dJ = self._buf(w, "J", n * nrhs * 48)
        dP = self._buf(w, "P", M * 24)
        dE, dH = self._buf(w, "E", nrhs * M * 48), self._buf(w, "H", nrhs * M * 48)
        fbytes = self._factor_bytes(n)
        ph: dict = {}
        e, builder = self.cache.pin((q.mesh, float(q.k)), fbytes)
        try:
            if builder:
                t0 = time.perf_counter()
                self._build_entry(w, m, q.k, e, fbytes)
                ph["build"] = time.perf_counter() - t0
            else:
                t0 = time.perf_counter()
                self.cache.wait_ready(e)
                ph["wait"] = time.perf_counter() - t0
            kvec = q.k * dirs
            t0 = time.perf_counter()
            ctx.upload(dP, pts)
            m.rhs_planewave_into(dB, n2, kvec, pols)
            iters, resid, fallback = 0, 0.0, False
            # decided once: another worker may add FP64 factors to the entry while this
            # one refines, and a refined solution must not be solved a second time
            fallback = self.form == "mixed" and "LU64" in e.bufs
            if self.form == "mixed" and not fallback:
                try:
                    iters, resid = ctx.refine_solve(e.bufs["Z"], n2, n2, e.bufs["F"], n2, e.bufs["piv"], dB, nrhs,
                                                    n2, self.refine_tol, self.refine_max_iter)
                except api.BielccError as err:
                    if err.code != 4:      # refinement stalled: FP64 factors for this entry
                        raise
                    self._upgrade(w, e, str(err))
                    m.rhs_planewave_into(dB, n2, kvec, pols)
                    fallback = True
            if fallback:
                ctx.lu_solve(e.bufs["LU64"], n2, n2, e.bufs["piv64"], dB, nrhs, n2)
            elif self.form == "fp64":
                ctx.lu_solve(e.bufs["Z"], n2, n2, e.bufs["piv"], dB, nrhs, n2)
            m.currents_into(dB, n2, nrhs, dJ)
            ctx.sync()
            ph["solve"] = time.perf_counter() - t0
            t0 = time.perf_counter()
            inc = np.concatenate([kvec, pols], axis=1)
            # near a resonance the currents carry a large non-radiating part whose
            # exterior field cancels: FP32 far sums would lose that cancellation
            prec = api.PREC_MIXED if self.form == "mixed" and not fallback else api.PREC_FP64
            m.field_batch_into(self.np, q.k, nrhs, dJ, dP, M, dE, dH, inc=inc, precision=prec)
            # normalised units (eps = mu = 1, omega = k): the solved currents are
            # j' = multE j_E with multE = i/k, so E = K[j'] + E_inc is the total E and
            # H = R[j_E] + H_inc = R[j'] / multE + PlaneWaveFieldH (freq = k / 2 pi, mu = 1)
            ctx.add_incident_h(M, nrhs, dP, inc, q.k / (2.0 * np.pi), dH, mu=1.0, scale=-1j * q.k)
            ctx.sync()
            ph["field"] = time.perf_counter() - t0
            self._charge_scratch(w, ctx)
            t0 = time.perf_counter()
            E = dE.to_numpy((nrhs, M, 3), np.complex128)
            H = dH.to_numpy((nrhs, M, 3), np.complex128)
            ph["copy"] = time.perf_counter() - t0
        finally:
            if e.ready or not builder:
                self.cache.release(e)
        return Result(q.qid, E, H, q.arrival, start, time.perf_counter(), builder, iters, resid, fallback, ph)


def percentiles(lat) -> dict:
    lat = np.sort(np.asarray(lat, dtype=np.float64))
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "max_s": float(lat.max())}


def poisson_run(server: Server, queries, qps: float, seed: int = 0, progress=None) -> dict:
    """Submit queries at Poisson arrival times (rate qps), wait for all, and
    return results with latency percentiles and throughput.  progress(done,
    total), if given, is called as results come in."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / qps, size=len(queries))
    t0 = time.perf_counter()
    t = t0
    for q, g in zip(queries, gaps):
        t += g
        now = time.perf_counter()
        if t > now:
            time.sleep(t - now)
        q.arrival = time.perf_counter()
        server.submit(q)
    if progress is not None:
        ids, last, t_last = [q.qid for q in queries], -1, time.monotonic()
        while True:
            with server.done_cv:
                done = sum(i in server.results or i in server.errors for i in ids)
            if done != last or time.monotonic() - t_last > 30.0:
                progress(done, len(ids))
                last, t_last = done, time.monotonic()
            if done == len(ids):
                break
            time.sleep(1.0)
    res = server.wait([q.qid for q in queries])
    t1 = max(r.end for r in res.values())
    solves = sum(q.nrhs for q in queries)
    out = {"queries": len(queries), "solves": solves, "wall_s": t1 - t0, "solves_per_s": solves / (t1 - t0),
           "factorisations": sum(r.factored for r in res.values()), "evictions": server.cache.evictions,
           "fp64_fallbacks": sum(r.fp64_fallback for r in res.values()),
           "fp64_upgrades": list(server.upgrades),
           "refine_corrections_max": max((r.refine_iters for r in res.values()), default=0),
           "lu_panel_aborts": sum(max(c.get_tuning("lu_panel_aborts"), 0) for c in server.ctxs),
           "trsv_expired": sum(c.get_tuning("trsv_expired") for c in server.ctxs),
           "trsv_ticket_faults": sum(c.get_tuning("trsv_ticket_faults") for c in server.ctxs),
           "mixed_fixup_points": sum(c.get_tuning("mixed_fixup_points") for c in server.ctxs),
           "hbm_budget_peak_gb": round(server.budget.peak / 1e9, 2),
           "worker_seconds": {k: round(sum(r.phases.get(k, 0.0) for r in res.values()), 3)
                              for k in ("build", "wait", "solve", "field", "copy")}}
    out.update(percentiles([r.latency for r in res.values()]))
    return {"results": res, "stats": out}
