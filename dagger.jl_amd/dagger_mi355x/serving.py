"""Serving form: a device-resident request loop for SPD factor-and-solve.

Beside the plain drop-in.  Requests (matrix id, n, kappa, nrhs, tolerance)
arrive over time; the server admits them by an HBM budget as worker slots
and memory free up, overlaps their device work on several HIP streams
(one host worker thread per stream, ctypes releases the GIL), and caches
factors on the GPU so a request reusing an earlier matrix only solves.

Per request (worker stream s):
  * new matrix: A = H diag(s) H^T written into a pooled resident DArray of
    padded size n_pad (a ladder of sizes; [A 0; 0 I] keeps x exact), then
    the FP32 tiled factor (MixedCholesky) on s;
  * solve with FP64 iterative refinement to the request's tolerance.
Pooled buffers keep their device addresses, so the cached plan of each
(buffer, size) is reused across requests.  Every computation depends
only on its own request, so a result is the same alone and in the mix.

Latency = completion - arrival (wall clock); solves/s = completed requests
/ (last completion - first arrival).
"""
import ctypes as C
import threading
import time
from collections import OrderedDict, deque

import numpy as np

from ._lib import call
from .darray import DArray, Blocks
from .datadeps import PlanOptions
from .mixed import MixedCholesky, HouseholderSpec, householder_fill
from .linalg import cholesky_, ldiv_

LADDER = (64, 128, 192, 256, 384, 512, 768, 1024, 1536, 2048, 3072, 4096, 6144, 8192)


def pad_size(n):
    for s in LADDER:
        if s >= n:
            return s
    return int(np.ceil(n / 1024) * 1024)


def tile_for(npad):
    if npad <= 1024:
        return npad
    return 512 if npad <= 4096 else 1024


class Request:
    __slots__ = ("rid", "mid", "n", "kappa", "nrhs", "tol", "t_arrive", "reuse", "seed",
                 "t_start", "t_done", "berr", "iters", "x")

    def __init__(self, rid, mid, n, kappa, nrhs, tol, t_arrive, reuse, seed):
        self.rid, self.mid, self.n, self.kappa = rid, mid, n, kappa
        self.nrhs, self.tol, self.t_arrive, self.reuse, self.seed = nrhs, tol, t_arrive, reuse, seed
        self.t_start = self.t_done = None
        self.berr = None
        self.iters = None
        self.x = None

    def rhs(self):
        """B (n x nrhs) of this request, from its own seeded generator."""
        rng = np.random.default_rng(1_000_003 * self.seed + 17)
        return rng.standard_normal((self.n, self.nrhs))


def make_trace(rate, count, seed, n_lo=64, n_hi=8192, k_lo=1e2, k_hi=1e6, reuse=0.3,
               tols=(1e-8, 1e-13), nrhs_max=32):
    """Poisson arrivals at `rate` req/s; n ~ LogUniform[n_lo, n_hi];
    kappa ~ LogUniform[k_lo, k_hi]; nrhs ~ U{1..nrhs_max}; tol uniform over
    `tols`; a fraction `reuse` of requests reuses an earlier matrix."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, count))
    reqs, mats = [], []
    for i in range(count):
        if mats and rng.random() < reuse:
            mid, n, kappa = mats[rng.integers(len(mats))]
            re = True
        else:
            n = int(np.clip(np.floor(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi + 1)))),
                            n_lo, n_hi))
            kappa = float(np.exp(rng.uniform(np.log(k_lo), np.log(k_hi))))
            mid = len(mats)
            mats.append((mid, n, kappa))
            re = False
        reqs.append(Request(i, mid, n, kappa, int(rng.integers(1, nrhs_max + 1)),
                            float(tols[rng.integers(len(tols))]), float(t[i]), re, seed * 7919 + i))
    return reqs


class _Entry:
    """A cached factor: the FP64 A, its FP32 factor, readiness."""

    def __init__(self, npad):
        self.npad = npad
        self.A = None
        self.F = None
        self.solver = None
        self.ready = threading.Event()
        self.users = 0
        self.error = None

    def nbytes(self):
        return self.npad * self.npad * 12


class Server:
    def __init__(self, workers=4, budget_bytes=200e9, cache=True, options=None):
        self.nworkers = workers
        self.budget = float(budget_bytes)
        self.cache_enabled = cache
        # Eager issue: a hipGraph capture on one worker's stream is
        # invalidated when another worker makes a synchronous runtime call
        # meanwhile, so concurrent workers never capture.
        self.options = options or PlanOptions(num_streams=2, batch=True, lookahead=1,
                                              use_graph=False)
        self.streams = []
        for _ in range(workers):
            h = C.c_void_p()
            call("dm_stream_create", C.byref(h))
            self.streams.append(h)
        self.lock = threading.Condition()
        self.cache = OrderedDict()       # mid -> _Entry (LRU order)
        self.pool = {}                   # npad -> [(A, F)] free buffers
        self.used = 0.0                  # bytes of entries + pooled buffers
        self.queue = deque()
        self.done = []
        self.stop = False

    # ---------------------------------------------------------- memory
    def _evict_for(self, need):
        """Free pooled buffers, then unused LRU cache entries, until `need`
        bytes fit in the budget.  Caller holds the lock."""
        for npad in list(self.pool):
            while self.pool[npad] and self.used + need > self.budget:
                A, F = self.pool[npad].pop()
                A.unsafe_free()
                F.unsafe_free()
                self.used -= npad * npad * 12
        for mid in list(self.cache):
            if self.used + need <= self.budget:
                break
            e = self.cache[mid]
            if e.users == 0 and e.ready.is_set():
                del self.cache[mid]
                self._release(e)
        return self.used + need <= self.budget

    def _release(self, e):
        if e.A is not None:
            self.pool.setdefault(e.npad, []).append((e.A, e.F))
            e.A = e.F = e.solver = None

    def _take_buffers(self, npad):
        lst = self.pool.get(npad)
        if lst:
            return lst.pop()
        nb = tile_for(npad)
        A = DArray(np.float64, (npad, npad), Blocks(nb, nb))
        F = DArray(np.float32, (npad, npad), Blocks(nb, nb))
        return A, F

    # ---------------------------------------------------------- work
    def _admit(self, req):
        """Reserve the request's factor (cached or new).  Returns the entry,
        or None when memory does not allow it yet.  Caller holds the lock."""
        key = (req.mid, req.n, req.kappa)  # a matrix is identified by all three
        e = self.cache.get(key) if self.cache_enabled else None
        if e is not None:
            self.cache.move_to_end(key)
            e.users += 1
            return e, False
        npad = pad_size(req.n)
        need = npad * npad * 12
        pooled = bool(self.pool.get(npad))
        if not pooled and not self._evict_for(need):
            return None, False
        e = _Entry(npad)
        e.users = 1
        if not pooled:
            self.used += need
        if self.cache_enabled:
            self.cache[key] = e
        return e, True

    def _run(self, req, e, fresh, stream):
        if fresh:
            try:
                with self.lock:
                    A, F = self._take_buffers(e.npad)
                spec = HouseholderSpec(req.n, req.kappa, 1_000_003 + req.mid, 64, stream)
                householder_fill(A, spec, stream)
                e.A, e.F = A, F
                e.solver = MixedCholesky(A, options=self.options, stream=stream, F=F)
            except Exception as ex:  # surfaced per request
                e.error = ex
            e.ready.set()
        else:
            e.ready.wait()
        if e.error is not None:
            raise e.error
        npad = e.npad
        B = np.zeros((npad, req.nrhs))
        B[: req.n] = req.rhs()
        nb = tile_for(npad)
        DB = DArray(np.float64, (npad, req.nrhs), Blocks(nb, nb))
        DB.copy_from_host(B, stream)
        solver = MixedCholesky.__new__(MixedCholesky)
        solver.__dict__.update(e.solver.__dict__)
        solver.stream = stream
        X, info = solver.solve(DB, tol=req.tol, max_iter=30)
        req.x = X.collect(stream)[: req.n]
        req.berr = float(np.max(info["berr"]))
        req.iters = info["iterations"]
        DB.unsafe_free()

    def _worker(self, w):
        stream = self.streams[w]
        while True:
            with self.lock:
                while not self.queue and not self.stop:
                    self.lock.wait()
                if not self.queue and self.stop:
                    return
                req, e, fresh = self.queue.popleft()
            req.t_start = time.perf_counter()
            try:
                self._run(req, e, fresh, stream)
            finally:
                req.t_done = time.perf_counter()
                with self.lock:
                    e.users -= 1
                    if not self.cache_enabled:
                        self._release(e)
                    self.done.append(req)
                    self.lock.notify_all()

    def run(self, trace, time_scale=1.0):
        """Replay `trace` (arrival times relative to start) and serve it.
        A server may run several traces one after another; pooled buffers,
        cached factors and plans carry over."""
        with self.lock:
            self.stop = False
            self.done = []
        threads = [threading.Thread(target=self._worker, args=(w,), daemon=True)
                   for w in range(self.nworkers)]
        for t in threads:
            t.start()
        t0 = time.perf_counter()
        pending = deque(sorted(trace, key=lambda r: r.t_arrive))
        waiting = deque()
        while pending or waiting:
            now = time.perf_counter() - t0
            while pending and pending[0].t_arrive * time_scale <= now:
                r = pending.popleft()
                r.t_arrive = t0 + r.t_arrive * time_scale
                waiting.append(r)
            with self.lock:
                while waiting and len(self.queue) < 2 * self.nworkers:
                    e, fresh = self._admit(waiting[0])
                    if e is None:
                        break
                    self.queue.append((waiting.popleft(), e, fresh))
                    self.lock.notify_all()
                if pending and not waiting:
                    delay = pending[0].t_arrive * time_scale - (time.perf_counter() - t0)
                    if delay > 0:
                        self.lock.wait(timeout=min(delay, 0.05))
                elif waiting:
                    self.lock.wait(timeout=0.002)
        with self.lock:
            self.stop = True
            self.lock.notify_all()
        for t in threads:
            t.join()
        return summarize(self.done)

    def close(self):
        for h in self.streams:
            call("dm_stream_destroy", h)
        self.streams = []


def serve_one_at_a_time(trace, options=None):
    """Baseline: requests strictly one at a time in arrival order, each a
    plain FP64 tiled cholesky! + ldiv! on its own freshly built matrix."""
    options = options or PlanOptions(num_streams=2, batch=True, lookahead=1, use_graph=True)
    t0 = time.perf_counter()
    done = []
    for r in sorted(trace, key=lambda r: r.t_arrive):
        arrive = t0 + r.t_arrive
        now = time.perf_counter()
        if now < arrive:
            time.sleep(arrive - now)
        r.t_arrive = arrive
        r.t_start = time.perf_counter()
        npad = pad_size(r.n)
        nb = tile_for(npad)
        A = DArray(np.float64, (npad, npad), Blocks(nb, nb))
        spec = HouseholderSpec(r.n, r.kappa, 1_000_003 + r.mid, 64)
        householder_fill(A, spec)
        F = cholesky_(A, options=options, check_hermitian=False)
        B = np.zeros((npad, r.nrhs))
        B[: r.n] = r.rhs()
        DB = DArray.from_array(B, Blocks(nb, nb))
        ldiv_(F, DB, options=options)
        r.x = DB.collect()[: r.n]
        r.t_done = time.perf_counter()
        A.unsafe_free()
        DB.unsafe_free()
        done.append(r)
    return summarize(done)


def summarize(done):
    if not done:
        return dict(completed=0)
    lat = np.array([r.t_done - r.t_arrive for r in done])
    span = max(r.t_done for r in done) - min(r.t_arrive for r in done)
    return dict(completed=len(done), solves_per_s=len(done) / span,
                p50_ms=float(np.percentile(lat, 50) * 1e3),
                p99_ms=float(np.percentile(lat, 99) * 1e3),
                mean_ms=float(lat.mean() * 1e3), span_s=span,
                max_berr=float(max((r.berr or 0.0) for r in done)))
