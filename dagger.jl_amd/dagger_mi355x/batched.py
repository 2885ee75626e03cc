"""Batching form: many uneven SPD systems factored and solved by one
persistent kernel (dm_batched_posv), with the panel width chosen per size
bucket by an autotuner written here.

Layout: every system's A_m (n x n) and B_m (n x nrhs) are column-major,
back to back in two device slabs (like a DArray chunk each); A_m's leading
dimension is n rounded up to even (so every A_m starts 16-byte aligned and
the tiled path's matrix-core GEMMs take their 2-wide vector loads for odd n
too), B_m's is n.  The persistent kernel's work queue is an atomic counter
over a largest-first order.  BatchedSystems.solve sends systems of at
least TILED_MIN_N rows (by default every system) through the tiled
level-synchronous plan instead (dm_batched_tiled_*): one step of batched
launches per 64-column block over all systems at once, the bulk of the
flops in batched MFMA GEMMs.
"""
import ctypes as C
import time

import numpy as np

from . import _lib
from ._lib import call
from .darray import DeviceBuffer, DeviceInt

_DT = {np.dtype(np.float64): _lib.DM_F64, np.dtype(np.float32): _lib.DM_F32}
PANELS = (16, 32, 64)
# systems of at least this many rows take the tiled level-synchronous path
# in BatchedSystems.solve (outer tiles of TILED_NB columns), the rest the
# persistent kernel.  Measured on 4096 systems n ~ LogUniform[32, 2048]
# (FP64): every system tiled 46.3 ms; n >= 64 tiled 52.7 ms; n >= 128
# 53.3 ms; all in the persistent kernel 155 ms.
TILED_MIN_N = 0
TILED_NB = 256


class BatchItem(C.Structure):
    _fields_ = [("a_off", C.c_int64), ("b_off", C.c_int64), ("n", C.c_int), ("nrhs", C.c_int),
                ("panel", C.c_int), ("lda", C.c_int)]


def bucket_of(n):
    """Size bucket: floor(log2 n) (32-63 -> 5, ..., 1024-2047 -> 10, 2048 -> 11)."""
    return int(np.floor(np.log2(max(n, 1))))


class BatchedSystems:
    """Device-resident batch of SPD systems (A_m's leading dimension: n
    rounded up to a multiple of ld_pad; 1 packs the systems back to back)."""

    def __init__(self, sizes, nrhs, dtype=np.float64, ld_pad=2):
        self.sizes = np.asarray(sizes, dtype=np.int64)
        self.nrhs = np.asarray(nrhs, dtype=np.int64)
        if len(self.sizes) != len(self.nrhs):
            raise ValueError("sizes and nrhs differ in length")
        if (self.nrhs < 0).any() or (self.nrhs > 8).any():
            raise ValueError("nrhs must be in [0, 8]")
        self.dtype = np.dtype(dtype)
        es = self.dtype.itemsize
        if ld_pad < 1:
            raise ValueError("ld_pad must be >= 1")
        self.ld = (self.sizes + ld_pad - 1) // ld_pad * ld_pad  # A_m's leading dimension
        self.a_off = np.concatenate([[0], np.cumsum(self.ld * self.sizes)[:-1]]).astype(np.int64)
        self.b_off = np.concatenate([[0], np.cumsum(self.sizes * self.nrhs)[:-1]]).astype(np.int64)
        self.a_elems = int((self.ld * self.sizes).sum())
        self.b_elems = int((self.sizes * self.nrhs).sum())
        self.A = DeviceBuffer(max(self.a_elems * es, 8))
        self.B = DeviceBuffer(max(self.b_elems * es, 8))
        call("dm_fill_i32", C.c_void_p(self.A.ptr), 0, C.c_size_t(self.A.nbytes // 4), None)
        self.info = DeviceInt(max(len(self.sizes), 1))
        self.count = len(self.sizes)

    def host_matrix(self, flat, m):
        """A_m (n x n) out of a host copy of the whole A slab."""
        n, ld, a0 = int(self.sizes[m]), int(self.ld[m]), int(self.a_off[m])
        return flat[a0:a0 + ld * n].reshape(ld, n, order="F")[:n]

    def a_ptr(self, m):
        return self.A.ptr + int(self.a_off[m]) * self.dtype.itemsize

    def b_ptr(self, m):
        return self.B.ptr + int(self.b_off[m]) * self.dtype.itemsize

    # ------------------------------------------------------------ data
    def generate(self, seed):
        """A_m = G G^T / n + I (G from the Philox stream seed + m) and B_m
        N(0,1) (stream seed + 10^6 + m), all on the device (FP64 generation,
        converted when the batch is FP32)."""
        nmax = int(self.sizes.max()) if self.count else 0
        G = DeviceBuffer(max(nmax * nmax * 8, 8))
        F = DeviceBuffer(max(nmax * nmax * 8, 8))
        S = DeviceBuffer(max(nmax * nmax * 8, 8))
        es = self.dtype.itemsize
        for m in range(self.count):
            n = int(self.sizes[m])
            if n == 0:  # an empty system: nothing to generate
                continue
            call("dm_fill_randn_d", C.c_void_p(G.ptr), C.c_size_t(n * n), C.c_uint64(seed + m),
                 C.c_uint64(0), None)
            call("dm_dgemm", b"N", b"T", n, n, n, 1.0 / n, C.c_void_p(G.ptr), n,
                 C.c_void_p(G.ptr), n, 0.0, C.c_void_p(F.ptr), n, None)
            t = _lib.dm_tiled(C.c_void_p(S.ptr), n, n, n, n, _DT[self.dtype])
            if self.dtype == np.float64:
                call("dm_tiled_pack", C.byref(t), C.c_void_p(F.ptr), n, 1.0, None)
            else:
                t64 = _lib.dm_tiled(C.c_void_p(G.ptr), n, n, n, n, _lib.DM_F64)
                call("dm_tiled_pack", C.byref(t64), C.c_void_p(F.ptr), n, 1.0, None)
                call("dm_tiled_convert", C.byref(t), C.byref(t64), None)
            call("dm_memcpy2d_d2d", C.c_void_p(self.a_ptr(m)), C.c_size_t(int(self.ld[m]) * es),
                 C.c_void_p(S.ptr), C.c_size_t(n * es), C.c_size_t(n * es), C.c_size_t(n), None)
            r = int(self.nrhs[m])
            if r:
                if self.dtype == np.float64:
                    call("dm_fill_randn_d", C.c_void_p(self.b_ptr(m)), C.c_size_t(n * r),
                         C.c_uint64(seed + 1000000 + m), C.c_uint64(0), None)
                else:
                    call("dm_fill_randn_d", C.c_void_p(G.ptr), C.c_size_t(n * r),
                         C.c_uint64(seed + 1000000 + m), C.c_uint64(0), None)
                    tb = _lib.dm_tiled(C.c_void_p(self.b_ptr(m)), n, r, n, r, _lib.DM_F32)
                    tg = _lib.dm_tiled(C.c_void_p(G.ptr), n, r, n, r, _lib.DM_F64)
                    call("dm_tiled_convert", C.byref(tb), C.byref(tg), None)
        call("dm_stream_sync", None)
        G.free()
        F.free()
        S.free()
        return self

    def get_system(self, m):
        """Host copies of (A_m, B_m) (column-major)."""
        n, r = int(self.sizes[m]), int(self.nrhs[m])
        es = self.dtype.itemsize
        A = np.empty((n, n), dtype=self.dtype, order="F")
        call("dm_memcpy2d_d2h", A.ctypes.data_as(C.c_void_p), C.c_size_t(n * es),
             C.c_void_p(self.a_ptr(m)), C.c_size_t(int(self.ld[m]) * es), C.c_size_t(n * es),
             C.c_size_t(n), None)
        B = np.empty((n, r), dtype=self.dtype, order="F")
        if r:
            call("dm_memcpy_d2h", B.ctypes.data_as(C.c_void_p), C.c_void_p(self.b_ptr(m)),
                 C.c_size_t(B.nbytes), None)
        return A, B

    def put_system(self, m, A=None, B=None):
        """Overwrite A_m and/or B_m from host arrays (n x n, n x nrhs)."""
        n, r = int(self.sizes[m]), int(self.nrhs[m])
        es = self.dtype.itemsize
        if A is not None:
            A = np.asfortranarray(A, dtype=self.dtype)
            if A.shape != (n, n):
                raise ValueError("A_m must be n x n")
            call("dm_memcpy2d_h2d", C.c_void_p(self.a_ptr(m)), C.c_size_t(int(self.ld[m]) * es),
                 A.ctypes.data_as(C.c_void_p), C.c_size_t(n * es), C.c_size_t(n * es),
                 C.c_size_t(n), None)
        if B is not None and r:
            B = np.asfortranarray(B, dtype=self.dtype)
            if B.shape != (n, r):
                raise ValueError("B_m must be n x nrhs")
            call("dm_memcpy_h2d", C.c_void_p(self.b_ptr(m)), B.ctypes.data_as(C.c_void_p),
                 C.c_size_t(B.nbytes), None)
        call("dm_stream_sync", None)

    def snapshot(self):
        """Device copies of both slabs (to restore between timed runs)."""
        a = DeviceBuffer(max(self.a_elems * self.dtype.itemsize, 8))
        b = DeviceBuffer(max(self.b_elems * self.dtype.itemsize, 8))
        call("dm_memcpy_d2d", C.c_void_p(a.ptr), C.c_void_p(self.A.ptr),
             C.c_size_t(self.a_elems * self.dtype.itemsize), None)
        call("dm_memcpy_d2d", C.c_void_p(b.ptr), C.c_void_p(self.B.ptr),
             C.c_size_t(self.b_elems * self.dtype.itemsize), None)
        return a, b

    def restore(self, snap):
        a, b = snap
        call("dm_memcpy_d2d", C.c_void_p(self.A.ptr), C.c_void_p(a.ptr),
             C.c_size_t(self.a_elems * self.dtype.itemsize), None)
        call("dm_memcpy_d2d", C.c_void_p(self.B.ptr), C.c_void_p(b.ptr),
             C.c_size_t(self.b_elems * self.dtype.itemsize), None)

    # ------------------------------------------------------------ solve
    def items(self, panels):
        """dm_batch_item array; panels: dict bucket -> P (default 32)."""
        it = (BatchItem * max(self.count, 1))()
        for m in range(self.count):
            n = int(self.sizes[m])
            it[m].a_off = int(self.a_off[m])
            it[m].b_off = int(self.b_off[m])
            it[m].n = n
            it[m].nrhs = int(self.nrhs[m])
            it[m].panel = int(panels.get(bucket_of(n), 32))
            it[m].lda = int(self.ld[m])
        return it

    def order(self, subset=None):
        idx = np.arange(self.count) if subset is None else np.asarray(subset)
        o = idx[np.argsort(-self.sizes[idx], kind="stable")]
        return np.ascontiguousarray(o, dtype=np.int32)

    def posv(self, panels=None, factor_only=False, subset=None, workgroups=0, stream=None):
        """Factor (and solve) the batch in one persistent launch."""
        panels = panels or {}
        it = self.items(panels)
        order = self.order(subset)
        call("dm_batched_posv", _DT[self.dtype], self.count, C.c_void_p(self.A.ptr),
             C.c_void_p(self.B.ptr), C.cast(it, C.c_void_p), len(order),
             order.ctypes.data_as(C.c_void_p), C.c_void_p(self.info.ptr),
             int(bool(factor_only)), int(workgroups), stream)
        return self

    def split(self, tiled_min):
        """(small, large): systems below / at least tiled_min rows."""
        idx = np.arange(self.count)
        large = self.sizes >= tiled_min
        return idx[~large], idx[large]

    def tiled_plan(self, subset, nb=TILED_NB, tile_nb=None):
        """Tiled level-synchronous plan over `subset` (cached per subset, nb
        and per-bucket tiles: the plan holds device tables and a working
        slab for them)."""
        key = (tuple(int(i) for i in subset), nb, tuple(sorted((tile_nb or {}).items())))
        cache = self.__dict__.setdefault("_plans", {})
        if key not in cache:
            cache[key] = TiledPlan(self, subset, nb, tile_nb)
        return cache[key]

    def solve(self, panels=None, factor_only=False, tiled_min=TILED_MIN_N, nb=TILED_NB,
              workgroups=0, overlap=True, tile_nb=None):
        """Hybrid: systems with n < tiled_min in one persistent launch (panel
        width per size bucket), the rest through the tiled plan.  overlap:
        the persistent launch runs on a second stream beside the plan's
        launches (forked from and joined back into the default stream).
        tile_nb: per-bucket outer tiles of the tiled plan (autotune_tiled)."""
        small, large = self.split(tiled_min)
        if len(small) and len(large) and overlap:
            side, fork, join = self._side()
            call("dm_event_record", fork, None)
            call("dm_stream_wait_event", side, fork)
            self.posv(panels=panels, factor_only=factor_only, subset=small,
                      workgroups=workgroups, stream=side)
            self.tiled_plan(large, nb, tile_nb).run(factor_only)
            call("dm_event_record", join, side)
            call("dm_stream_wait_event", None, join)
            return self
        if len(small):
            self.posv(panels=panels, factor_only=factor_only, subset=small,
                      workgroups=workgroups)
        if len(large):
            self.tiled_plan(large, nb, tile_nb).run(factor_only)
        return self

    def _side(self):
        if "_side_res" not in self.__dict__:
            side, fork, join = C.c_void_p(), C.c_void_p(), C.c_void_p()
            call("dm_stream_create", C.byref(side))
            call("dm_event_create", C.byref(fork))
            call("dm_event_create", C.byref(join))
            self._side_res = (side, fork, join)
        return self._side_res

    def close(self):
        """Release the cached tiled plans and the side stream."""
        for p in self.__dict__.pop("_plans", {}).values():
            p.close()
        res = self.__dict__.pop("_side_res", None)
        if res:
            call("dm_stream_sync", res[0])
            call("dm_event_destroy", res[1])
            call("dm_event_destroy", res[2])
            call("dm_stream_destroy", res[0])

    def infos(self):
        return self.info.get()[: self.count]

    def flops(self):
        """sum n^3/3 (factor) and sum 2 n^2 nrhs (solves)."""
        n = self.sizes.astype(np.float64)
        return float((n ** 3 / 3).sum()), float((2 * n * n * self.nrhs).sum())


class TiledPlan:
    """dm_batched_tiled_* over a subset of a BatchedSystems batch.  tile_nb:
    dict size bucket -> outer tile of the factorization for that bucket's
    systems (others: nb, which is also the solves' tile)."""

    def __init__(self, systems, subset, nb=TILED_NB, tile_nb=None):
        self.systems = systems
        it = systems.items({})
        order = systems.order(subset)
        tile_nb = tile_nb or {}
        per = np.array([int(tile_nb.get(bucket_of(int(n)), 0)) for n in systems.sizes],
                       dtype=np.int32)
        self.item_nb = np.ascontiguousarray(per if len(per) else np.zeros(1, np.int32))
        self.handle = C.c_void_p()
        call("dm_batched_tiled_create", C.byref(self.handle), _DT[systems.dtype], systems.count,
             C.c_void_p(systems.A.ptr), C.c_void_p(systems.B.ptr), C.cast(it, C.c_void_p),
             len(order), order.ctypes.data_as(C.c_void_p), int(nb),
             self.item_nb.ctypes.data_as(C.c_void_p))

    def run(self, factor_only=False, stream=None):
        call("dm_batched_tiled_run", self.handle, C.c_void_p(self.systems.info.ptr),
             int(bool(factor_only)), stream)
        return self

    def close(self):
        if self.handle:
            call("dm_batched_tiled_destroy", self.handle)
            self.handle = C.c_void_p()


def autotune(systems, sample=24, reps=2):
    """Pick the panel width per size bucket: for each bucket, time the
    factorization of up to `sample` of its systems with each candidate P
    (HIP events, factor only, systems restored between trials) and keep the
    fastest.  Returns dict bucket -> P and the timing table."""
    snap = systems.snapshot()
    e0, e1 = C.c_void_p(), C.c_void_p()
    call("dm_event_create", C.byref(e0))
    call("dm_event_create", C.byref(e1))
    buckets = {}
    for m in range(systems.count):
        buckets.setdefault(bucket_of(int(systems.sizes[m])), []).append(m)
    choice, table = {}, {}
    for b, members in sorted(buckets.items()):
        sub = members[:sample]
        best = None
        for P in PANELS:
            if P > 2 * int(systems.sizes[sub].max()):
                continue
            ms_best = None
            for _ in range(reps):
                systems.restore(snap)
                call("dm_event_record", e0, None)
                systems.posv(panels={b: P}, factor_only=True, subset=sub)
                call("dm_event_record", e1, None)
                call("dm_event_sync", e1)
                ms = C.c_float()
                call("dm_event_elapsed_ms", C.byref(ms), e0, e1)
                ms_best = ms.value if ms_best is None else min(ms_best, ms.value)
            table[(b, P)] = ms_best
            if best is None or ms_best < best[1]:
                best = (P, ms_best)
        choice[b] = best[0]
    systems.restore(snap)
    call("dm_event_destroy", e0)
    call("dm_event_destroy", e1)
    return choice, table


TILE_CANDIDATES = (128, 192, 256, 320, 384, 512)


def autotune_tiled(systems, candidates=TILE_CANDIDATES, sample=64, reps=2):
    """Pick the tiled plan's outer tile per size bucket: for each bucket,
    time the factorization of up to `sample` of its systems (a tiled plan
    over just them, HIP events, systems restored between trials) with each
    candidate no larger than twice the bucket's largest n, and keep the
    fastest.  Returns dict bucket -> tile and the timing table."""
    snap = systems.snapshot()
    e0, e1 = C.c_void_p(), C.c_void_p()
    call("dm_event_create", C.byref(e0))
    call("dm_event_create", C.byref(e1))
    buckets = {}
    for m in range(systems.count):
        buckets.setdefault(bucket_of(int(systems.sizes[m])), []).append(m)
    choice, table = {}, {}
    try:
        for b, members in sorted(buckets.items()):
            sub = members[:sample]
            nmax = int(systems.sizes[sub].max())
            best = None
            for nb in candidates:
                if nb > max(2 * nmax, candidates[0]):
                    continue
                plan = TiledPlan(systems, sub, nb)
                ms_best = None
                for _ in range(reps):
                    systems.restore(snap)
                    call("dm_event_record", e0, None)
                    plan.run(factor_only=True)
                    call("dm_event_record", e1, None)
                    call("dm_event_sync", e1)
                    ms = C.c_float()
                    call("dm_event_elapsed_ms", C.byref(ms), e0, e1)
                    ms_best = ms.value if ms_best is None else min(ms_best, ms.value)
                plan.close()
                table[(b, nb)] = ms_best
                if best is None or ms_best < best[1]:
                    best = (nb, ms_best)
            choice[b] = best[0]
    finally:
        systems.restore(snap)
        for buf in snap:
            buf.free()
        call("dm_event_destroy", e0)
        call("dm_event_destroy", e1)
    return choice, table


def log_uniform_sizes(count, lo, hi, seed):
    """n ~ LogUniform[lo, hi] (integers), nrhs ~ U{1..8}, from numpy's
    PCG64 seeded generator (host-side problem description only)."""
    rng = np.random.default_rng(seed)
    sizes = np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi + 1), count))).astype(np.int64)
    sizes = np.clip(sizes, lo, hi)
    nrhs = rng.integers(1, 9, count)
    return sizes, nrhs
