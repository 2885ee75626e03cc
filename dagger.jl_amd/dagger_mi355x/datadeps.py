"""spawn_datadeps / In / Out / InOut over resident tiles.

Host mirror of Dagger's datadeps API (src/datadeps/queue.jl:64-106,
src/datadeps/aliasing.jl:164-181): inside a region, `spawn(f, args...)`
records a task whose array arguments carry In/Out/InOut annotations (bare
arguments count as In).  When the region closes, the tasks are handed in
submission order to the native executor (dm_dag_*), which derives the
dependencies with the datadeps rules, batches independent tasks of one
kernel shape, maps them onto HIP streams with events (optionally one
hipGraph) and runs them.  The region waits for completion, as
spawn_datadeps does (`wait_all`, src/datadeps/queue.jl:79).

The task functions are the tile BLAS/LAPACK calls the reference spawns:
`BLAS.gemm`, `BLAS.syrk`, `BLAS.trsm`, `BLAS.gemv`, `BLAS.trsv`,
`LAPACK.potrf` and `potrf_checked` (src/array/cholesky.jl:2-18), with
Julia's argument order.  This is the function remap ROCExt performs with
`move(::CPUProc, ::ROCProc, ::typeof(BLAS.gemm!))` (ext/ROCExt.jl:358-377).
"""
import ctypes as C
import threading

import numpy as np

from . import _lib
from ._lib import call, ch
from .darray import Tile, DeviceInt


class _Dep:
    __slots__ = ("x",)
    mode = _lib.DM_IN

    def __init__(self, x):
        self.x = x


class In(_Dep):
    mode = _lib.DM_IN


class Out(_Dep):
    mode = _lib.DM_OUT


class InOut(_Dep):
    mode = _lib.DM_INOUT


class _TileOp:
    def __init__(self, name, code):
        self.name, self.code = name, code

    def __repr__(self):
        return self.name


class _Namespace:
    def __init__(self, **kw):
        self.__dict__.update(kw)


BLAS = _Namespace(
    gemm=_TileOp("BLAS.gemm!", _lib.DM_OP_GEMM),
    syrk=_TileOp("BLAS.syrk!", _lib.DM_OP_SYRK),
    herk=_TileOp("BLAS.herk!", _lib.DM_OP_SYRK),  # complex tiles: C := a A^H A + b C
    trsm=_TileOp("BLAS.trsm!", _lib.DM_OP_TRSM),
    gemv=_TileOp("BLAS.gemv!", _lib.DM_OP_GEMV),
    trsv=_TileOp("BLAS.trsv!", _lib.DM_OP_TRSV),
)
LAPACK = _Namespace(potrf=_TileOp("LAPACK.potrf!", _lib.DM_OP_POTRF))
potrf_checked = _TileOp("potrf_checked!", _lib.DM_OP_POTRF)

_DT = {np.dtype(np.float64): _lib.DM_F64, np.dtype(np.float32): _lib.DM_F32,
       np.dtype(np.complex128): _lib.DM_C64, np.dtype(np.complex64): _lib.DM_C32}


def _unwrap(a):
    if isinstance(a, _Dep):
        return a.x, a.mode
    return a, _lib.DM_IN


def _tile(a, what):
    if not isinstance(a, Tile):
        raise TypeError(f"{what}: expected a DArray tile, got {type(a).__name__}")
    return a


def _dims(ok, what):
    """BLAS DimensionMismatch: raised before a task with inconsistent
    operand shapes reaches the executor (the kernels trust the shapes)."""
    if not ok:
        raise ValueError(f"DimensionMismatch: {what}")


def make_task(f, args):
    """Translate a Julia-style tile call into a dm_task.  Operand shapes are
    checked as BLAS.gemm!/syrk!/trsm!/gemv!/trsv! and LAPACK.potrf! check
    them (DimensionMismatch, raised as ValueError)."""
    t = _lib.dm_task()
    t.op = f.code
    t.side = t.uplo = t.transa = t.transb = t.diag = b"N"
    vals = [_unwrap(a) for a in args]
    if f.code == _lib.DM_OP_GEMM:
        (ta, _), (tb, _), (alpha, _), (A, ma), (B, mb), (beta, _), (Cm, mc) = vals
        A, B, Cm = _tile(A, "A"), _tile(B, "B"), _tile(Cm, "C")
        t.transa, t.transb = ch(ta), ch(tb)
        t.m, t.n = Cm.rows, Cm.cols
        t.k = A.rows if ta in ("T", "C") else A.cols
        am = A.cols if ta in ("T", "C") else A.rows
        bk, bn = (B.cols, B.rows) if tb in ("T", "C") else (B.rows, B.cols)
        _dims(am == t.m and bn == t.n and bk == t.k,
              f"gemm: op(A) {am}x{t.k}, op(B) {bk}x{bn}, C {t.m}x{t.n}")
        t.alpha, t.beta = alpha, beta
        t.a, t.lda, t.b, t.ldb, t.c, t.ldc = A.ptr, A.ld, B.ptr, B.ld, Cm.ptr, Cm.ld
        t.access[0], t.access[1], t.access[2] = ma, mb, mc
        dt = Cm.dtype
    elif f.code == _lib.DM_OP_SYRK:
        (uplo, _), (tr, _), (alpha, _), (A, ma), (beta, _), (Cm, mc) = vals
        A, Cm = _tile(A, "A"), _tile(Cm, "C")
        t.uplo, t.transa = ch(uplo), ch(tr)
        t.n = Cm.rows
        t.k = A.rows if tr in ("T", "C") else A.cols
        an = A.cols if tr in ("T", "C") else A.rows
        _dims(Cm.rows == Cm.cols and an == t.n,
              f"syrk: C {Cm.rows}x{Cm.cols} must be square with op(A)'s {an} rows")
        t.alpha, t.beta = alpha, beta
        t.a, t.lda, t.c, t.ldc = A.ptr, A.ld, Cm.ptr, Cm.ld
        t.access[0], t.access[2] = ma, mc
        dt = Cm.dtype
    elif f.code == _lib.DM_OP_TRSM:
        (side, _), (uplo, _), (tr, _), (diag, _), (alpha, _), (A, ma), (B, mb) = vals
        A, B = _tile(A, "A"), _tile(B, "B")
        t.side, t.uplo, t.transa, t.diag = ch(side), ch(uplo), ch(tr), ch(diag)
        t.m, t.n = B.rows, B.cols
        ka = B.rows if side == "L" else B.cols
        _dims(A.rows == A.cols == ka,
              f"trsm: A {A.rows}x{A.cols} must be square of order {ka}")
        t.alpha = alpha
        t.a, t.lda, t.c, t.ldc = A.ptr, A.ld, B.ptr, B.ld
        t.access[0], t.access[2] = ma, mb
        dt = B.dtype
    elif f.code == _lib.DM_OP_POTRF:
        uplo = vals[0][0]
        A, ma = vals[1]
        A = _tile(A, "A")
        _dims(A.rows == A.cols, f"potrf: A {A.rows}x{A.cols} must be square")
        t.uplo = ch(uplo)
        t.n = A.rows
        t.c, t.ldc = A.ptr, A.ld
        t.access[2] = ma if ma != _lib.DM_IN else _lib.DM_INOUT
        if len(vals) > 2:
            info = vals[2][0]
            if not isinstance(info, DeviceInt):
                raise TypeError("potrf info must be a DeviceInt")
            t.info = info.ptr
        # global row offset of a DArray diagonal tile, for the global info
        if A.owner is not None and A.index is not None:
            t.k = A.index[0] * A.owner.mb
        dt = A.dtype
    elif f.code == _lib.DM_OP_GEMV:
        (tr, _), (alpha, _), (A, ma), (x, mx), (beta, _), (y, my) = vals
        A, x, y = _tile(A, "A"), _tile(x, "x"), _tile(y, "y")
        t.transa = ch(tr)
        t.m, t.n = A.rows, A.cols
        lx, ly = (A.rows, A.cols) if tr in ("T", "C") else (A.cols, A.rows)
        _dims(x.rows * x.cols == lx and y.rows * y.cols == ly and x.cols == 1 and y.cols == 1,
              f"gemv: op(A) {ly}x{lx}, x {x.rows}x{x.cols}, y {y.rows}x{y.cols}")
        t.alpha, t.beta = alpha, beta
        t.a, t.lda, t.b, t.ldb, t.c, t.ldc = A.ptr, A.ld, x.ptr, 1, y.ptr, 1
        t.access[0], t.access[1], t.access[2] = ma, mx, my
        dt = A.dtype
    elif f.code == _lib.DM_OP_TRSV:
        (uplo, _), (tr, _), (diag, _), (A, ma), (x, mx) = vals
        A, x = _tile(A, "A"), _tile(x, "x")
        t.uplo, t.transa, t.diag = ch(uplo), ch(tr), ch(diag)
        t.n = A.rows
        _dims(A.rows == A.cols and x.rows == A.rows and x.cols == 1,
              f"trsv: A {A.rows}x{A.cols} must be square, x {x.rows}x{x.cols} of its order")
        t.a, t.lda, t.c, t.ldc = A.ptr, A.ld, x.ptr, 1
        t.access[0], t.access[2] = ma, mx
        dt = A.dtype
    else:
        raise ValueError(f"unsupported task function {f!r}")
    t.dtype = _DT[np.dtype(dt)]
    return t


class PlanOptions:
    """Executor knobs: streams, batching, look-ahead depth, hipGraph, the
    arithmetic of FP32 GEMM/SYRK groups (f32_update: 0 FP32 matrix cores,
    6 or 3 BF16 matrix cores with 3- or 2-term operand splitting), and
    whether FP32 plans factor their diagonal tiles in FP64 (f64_diag)."""

    def __init__(self, num_streams=2, batch=True, lookahead=1, use_graph=False, profile=False,
                 f32_update=0, f64_diag=False):
        if f32_update not in (0, 3, 6):
            raise ValueError("f32_update must be 0, 3 or 6")
        self.num_streams, self.batch = num_streams, batch
        self.lookahead, self.use_graph, self.profile = lookahead, use_graph, profile
        self.f32_update = f32_update
        self.f64_diag = bool(f64_diag)

    def c(self):
        return _lib.dm_plan_opts(int(self.num_streams), int(bool(self.batch)),
                                 int(self.lookahead), int(bool(self.use_graph)),
                                 int(bool(self.profile)), int(self.f32_update),
                                 int(self.f64_diag))

    def replace(self, **kw):
        d = dict(num_streams=self.num_streams, batch=self.batch, lookahead=self.lookahead,
                 use_graph=self.use_graph, profile=self.profile, f32_update=self.f32_update,
                 f64_diag=self.f64_diag)
        d.update(kw)
        return PlanOptions(**d)


SEQUENTIAL = PlanOptions(num_streams=1, batch=False, lookahead=0, use_graph=False)


class Dag:
    """A recorded task list planned by the native executor."""

    def __init__(self, tasks=(), handle=None):
        self.tasks = list(tasks)
        if handle is None:
            handle = C.c_void_p()
            call("dm_dag_create", C.byref(handle))
        self.h = handle
        for t in self.tasks:
            call("dm_dag_add_task", self.h, C.byref(t))
        self.planned = False

    @classmethod
    def cholesky(cls, A, uplo="U", info=None):
        """The tiled Cholesky DAG of a resident DArray (dm_dag_build_cholesky)."""
        h = C.c_void_p()
        t = A.tiled()
        call("dm_dag_build_cholesky", C.byref(h), uplo.encode(), C.byref(t),
             C.c_void_p(info.ptr if info is not None else 0))
        return cls(handle=h)

    @classmethod
    def potrs(cls, A, B, uplo="U"):
        """The two trsm! regions of ldiv! (dm_dag_build_potrs)."""
        h = C.c_void_p()
        ta, tb = A.tiled(), B.tiled()
        call("dm_dag_build_potrs", C.byref(h), uplo.encode(), C.byref(ta), C.byref(tb))
        return cls(handle=h)

    def profile(self):
        """Per-group [op, ntasks, stream, level, flops, ms, variant] of the
        last profiled run (variant: GEMM tile configuration, id*1000000 +
        BM*1000 + BN; -1 for other groups)."""
        cnt = C.c_int(0)
        call("dm_dag_profile", self.h, None, C.byref(cnt))
        buf = (_lib.dm_group_profile * max(cnt.value, 1))()
        cnt2 = C.c_int(cnt.value)
        call("dm_dag_profile", self.h, buf, C.byref(cnt2))
        return [dict(op=b.op, ntasks=b.ntasks, stream=b.stream, level=b.level, flops=b.flops,
                     ms=b.ms, variant=b.variant) for b in buf[:cnt2.value]]

    def edges(self):
        cnt = C.c_int(0)
        call("dm_dag_edges", self.h, None, C.byref(cnt))
        buf = (C.c_int * (2 * max(cnt.value, 1)))()
        cnt2 = C.c_int(cnt.value)
        call("dm_dag_edges", self.h, buf, C.byref(cnt2))
        return [(buf[2 * i], buf[2 * i + 1]) for i in range(cnt2.value)]

    def plan(self, opts):
        o = opts.c()
        call("dm_dag_plan", self.h, C.byref(o))
        self.planned = True
        return self

    def run(self, stream=None, sync=True):
        call("dm_dag_run", self.h, stream)
        if sync:
            call("dm_stream_sync", stream)

    def stats(self):
        s = _lib.dm_dag_stats()
        call("dm_dag_get_stats", self.h, C.byref(s))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def __del__(self):
        try:
            if self.h:
                _lib.load().dm_dag_destroy(self.h)
        except Exception:
            pass


class _Region:
    def __init__(self):
        self.tasks = []


_tls = threading.local()


def _stack():
    if not hasattr(_tls, "stack"):
        _tls.stack = []
    return _tls.stack


def spawn(f, *args):
    """`Dagger.@spawn f(args...)`.  Inside a datadeps region the task is
    recorded; outside one it runs immediately as a one-task region."""
    t = make_task(f, args)
    st = _stack()
    if st:
        st[-1].tasks.append(t)
        return None
    d = Dag([t]).plan(SEQUENTIAL)
    d.run()
    return None


def spawn_datadeps(f=None, *, options=None, stream=None):
    """Run `f()` as a datadeps region on the device.

    Returns f's result after every recorded task has completed.  As a
    context manager (`with spawn_datadeps() as region:`) the region runs on
    exit.  `options` is a PlanOptions (default: 2 streams, batching,
    look-ahead 1, no graph)."""
    if f is None:
        return _RegionCtx(options, stream)
    with _RegionCtx(options, stream):
        return f()


class _RegionCtx:
    def __init__(self, options, stream):
        self.options = options or PlanOptions()
        self.stream = stream
        self.region = _Region()
        self.dag = None

    def __enter__(self):
        _stack().append(self.region)
        return self

    def __exit__(self, et, ev, tb):
        _stack().pop()
        if et is not None:
            return False
        if self.region.tasks:
            self.dag = Dag(self.region.tasks).plan(self.options)
            self.dag.run(self.stream)
        return False
