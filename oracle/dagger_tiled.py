"""CPU oracle: a restatement of Dagger.jl's tiled Cholesky / ldiv! path.

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may import this module, and only as the
checker (or the timed CPU baseline), never as the product path.

What it restates (paths inside JuliaParallel/Dagger.jl):
  * partition / _cumlength              src/array/alloc.jl:23-33
  * each DArray chunk is its own dense column-major array
                                        src/array/darray.jl:568-581
  * _chol!(A, UpperTriangular)          src/array/cholesky.jl:19-63 (loop :38-53)
  * _chol!(A, LowerTriangular)          src/array/cholesky.jl:64-108 (loop :83-98)
  * potrf_checked! info rule            src/array/cholesky.jl:2-18
  * trsm! (all side/uplo/trans branches) src/array/trsm.jl:59-192
  * trsv! (vector right-hand side)      src/array/trsm.jl:1-57
  * ldiv!(::Cholesky{T,<:DMatrix}, B)   src/array/linalg.jl:352-405
  * datadeps dependency rules           src/datadeps/queue.jl:44-54,
                                        src/datadeps/aliasing.jl:740-773
  * norm2 of a DArray and of its UpperTriangular / LowerTriangular views
                                        src/array/linalg.jl:1-33
  * LAPACK.chkfinite(::DArray)          src/array/linalg.jl:127-148
  * ishermitian(::DArray) with is_cross_hermitian
                                        src/array/linalg.jl:38-48, 90-124
  * lu!(A, RowMaximum()) / lu!(A, NoPivot()) and ldiv!(::LU, B)
                                        src/array/lu.jl:5-55, 160-204,
                                        239-262, 289-371;
                                        src/array/linalg.jl:248-310
The tile kernels are the same LAPACK/BLAS routines the reference calls on
its CPU path (LAPACK.potrf!, BLAS.trsm!/syrk!/gemm!/gemv!/trsv!), here
through scipy's bundled OpenBLAS.

Execution: `run_sequential` runs the recorded tasks in submission order
(the datadeps semantics: the result equals sequential execution);
`run_threaded` runs the same task list on a thread pool with one
single-threaded BLAS call per worker, honoring the datadeps dependencies
-- the shape of Dagger's CPU datadeps execution (ThreadProc workers,
src/threadproc.jl:13-41), used as the timed CPU baseline.
"""
import concurrent.futures as cf
import os
import threading

import numpy as np
import scipy.linalg.blas as blas
import scipy.linalg.lapack as lapack

IN, OUT, INOUT = 1, 2, 3


class PosDefException(Exception):
    def __init__(self, info):
        super().__init__(f"PosDefException({info})")
        self.info = info


def cumlength(length, step):
    """src/array/alloc.jl:23-28."""
    nice, extra = divmod(length, step)
    return [step] * nice + ([extra] if extra else [])


class TiledMatrix:
    """A Blocks(mb, nb) partitioned matrix whose chunks are separate
    Fortran-ordered arrays (Dagger's DArray chunks)."""

    def __init__(self, A, mb, nb=None):
        A = np.asarray(A)
        if A.ndim == 1:
            A = A.reshape(-1, 1)
            nb = 1
        nb = mb if nb is None else nb
        self.m, self.n = A.shape
        self.mb, self.nb = mb, nb
        self.rows = cumlength(self.m, mb)
        self.cols = cumlength(self.n, nb)
        self.chunks = np.empty((len(self.rows), len(self.cols)), dtype=object)
        r0 = 0
        for i, r in enumerate(self.rows):
            c0 = 0
            for j, c in enumerate(self.cols):
                self.chunks[i, j] = np.array(A[r0:r0 + r, c0:c0 + c], order="F", copy=True)
                c0 += c
            r0 += r

    @property
    def mt(self):
        return len(self.rows)

    @property
    def nt(self):
        return len(self.cols)

    def collect(self):
        return np.block([[self.chunks[i, j] for j in range(self.nt)] for i in range(self.mt)])


# ---------------------------------------------------------------- tile ops
# Each op mutates its output tile in place, like BLAS.xxx!/LAPACK.potrf!.

def _tr(t):
    """BLAS trans code: 'N' 0, 'T' 1, 'C' 2 (the conjugate transpose; for
    real element types 'C' is 'T')."""
    return {"N": 0, "T": 1, "C": 2}[t]


_PREFIX = {np.dtype(np.float64): "d", np.dtype(np.float32): "s",
           np.dtype(np.complex128): "z", np.dtype(np.complex64): "c"}


def _blas(name, dtype):
    return getattr(blas, _PREFIX[np.dtype(dtype)] + name)


def tile_potrf(uplo, A):
    """LAPACK.potrf!(uplo, A) -> info (LAPACK convention; complex: the
    Hermitian A = U^H U / L L^H)."""
    f = getattr(lapack, _PREFIX[A.dtype] + "potrf")
    c, info = f(A, lower=1 if uplo == "L" else 0, clean=0, overwrite_a=0)
    A[...] = c
    return int(info)


def tile_trsm(side, uplo, trans, diag, alpha, A, B):
    B[...] = _blas("trsm", B.dtype)(alpha, A, B, side=0 if side == "L" else 1,
                                     lower=1 if uplo == "L" else 0, trans_a=_tr(trans),
                                     diag=1 if diag == "U" else 0)


def tile_syrk(uplo, trans, alpha, A, beta, C):
    C[...] = _blas("syrk", C.dtype)(alpha, A, beta=beta, c=C, trans=_tr(trans),
                                    lower=1 if uplo == "L" else 0)


def tile_herk(uplo, trans, alpha, A, beta, C):
    """BLAS.herk!(uplo, trans, alpha, A, beta, C) on complex tiles: trans 'N'
    C = a A A^H + b C, 'C' C = a A^H A + b C on the uplo triangle, with the
    imaginary part of the diagonal set to zero."""
    C[...] = _blas("herk", C.dtype)(alpha, A, beta=beta, c=C, trans=_tr(trans),
                                    lower=1 if uplo == "L" else 0)


def tile_gemm(ta, tb, alpha, A, B, beta, C):
    C[...] = _blas("gemm", C.dtype)(alpha, A, B, beta=beta, c=C, trans_a=_tr(ta),
                                    trans_b=_tr(tb))


def tile_gemv(trans, alpha, A, x, beta, y):
    f = blas.dgemv if y.dtype == np.float64 else blas.sgemv
    y[...] = f(alpha, A, x.ravel(), beta=beta, y=y.ravel(), trans=_tr(trans)).reshape(y.shape)


def tile_trsv(uplo, trans, diag, A, x):
    f = blas.dtrsv if x.dtype == np.float64 else blas.strsv
    x[...] = f(A, x.ravel(), lower=1 if uplo == "L" else 0, trans=_tr(trans),
               diag=1 if diag == "U" else 0).reshape(x.shape)


# ---------------------------------------------------------------- tasks
class Task:
    """One recorded tile task: fn(*args) with (array, mode) accesses."""

    __slots__ = ("name", "fn", "args", "acc")

    def __init__(self, name, fn, args, acc):
        self.name, self.fn, self.args, self.acc = name, fn, args, acc

    def run(self):
        return self.fn(*self.args)


class Region:
    """A datadeps region: tasks recorded in submission order."""

    def __init__(self):
        self.tasks = []

    def spawn(self, name, fn, args, acc):
        self.tasks.append(Task(name, fn, args, acc))


def datadeps_edges(tasks):
    """Dependencies by the datadeps rules: In waits for the last writer of
    the same array; Out/InOut waits for the last writer and every reader
    since (src/datadeps/aliasing.jl:740-773).  Arrays are whole tiles here
    (ContiguousAliasing of a chunk), identified by object identity."""
    last_writer, readers = {}, {}
    edges = set()
    for i, t in enumerate(tasks):
        deps = set()
        for arr, mode in t.acc:
            key = id(arr)
            if key in last_writer:
                deps.add(last_writer[key])
            if mode & OUT:
                deps.update(readers.get(key, ()))
        deps.discard(i)
        for d in deps:
            edges.add((d, i))
        for arr, mode in t.acc:
            key = id(arr)
            if mode & OUT:
                last_writer[key] = i
                readers[key] = []
            else:
                readers.setdefault(key, []).append(i)
    return edges


def run_sequential(region):
    for t in region.tasks:
        t.run()


def run_threaded(region, workers=None):
    """Execute the region on a thread pool honoring the datadeps edges, one
    tile BLAS call per worker (BLAS itself single-threaded via threadpoolctl
    when available)."""
    tasks = region.tasks
    n = len(tasks)
    workers = workers or os.cpu_count() or 1
    preds = [0] * n
    succs = [[] for _ in range(n)]
    for a, b in datadeps_edges(tasks):
        preds[b] += 1
        succs[a].append(b)
    lock = threading.Lock()
    done = threading.Event()
    remaining = [n]
    errors = []
    ex = cf.ThreadPoolExecutor(max_workers=workers)

    def submit(i):
        ex.submit(work, i)

    def work(i):
        try:
            tasks[i].run()
        except Exception as e:  # surfaces after the region, like wait_all
            errors.append(e)
        ready = []
        with lock:
            for s in succs[i]:
                preds[s] -= 1
                if preds[s] == 0:
                    ready.append(s)
            remaining[0] -= 1
            if remaining[0] == 0:
                done.set()
        for s in ready:
            submit(s)

    if n == 0:
        return
    # snapshot the sources first: once submitted, completing tasks release
    # their successors themselves (a live scan could submit a task twice)
    sources = [i for i in range(n) if preds[i] == 0]
    for i in sources:
        submit(i)
    done.wait()
    ex.shutdown(wait=True)
    if errors:
        raise errors[0]


def blas_single_thread():
    """Context limiting BLAS to one thread per call (threadpoolctl)."""
    try:
        from threadpoolctl import threadpool_limits
        return threadpool_limits(limits=1, user_api="blas")
    except Exception:  # pragma: no cover
        import contextlib
        return contextlib.nullcontext()


# ---------------------------------------------------------------- Cholesky
def chol_region(T, uplo, info, steps=None):
    """Record the task loop of src/array/cholesky.jl:38-53 ('U') / :83-98 ('L').

    info is the 1-element array filled by potrf_checked! on failure; `steps`
    (default all) records only the first `steps` iterations of the k loop
    (a bounded sample of the same task list, for timing)."""
    Ac = T.chunks
    mt, nt = Ac.shape
    kmax = mt if steps is None else min(mt, steps)
    R = Region()

    def potrf_checked(uplo_, A, info_arr):
        # src/array/cholesky.jl:2-18: tile-local LAPACK info; throws
        if info_arr[0] != 0:
            return  # a dependent of the failed task does not run
        i = tile_potrf(uplo_, A)
        if i != 0:
            info_arr[0] = i
            raise PosDefException(i)

    def guarded(fn):
        def g(*a):
            if info[0] != 0:
                return  # tasks after a failed POTRF fail with it
            fn(*a)
        return g

    iscomplex = np.issubdtype(Ac[0, 0].dtype, np.complexfloating)
    tr = "C" if iscomplex else "T"  # src/array/cholesky.jl:27-28
    trsm, gemm = guarded(tile_trsm), guarded(tile_gemm)
    syrk = guarded(tile_herk if iscomplex else tile_syrk)  # herk! for complex (:44-48)
    if uplo == "U":
        for k in range(kmax):
            R.spawn("potrf", potrf_checked, ("U", Ac[k, k], info), [(Ac[k, k], INOUT), (info, OUT)])
            for n in range(k + 1, nt):
                R.spawn("trsm", trsm, ("L", "U", tr, "N", 1.0, Ac[k, k], Ac[k, n]),
                        [(Ac[k, k], IN), (Ac[k, n], INOUT)])
            for m in range(k + 1, mt):
                R.spawn("syrk", syrk, ("U", tr, -1.0, Ac[k, m], 1.0, Ac[m, m]),
                        [(Ac[k, m], IN), (Ac[m, m], INOUT)])
                for n in range(m + 1, nt):
                    R.spawn("gemm", gemm, (tr, "N", -1.0, Ac[k, m], Ac[k, n], 1.0, Ac[m, n]),
                            [(Ac[k, m], IN), (Ac[k, n], IN), (Ac[m, n], INOUT)])
    else:
        for k in range(kmax):
            R.spawn("potrf", potrf_checked, ("L", Ac[k, k], info), [(Ac[k, k], INOUT), (info, OUT)])
            for m in range(k + 1, mt):
                R.spawn("trsm", trsm, ("R", "L", tr, "N", 1.0, Ac[k, k], Ac[m, k]),
                        [(Ac[k, k], IN), (Ac[m, k], INOUT)])
            for n in range(k + 1, nt):
                R.spawn("syrk", syrk, ("L", "N", -1.0, Ac[n, k], 1.0, Ac[n, n]),
                        [(Ac[n, k], IN), (Ac[n, n], INOUT)])
                for m in range(n + 1, mt):
                    R.spawn("gemm", gemm, ("N", tr, -1.0, Ac[m, k], Ac[n, k], 1.0, Ac[m, n]),
                            [(Ac[m, k], IN), (Ac[n, k], IN), (Ac[m, n], INOUT)])
    return R


def cholesky(A, nb, uplo="U", threaded=False, workers=None, steps=None):
    """_chol!(DArray(A, Blocks(nb, nb)), uplo): returns (factor matrix, info).

    The factor matrix is the collected DArray: the uplo triangle holds the
    factor, the other triangle keeps the input (as the reference leaves it).
    `steps` limits the k loop (chol_region) to a leading sample."""
    T = TiledMatrix(np.array(A, copy=True), nb, nb)
    info = np.zeros(1, dtype=np.int64)
    R = chol_region(T, uplo, info, steps)
    try:
        if threaded:
            run_threaded(R, workers)
        else:
            for t in R.tasks:
                try:
                    t.run()
                except PosDefException:
                    pass
    except PosDefException:
        pass
    return T.collect(), int(info[0])


# ---------------------------------------------------------------- trsm!
def trsm_region(side, uplo, trans, diag, alpha, TA, TB):
    """Record the task loop of src/array/trsm.jl:88-190."""
    Ac, Bc = TA.chunks, TB.chunks
    Bmt, Bnt = Bc.shape
    R = Region()
    # the reference hands trans through to its tile calls ('C' is the
    # conjugate transpose on complex tiles, the same as 'T' on real ones)
    tr = trans if np.iscomplexobj(TB.chunks[0, 0]) else ("T" if trans in ("T", "C") else "N")

    def trsm(al, a, b):
        R.spawn("trsm", tile_trsm, (side, uplo, tr, diag, al, a, b), [(a, IN), (b, INOUT)])

    def gemm(ta, tb, x, y, beta, c):
        R.spawn("gemm", tile_gemm, (ta, tb, -1.0, x, y, beta, c), [(x, IN), (y, IN), (c, INOUT)])

    if side == "L":
        if uplo == "U":
            if tr == "N":
                for k in range(1, Bmt + 1):
                    la = alpha if k == 1 else 1.0
                    kk = Bmt - k
                    for n in range(Bnt):
                        trsm(la, Ac[kk, kk], Bc[kk, n])
                    for m in range(k + 1, Bmt + 1):
                        for n in range(Bnt):
                            gemm("N", "N", Ac[Bmt - m, kk], Bc[kk, n], la, Bc[Bmt - m, n])
            else:
                for k in range(Bmt):
                    la = alpha if k == 0 else 1.0
                    for n in range(Bnt):
                        trsm(la, Ac[k, k], Bc[k, n])
                    for m in range(k + 1, Bmt):
                        for n in range(Bnt):
                            gemm(tr, "N", Ac[k, m], Bc[k, n], la, Bc[m, n])
        else:
            if tr == "N":
                for k in range(Bmt):
                    la = alpha if k == 0 else 1.0
                    for n in range(Bnt):
                        trsm(la, Ac[k, k], Bc[k, n])
                    for m in range(k + 1, Bmt):
                        for n in range(Bnt):
                            gemm("N", "N", Ac[m, k], Bc[k, n], la, Bc[m, n])
            else:
                for k in range(1, Bmt + 1):
                    la = alpha if k == 1 else 1.0
                    kk = Bmt - k
                    for n in range(Bnt):
                        trsm(la, Ac[kk, kk], Bc[kk, n])
                    for m in range(k + 1, Bmt + 1):
                        for n in range(Bnt):
                            gemm(tr, "N", Ac[kk, Bmt - m], Bc[kk, n], la, Bc[Bmt - m, n])
    else:
        if uplo == "U":
            if tr == "N":
                for k in range(Bnt):
                    la = alpha if k == 0 else 1.0
                    for m in range(Bmt):
                        trsm(la, Ac[k, k], Bc[m, k])
                    for m in range(Bmt):
                        for n in range(k + 1, Bnt):
                            gemm("N", "N", Bc[m, k], Ac[k, n], la, Bc[m, n])
            else:
                for k in range(1, Bnt + 1):
                    kk = Bnt - k
                    for m in range(Bmt):
                        trsm(alpha, Ac[kk, kk], Bc[m, kk])
                        for n in range(k + 1, Bnt + 1):
                            gemm("N", tr, Bc[m, kk], Ac[Bnt - n, kk], 1.0, Bc[m, Bnt - n])
        else:
            if tr == "N":
                for k in range(1, Bnt + 1):
                    la = alpha if k == 1 else 1.0
                    kk = Bnt - k
                    for m in range(Bmt):
                        trsm(la, Ac[kk, kk], Bc[m, kk])
                        for n in range(k + 1, Bnt + 1):
                            gemm("N", "N", Bc[m, kk], Ac[kk, Bnt - n], la, Bc[m, Bnt - n])
            else:
                for k in range(Bnt):
                    for m in range(Bmt):
                        trsm(alpha, Ac[k, k], Bc[m, k])
                        for n in range(k + 1, Bnt):
                            gemm("N", tr, Bc[m, k], Ac[n, k], 1.0, Bc[m, n])
    return R


def trsm(side, uplo, trans, diag, alpha, A, B, nb):
    TA = TiledMatrix(np.array(A, copy=True), nb, nb)
    TB = TiledMatrix(np.array(B, copy=True), nb, nb)
    run_sequential(trsm_region(side, uplo, trans, diag, alpha, TA, TB))
    return TB.collect()


def trsv_region(uplo, trans, diag, TA, TB, alpha=1.0):
    """src/array/trsm.jl:16-54.  alpha enters only as the GEMV beta
    (`lalpha`) of the first step's updates, as in the reference: the first
    diagonal block solved is never scaled, so for one block alpha has no
    effect and for more the result is not alpha * op(A)^-1 b."""
    Ac = TA.chunks
    Bc = [TB.chunks[i, 0] for i in range(TB.mt)]
    Bnt = len(Bc)
    R = Region()

    def trsv(u, t, a, x):
        R.spawn("trsv", tile_trsv, (u, t, diag, a, x), [(a, IN), (x, INOUT)])

    def gemv(t, a, x, y, beta):
        R.spawn("gemv", tile_gemv, (t, -1.0, a, x, beta, y), [(a, IN), (x, IN), (y, INOUT)])

    if uplo == "U":
        if trans == "N":
            for k in reversed(range(Bnt)):
                la = alpha if k == Bnt - 1 else 1.0
                trsv("U", "N", Ac[k, k], Bc[k])
                for i in range(k):
                    gemv("N", Ac[i, k], Bc[k], Bc[i], la)
        else:
            for k in range(Bnt):
                la = alpha if k == 0 else 1.0
                trsv("U", trans, Ac[k, k], Bc[k])
                for i in range(k + 1, Bnt):
                    gemv(trans, Ac[k, i], Bc[k], Bc[i], la)
    else:
        if trans == "N":
            for k in range(Bnt):
                la = alpha if k == 0 else 1.0
                trsv("L", "N", Ac[k, k], Bc[k])
                for i in range(k + 1, Bnt):
                    gemv("N", Ac[i, k], Bc[k], Bc[i], la)
        else:
            for k in reversed(range(Bnt)):
                la = alpha if k == Bnt - 1 else 1.0
                trsv("L", trans, Ac[k, k], Bc[k])
                for i in range(k):
                    gemv(trans, Ac[k, i], Bc[k], Bc[i], la)
    return R


def gemm_region(ta, tb, alpha, TA, TB, beta, TC):
    """gemm_dagger!'s tile loop (src/array/mul.jl:107-210): for each C tile,
    the k-ordered tile products, beta applied by the first."""
    Ac, Bc, Cc = TA.chunks, TB.chunks, TC.chunks
    K = TA.mt if ta == "T" else TA.nt
    R = Region()
    for j in range(TC.nt):
        for i in range(TC.mt):
            for k in range(K):
                a = Ac[k, i] if ta == "T" else Ac[i, k]
                b = Bc[j, k] if tb == "T" else Bc[k, j]
                R.spawn("gemm", tile_gemm, (ta, tb, alpha, a, b, beta if k == 0 else 1.0, Cc[i, j]),
                        [(a, IN), (b, IN), (Cc[i, j], INOUT)])
    return R


def gemm(ta, tb, alpha, A, B, beta, Cm, nb):
    """mul!(C, op(A), op(B), alpha, beta) on matrices blocked nb x nb
    (src/array/mul.jl:11-43 then gemm_dagger!); returns C."""
    TA = TiledMatrix(np.array(A, copy=True), nb, nb)
    TB = TiledMatrix(np.array(B, copy=True), nb, nb)
    TC = TiledMatrix(np.array(Cm, copy=True), nb, nb)
    run_sequential(gemm_region(ta, tb, alpha, TA, TB, beta, TC))
    return TC.collect()


def trsv(uplo, trans, diag, A, b, nb, alpha=1.0):
    """Dagger.trsv!(uplo, trans, diag, alpha, A, b) (src/array/trsm.jl:1-57)
    on A tiled nb x nb and b tiled nb; returns x."""
    TA = TiledMatrix(np.array(A, copy=True), nb, nb)
    TB = TiledMatrix(np.array(b, copy=True), nb)
    run_sequential(trsv_region(uplo, trans, diag, TA, TB, alpha))
    return TB.collect().ravel()


def ldiv(F, uplo, B, nb, threaded=False, workers=None):
    """ldiv!(Cholesky(F, uplo), B) (src/array/linalg.jl:352-405): the two
    triangular-solve regions over the tiled factor; returns X."""
    B = np.asarray(B)
    TA = TiledMatrix(np.array(F, copy=True), nb, nb)
    run = (lambda R: run_threaded(R, workers)) if threaded else run_sequential
    if np.iscomplexobj(F) and B.ndim == 1:
        return ldiv(F, uplo, B.reshape(-1, 1), nb, threaded, workers).ravel()  # n x 1 form
    if B.ndim == 1:
        TB = TiledMatrix(np.array(B, copy=True), nb)
        if uplo == "U":
            run(trsv_region("U", "T", "N", TA, TB))
            run(trsv_region("U", "N", "N", TA, TB))
        else:
            run(trsv_region("L", "N", "N", TA, TB))
            run(trsv_region("L", "T", "N", TA, TB))
        return TB.collect().ravel()
    tc = "C" if np.iscomplexobj(F) else "T"  # Uᴴ / Lᴴ for complex factors
    TB = TiledMatrix(np.array(B, copy=True), nb, nb)
    if uplo == "U":
        run(trsm_region("L", "U", tc, "N", 1.0, TA, TB))
        run(trsm_region("L", "U", "N", "N", 1.0, TA, TB))
    else:
        run(trsm_region("L", "L", "N", "N", 1.0, TA, TB))
        run(trsm_region("L", "L", tc, "N", 1.0, TA, TB))
    return TB.collect()


# ---------------------------------------------------------------- metrics
def rel_factor_diff(F1, F2, uplo="U"):
    """Frobenius relative difference of the uplo triangles."""
    tri = np.triu if uplo == "U" else np.tril
    a, b = tri(F1), tri(F2)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def reconstruction_error(F, A, uplo="U"):
    """||A - U^T U||_F / ||A||_F (or L L^T)."""
    if uplo == "U":
        U = np.triu(F)
        R = U.T @ U
    else:
        L = np.tril(F)
        R = L @ L.T
    return float(np.linalg.norm(A - R) / np.linalg.norm(A))


def backward_error_columns(A, X, B):
    """Largest per-column normwise backward error
    max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf + ||b_j||_inf) -- the
    per-right-hand-side criterion of LAPACK's iterative refinement (dsposv)."""
    X2 = X.reshape(X.shape[0], -1)
    B2 = B.reshape(B.shape[0], -1)
    return max(backward_error(A, X2[:, j], B2[:, j]) for j in range(X2.shape[1]))


def backward_error(A, X, B):
    """Normwise ||B - A X||_inf / (||A||_inf ||X||_inf + ||B||_inf)."""
    X2 = X.reshape(X.shape[0], -1)
    B2 = B.reshape(B.shape[0], -1)
    R = B2 - A @ X2
    return float(np.abs(R).sum(axis=1).max() /
                 (np.abs(A).sum(axis=1).max() * np.abs(X2).sum(axis=1).max() +
                  np.abs(B2).sum(axis=1).max()))


# ---------------------------------------------------------------- norms
def norm2(A, nb, part="full"):
    """norm2 (src/array/linalg.jl:1-33): per chunk mapreduce(norm_sqr, +),
    then the sum over chunks, then sqrt.  part "upper" takes the chunks
    right of the diagonal chunk plus UpperTriangular of each diagonal chunk
    (:7-19); "lower" the mirror image (:21-33).  Square blocks."""
    T = TiledMatrix(A, nb)
    sq = lambda X: float(np.sum(np.abs(X.astype(np.float64)) ** 2))
    if part == "full":
        return float(np.sqrt(sum(sq(T.chunks[i, j]) for i in range(T.mt) for j in range(T.nt))))
    off, diag = 0.0, 0.0
    for i in range(T.mt):
        if part == "upper":
            off += sum(sq(T.chunks[i, j]) for j in range(i + 1, T.nt))
            diag += sq(np.triu(T.chunks[i, i]))
        else:
            off += sum(sq(T.chunks[j, i]) for j in range(i + 1, T.mt))
            diag += sq(np.tril(T.chunks[i, i]))
    return float(np.sqrt(off + diag))


def chkfinite(A, nb):
    """LAPACK.chkfinite(::DArray) (src/array/linalg.jl:127-148): every
    chunk all-finite, else ArgumentError (here ValueError)."""
    T = TiledMatrix(A, nb)
    for i in range(T.mt):
        for j in range(T.nt):
            if not np.all(np.isfinite(T.chunks[i, j])):
                raise ValueError("matrix has Inf or NaN")
    return True


def ishermitian(A, nb):
    """ishermitian(::DArray) (src/array/linalg.jl:90-124) on A tiled
    Blocks(nb, nb): every diagonal chunk exactly symmetric (stdlib
    ishermitian: element equality, so a NaN fails), and for every chunk pair
    below the diagonal is_cross_hermitian (:38-48):
        ||A_ij - A_ji^T||_F <= sqrt(eps) * max(||A_ij||_F, ||A_ji^T||_F),
    eps of the element type, each Frobenius norm sqrt(sum(abs2, .)).
    A non-square matrix is not Hermitian."""
    A = np.asarray(A)
    if A.ndim != 2 or A.shape[0] != A.shape[1]:
        return False
    rtol = np.sqrt(np.finfo(A.dtype).eps)
    T = TiledMatrix(A, nb, nb)
    mt = T.mt
    for i in range(mt):
        d = T.chunks[i, i]
        if not np.array_equal(d, d.T):
            return False
    for i in range(1, mt):
        for j in range(i):
            a1 = T.chunks[i, j].astype(np.float64)
            b = T.chunks[j, i].T.astype(np.float64)
            lhs = np.sqrt(np.sum((a1 - b) ** 2))
            rhs = rtol * max(np.sqrt(np.sum(a1 ** 2)), np.sqrt(np.sum(b ** 2)))
            if not lhs <= rhs:
                return False
    return True


# ---------------------------------------------------------------- LU
def lu_panel(panel_tiles, k, mb, ipiv_chunk, info):
    """factor_panel_lapack! (src/array/lu.jl:160-204): stack the panel's
    tiles, LAPACK getrf, write the tiles back; ipiv_chunk gets the global
    1-based pivot rows, info[0] the global first zero pivot if unset."""
    stacked = np.vstack(panel_tiles)
    f = lapack.dgetrf if stacked.dtype == np.float64 else lapack.sgetrf
    lu, piv, linfo = f(stacked, overwrite_a=0)
    base = k * mb
    if info[0] == 0 and linfo > 0:
        info[0] = base + int(linfo)
    ipiv_chunk[: len(piv)] = base + piv + 1       # scipy: 0-based
    r0 = 0
    for t in panel_tiles:
        t[...] = lu[r0:r0 + t.shape[0]]
        r0 += t.shape[0]


def swaprows_trail(top, blocks, ipiv_chunk, k, mb):
    """swaprows_trail! over tiles (k..mt-1) of one block column
    (src/array/lu.jl:239-262): pivot p of panel k swaps row p of the top
    tile with row ipiv[p] (global, 1-based) of the block column, in order
    p = 1..npiv (the reference spawns one task per target tile in order,
    which gives the same sequence)."""
    col = np.vstack(blocks)
    for p, g in enumerate(ipiv_chunk):
        q = int(g) - 1 - k * mb
        if q != p:
            col[[p, q]] = col[[q, p]]
    r0 = 0
    for t in blocks:
        t[...] = col[r0:r0 + t.shape[0]]
        r0 += t.shape[0]


def lu_nopivot_tile(A, info):
    """generic_lufact!(A, NoPivot()) on one tile (Doolittle, no row
    exchanges): unit lower L below the diagonal, U on and above; a zero
    pivot sets info (1-based, local) and its column is not scaled."""
    n = min(A.shape)
    for j in range(n):
        d = A[j, j]
        if d == 0:
            if info[0] == 0:
                info[0] = j + 1
            continue
        A[j + 1:, j] /= d
        A[j + 1:, j + 1:] -= np.outer(A[j + 1:, j], A[j, j + 1:])


def lu(A, nb, pivot=True):
    """lu!(A, RowMaximum()) (src/array/lu.jl:289-371) or lu!(A, NoPivot())
    (:10-40) on A tiled Blocks(nb, nb); returns (factors, ipiv (global
    1-based), info)."""
    T = TiledMatrix(np.array(A, copy=True), nb, nb)
    Ac = T.chunks
    mt, nt = T.mt, T.nt
    m, n = T.m, T.n
    ipiv = np.arange(1, min(m, n) + 1)
    info = [0]
    for k in range(min(mt, nt)):
        if pivot:
            chunk = ipiv[k * nb:k * nb + min(m - k * nb, T.cols[k])]
            lu_panel([Ac[i, k] for i in range(k, mt)], k, nb, chunk, info)
            for j in list(range(k)) + list(range(k + 1, nt)):
                swaprows_trail(Ac[k, j], [Ac[i, j] for i in range(k, mt)], chunk, k, nb)
        else:
            li = [0]
            lu_nopivot_tile(Ac[k, k], li)
            if info[0] == 0 and li[0]:
                info[0] = k * nb + li[0]
            for i in range(k + 1, mt):
                tile_trsm("R", "U", "N", "N", 1.0, Ac[k, k][:T.cols[k], :T.cols[k]], Ac[i, k])
        for j in range(k + 1, nt):
            r = T.rows[k]
            tile_trsm("L", "L", "N", "U", 1.0, Ac[k, k][:r, :r], Ac[k, j])
            for i in range(k + 1, mt):
                tile_gemm("N", "N", -1.0, Ac[i, k], Ac[k, j], 1.0, Ac[i, j])
    return T.collect(), ipiv, info[0]


def lu_ldiv(F, ipiv, B, nb):
    """ldiv!(::LU, B) (src/array/linalg.jl:248-310): the pivots applied to
    B's rows in order, then the unit lower and the upper tiled solves."""
    X = np.array(B, dtype=F.dtype, copy=True)
    vec = X.ndim == 1
    X = X.reshape(X.shape[0], -1)
    for p, g in enumerate(ipiv):
        q = int(g) - 1
        if q != p:
            X[[p, q]] = X[[q, p]]
    X = trsm("L", "L", "N", "U", 1.0, F, X, nb)
    X = trsm("L", "U", "N", "N", 1.0, F, X, nb)
    return X.ravel() if vec else X


def lu_unpack(F, ipiv):
    """(L, U, P) with P A = L U from the packed factors and ipiv."""
    m, n = F.shape
    r = min(m, n)
    L = np.tril(F[:, :r], -1) + np.eye(m, r)
    U = np.triu(F[:r, :])
    perm = np.arange(m)
    for p, g in enumerate(ipiv):
        q = int(g) - 1
        perm[[p, q]] = perm[[q, p]]
    P = np.eye(m)[perm]
    return L, U, P
