"""cholesky! / cholesky / ldiv! / trsm! / trsv! on resident DArrays.

Host mirror of Dagger's tiled linear algebra for this path:
  * `cholesky_(A)`       LinearAlgebra.cholesky!(::DArray) -> _chol!
                         (src/array/cholesky.jl:19-108)
  * `cholesky(A)`        out of place via cholcopy = copy (src/array/cholesky.jl:1)
  * `ldiv_(C, B)`        ldiv!(::Cholesky{T,<:DMatrix}, B) (src/array/linalg.jl:352-405)
  * `ldiv(C, B)`         C \\ B staged in the factor's blocking (src/array/linalg.jl:208-246)
  * `ldiv_triangular_`   ldiv!(UpperTriangular(A), B) etc. (src/array/linalg.jl:310-335)
  * `trsm_`, `trsv_`     Dagger.trsm!/trsv! (src/array/trsm.jl:1-192)
  * `mul_`, `mul`        mul!(C, op(A), op(B), alpha, beta) and op(A) * op(B)
                         (src/array/mul.jl:11-210)
Block re-partitioning follows maybe_copy_buffered (src/array/copy.jl:3-10):
arguments whose partitioning differs from the requested square Blocks are
copied into a buffer of that partitioning, processed, and copied back.

Two executions are available:
  * via="native": the C++ builders (dm_cholesky_tiled / dm_trsm_tiled)
    record the reference's loop nest natively and cache the plan + hipGraph
    per matrix (the steady-state fast path); ldiv! (dm_potrs_tiled) runs the
    same two triangular sweeps as GEMM-only substitution over the inverted
    diagonal tiles (the arithmetic reordered, equal to 1e-12 relative);
  * via="datadeps": the loop nest is written here with spawn/In/InOut and
    handed to the executor, exactly as the Julia source spawns it.
Cholesky and trsm! give bit-identical results both ways (same tasks, same
kernels).
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import call
from .darray import DArray, Blocks, DeviceBuffer, DeviceInt, zeros
from .datadeps import (spawn, spawn_datadeps, In, Out, InOut, BLAS, potrf_checked,
                       PlanOptions)


class PosDefException(Exception):
    """LinearAlgebra.PosDefException(info)."""

    def __init__(self, info):
        super().__init__(f"matrix is not positive definite; factorization failed (info={info})")
        self.info = info


class Cholesky:
    """`Cholesky{T,<:DMatrix}`: factors + uplo + info."""

    def __init__(self, factors, uplo, info, global_info=0):
        self.factors, self.uplo, self.info = factors, uplo, int(info)
        self.global_info = int(global_info)

    @property
    def U(self):
        F = self.factors.collect()
        return np.triu(F) if self.uplo == "U" else np.triu(F.conj().T)

    @property
    def L(self):
        F = self.factors.collect()
        return np.tril(F) if self.uplo == "L" else np.tril(F.conj().T)

    def issuccess(self):
        return self.info == 0


DEFAULT_OPTIONS = PlanOptions(num_streams=2, batch=True, lookahead=1, use_graph=True)


def _square_blocks(A):
    bs = min(A.partitioning.blocksize)
    return Blocks(bs, bs)


def _maybe_copy_buffered(pairs, f):
    """maybe_copy_buffered(A => Blocks(...), ...) (src/array/copy.jl:3-34)."""
    if all(a.partitioning == p for a, p in pairs):
        return f(*[a for a, _ in pairs])
    bufs = [a.reblock(p) if a.partitioning != p else a for a, p in pairs]
    res = f(*bufs)
    for (a, p), b in zip(pairs, bufs):
        if b is not a:
            b.copy_into(a)
            b.unsafe_free()
    return res


def _chol_datadeps(Ac, uplo, info, options):
    """The loop nest of src/array/cholesky.jl:38-53 (U) / :83-98 (L); for
    complex tiles trans = 'C' and herk! in place of syrk! (:27-28, :44-45)."""
    mt, nt = Ac.shape
    iscomplex = np.issubdtype(Ac[0, 0].dtype, np.complexfloating)
    trans = "C" if iscomplex else "T"
    rk = BLAS.herk if iscomplex else BLAS.syrk
    with spawn_datadeps(options=options):
        if uplo == "U":
            for k in range(mt):
                spawn(potrf_checked, "U", InOut(Ac[k, k]), Out(info))
                for n in range(k + 1, nt):
                    spawn(BLAS.trsm, "L", "U", trans, "N", 1.0, In(Ac[k, k]), InOut(Ac[k, n]))
                for m in range(k + 1, mt):
                    spawn(rk, "U", trans, -1.0, In(Ac[k, m]), 1.0, InOut(Ac[m, m]))
                    for n in range(m + 1, nt):
                        spawn(BLAS.gemm, trans, "N", -1.0, In(Ac[k, m]), In(Ac[k, n]), 1.0,
                              InOut(Ac[m, n]))
        else:
            for k in range(mt):
                spawn(potrf_checked, "L", InOut(Ac[k, k]), Out(info))
                for m in range(k + 1, mt):
                    spawn(BLAS.trsm, "R", "L", trans, "N", 1.0, In(Ac[k, k]), InOut(Ac[m, k]))
                for n in range(k + 1, nt):
                    spawn(rk, "L", "N", -1.0, In(Ac[n, k]), 1.0, InOut(Ac[n, n]))
                    for m in range(n + 1, mt):
                        spawn(BLAS.gemm, "N", trans, -1.0, In(Ac[m, k]), In(Ac[n, k]), 1.0,
                              InOut(Ac[m, n]))


def _chol(A, uplo, via, options):
    """LinearAlgebra._chol!(A, UpperTriangular/LowerTriangular): returns info."""
    if A.ndim != 2 or A.m != A.n:
        raise ValueError("DimensionMismatch: matrix is not square")
    # One info pair per matrix: the cached plan/graph is keyed on these
    # device pointers, so reusing them keeps repeated factorizations on the
    # captured graph.
    if getattr(A, "_chol_info", None) is None:
        A._chol_info = (DeviceInt(1), DeviceInt(1))
    info, ginfo = A._chol_info

    def run(Ab):
        if via == "native":
            t = Ab.tiled()
            o = options.c()
            call("dm_cholesky_tiled", uplo.encode(), C.byref(t), C.c_void_p(info.ptr),
                 C.c_void_p(ginfo.ptr), C.byref(o), None)
            call("dm_stream_sync", None)
        elif via == "datadeps":
            _chol_datadeps(Ab.chunks, uplo, info, options)
        else:
            raise ValueError(f"unknown via={via!r}")

    _maybe_copy_buffered([(A, _square_blocks(A))], run)
    loc = int(info.get()[0])
    g = int(ginfo.get()[0]) if via == "native" else 0
    return loc, g


def ishermitian(A):
    """`ishermitian(::DArray)` (src/array/linalg.jl:90-124) on the device:
    exact symmetry of diagonal tiles, sqrt(eps)-relative Frobenius test of
    off-diagonal tile pairs; non-square blocks are re-blocked first."""
    if A.ndim != 2 or A.m != A.n:
        return False
    sq = _square_blocks(A)
    B = A if A.partitioning == sq else A.reblock(sq)
    mt = B.mt
    work = DeviceInt(max(mt * (mt + 1) // 2, 1))
    res = DeviceInt(1)
    t = B.tiled()
    call("dm_ishermitian_tiled", C.byref(t), C.c_void_p(res.ptr), C.c_void_p(work.ptr), None)
    ok = bool(res.get()[0])
    if B is not A:
        B.unsafe_free()
    return ok


_PARTS = {"full": 0, "upper": 1, "lower": 2}


def _norm2_pass(A, part, stream=None):
    """One dm_tiled_norm2 pass: (sqrt of the sum of squares, any non-finite)."""
    if A.ndim == 1:
        A = DArray(A.dtype, (A.m, 1), Blocks(A.mb, 1), _buffer=A.buffer)
    work = DeviceBuffer(8 * 8 * max(A.mt * A.nt, 1))
    out = DeviceBuffer(8)
    flag = DeviceInt(1)
    t = A.tiled()
    call("dm_tiled_norm2", C.byref(t), _PARTS[part], C.c_void_p(out.ptr), C.c_void_p(flag.ptr),
         C.c_void_p(work.ptr), stream)
    val = np.zeros(1)
    call("dm_memcpy_d2h", val.ctypes.data_as(C.c_void_p), C.c_void_p(out.ptr), C.c_size_t(8),
         stream)
    bad = bool(flag.get(stream)[0])
    for b in (work, out, flag):
        b.free()
    return float(val[0]), bad


def norm2(A, part="full"):
    """`LinearAlgebra.norm2(::DArray)` (src/array/linalg.jl:1-5) and of its
    UpperTriangular / LowerTriangular views (part "upper" / "lower",
    src/array/linalg.jl:7-33): sqrt of the sum of squares over the stored
    tiles, one device pass (FP64 accumulation)."""
    if part not in _PARTS:
        raise ValueError("part must be 'full', 'upper' or 'lower'")
    if part != "full" and (A.ndim != 2 or A.partitioning != _square_blocks(A)):
        raise ValueError("triangular norm2 needs a matrix with square blocks")
    return _norm2_pass(A, part)[0]


def chkfinite(A):
    """`LAPACK.chkfinite(::DArray)` (src/array/linalg.jl:127-148): True when
    every element is finite, else raises ValueError("matrix has Inf or NaN")
    (the reference's ArgumentError), from one device pass."""
    if _norm2_pass(A, "full")[1]:
        raise ValueError("matrix has Inf or NaN")
    return True


def cholesky_(A, uplo="U", check=True, via="native", options=None, check_hermitian=True):
    """`cholesky!(A)` / `cholesky!(Hermitian(A, :L))` on a DArray, in place.

    As the stdlib cholesky!(::AbstractMatrix) does, a non-Hermitian input is
    rejected with info = -1 before factoring (check_hermitian=False skips the
    pass, as cholesky!(Hermitian(A)) would).  Otherwise info is the
    reference's: the LAPACK info of the first diagonal tile whose POTRF fails
    (tile-local, src/array/cholesky.jl:13-16,62); the global leading-minor
    order is kept in `Cholesky.global_info`."""
    uplo = uplo.upper()
    if uplo not in ("U", "L"):
        raise ValueError("uplo must be 'U' or 'L'")
    options = options or DEFAULT_OPTIONS
    if check_hermitian and not ishermitian(A):
        if check:
            raise PosDefException(-1)
        return Cholesky(A, uplo, -1, -1)
    info, ginfo = _chol(A, uplo, via, options)
    if check and info != 0:
        raise PosDefException(info)
    return Cholesky(A, uplo, info, ginfo)


def cholesky(A, uplo="U", check=True, via="native", options=None):
    """Out-of-place `cholesky(A)`: cholcopy (a copy) then cholesky!."""
    return cholesky_(A.copy(), uplo=uplo, check=check, via=via, options=options)


def trsm_(side, uplo, trans, diag, alpha, A, B, via="native", options=None):
    """Dagger.trsm!(side, uplo, trans, diag, alpha, A, B) (src/array/trsm.jl:59-192)."""
    options = options or DEFAULT_OPTIONS
    if via == "native":
        ta, tb = A.tiled(), B.tiled()
        o = options.c()
        call("dm_trsm_tiled", side.encode(), uplo.encode(), trans.encode(), diag.encode(),
             float(alpha), C.byref(ta), C.byref(tb), C.byref(o), None)
        call("dm_stream_sync", None)
        return B
    _trsm_datadeps(side, uplo, trans, diag, alpha, A.chunks, B.chunks, options)
    return B


def _trsm_datadeps_right(uplo, trans, diag, alpha, Ac, Bc):
    """src/array/trsm.jl:143-189 (side 'R'), inside an open region.  The
    transposed branches pass alpha to every diagonal solve and 1 to every
    update, as the reference does."""
    Bmt, Bnt = Bc.shape
    if uplo == "U" and trans == "N":
        for k in range(Bnt):
            la = alpha if k == 0 else 1.0
            for m in range(Bmt):
                spawn(BLAS.trsm, "R", "U", "N", diag, la, In(Ac[k, k]), InOut(Bc[m, k]))
            for m in range(Bmt):
                for n in range(k + 1, Bnt):
                    spawn(BLAS.gemm, "N", "N", -1.0, In(Bc[m, k]), In(Ac[k, n]), la,
                          InOut(Bc[m, n]))
    elif uplo == "U":
        for k in range(1, Bnt + 1):
            kk = Bnt - k
            for m in range(Bmt):
                spawn(BLAS.trsm, "R", "U", trans, diag, alpha, In(Ac[kk, kk]), InOut(Bc[m, kk]))
                for n in range(k + 1, Bnt + 1):
                    spawn(BLAS.gemm, "N", trans, -1.0, In(Bc[m, kk]), In(Ac[Bnt - n, kk]), 1.0,
                          InOut(Bc[m, Bnt - n]))
    elif trans == "N":
        for k in range(1, Bnt + 1):
            la = alpha if k == 1 else 1.0
            kk = Bnt - k
            for m in range(Bmt):
                spawn(BLAS.trsm, "R", "L", "N", diag, la, In(Ac[kk, kk]), InOut(Bc[m, kk]))
                for n in range(k + 1, Bnt + 1):
                    spawn(BLAS.gemm, "N", "N", -1.0, In(Bc[m, kk]), In(Ac[kk, Bnt - n]), la,
                          InOut(Bc[m, Bnt - n]))
    else:
        for k in range(Bnt):
            for m in range(Bmt):
                spawn(BLAS.trsm, "R", "L", trans, diag, alpha, In(Ac[k, k]), InOut(Bc[m, k]))
                for n in range(k + 1, Bnt):
                    spawn(BLAS.gemm, "N", trans, -1.0, In(Bc[m, k]), In(Ac[n, k]), 1.0,
                          InOut(Bc[m, n]))


def _trsm_datadeps(side, uplo, trans, diag, alpha, Ac, Bc, options):
    """src/array/trsm.jl:88-190, every side/uplo/trans branch."""
    Bmt, Bnt = Bc.shape
    if side == "R":
        with spawn_datadeps(options=options):
            _trsm_datadeps_right(uplo, trans, diag, alpha, Ac, Bc)
        return
    with spawn_datadeps(options=options):
        if uplo == "U" and trans == "N":
            for k in range(1, Bmt + 1):
                la = alpha if k == 1 else 1.0
                kk = Bmt - k
                for n in range(Bnt):
                    spawn(BLAS.trsm, side, uplo, trans, diag, la, In(Ac[kk, kk]), InOut(Bc[kk, n]))
                for m in range(k + 1, Bmt + 1):
                    mm = Bmt - m
                    for n in range(Bnt):
                        spawn(BLAS.gemm, "N", "N", -1.0, In(Ac[mm, kk]), In(Bc[kk, n]), la,
                              InOut(Bc[mm, n]))
        elif uplo == "U":
            for k in range(Bmt):
                la = alpha if k == 0 else 1.0
                for n in range(Bnt):
                    spawn(BLAS.trsm, side, uplo, trans, diag, la, In(Ac[k, k]), InOut(Bc[k, n]))
                for m in range(k + 1, Bmt):
                    for n in range(Bnt):
                        spawn(BLAS.gemm, trans, "N", -1.0, In(Ac[k, m]), In(Bc[k, n]), la,
                              InOut(Bc[m, n]))
        elif trans == "N":
            for k in range(Bmt):
                la = alpha if k == 0 else 1.0
                for n in range(Bnt):
                    spawn(BLAS.trsm, side, uplo, trans, diag, la, In(Ac[k, k]), InOut(Bc[k, n]))
                for m in range(k + 1, Bmt):
                    for n in range(Bnt):
                        spawn(BLAS.gemm, "N", "N", -1.0, In(Ac[m, k]), In(Bc[k, n]), la,
                              InOut(Bc[m, n]))
        else:
            for k in range(1, Bmt + 1):
                la = alpha if k == 1 else 1.0
                kk = Bmt - k
                for n in range(Bnt):
                    spawn(BLAS.trsm, side, uplo, trans, diag, la, In(Ac[kk, kk]), InOut(Bc[kk, n]))
                for m in range(k + 1, Bmt + 1):
                    mm = Bmt - m
                    for n in range(Bnt):
                        spawn(BLAS.gemm, trans, "N", -1.0, In(Ac[kk, mm]), In(Bc[kk, n]), la,
                              InOut(Bc[mm, n]))


def trsv_(uplo, trans, diag, alpha, A, b, options=None):
    """Dagger.trsv! (src/array/trsm.jl:1-57): trsv/gemv tile tasks.  As in
    the reference, alpha is the GEMV beta of the first step's updates only
    (`lalpha`); the first diagonal block solved is not scaled."""
    options = options or PlanOptions(num_streams=1, batch=False, lookahead=0)
    Ac, Bc = A.chunks, b.chunks
    Bnt = len(Bc)
    with spawn_datadeps(options=options):
        if uplo == "U" and trans == "N":
            for k in reversed(range(Bnt)):
                la = alpha if k == Bnt - 1 else 1.0
                spawn(BLAS.trsv, "U", "N", diag, In(Ac[k, k]), InOut(Bc[k]))
                for i in range(k):
                    spawn(BLAS.gemv, "N", -1.0, In(Ac[i, k]), In(Bc[k]), la, InOut(Bc[i]))
        elif uplo == "U":
            for k in range(Bnt):
                la = alpha if k == 0 else 1.0
                spawn(BLAS.trsv, "U", trans, diag, In(Ac[k, k]), InOut(Bc[k]))
                for i in range(k + 1, Bnt):
                    spawn(BLAS.gemv, trans, -1.0, In(Ac[k, i]), In(Bc[k]), la, InOut(Bc[i]))
        elif trans == "N":
            for k in range(Bnt):
                la = alpha if k == 0 else 1.0
                spawn(BLAS.trsv, "L", "N", diag, In(Ac[k, k]), InOut(Bc[k]))
                for i in range(k + 1, Bnt):
                    spawn(BLAS.gemv, "N", -1.0, In(Ac[i, k]), In(Bc[k]), la, InOut(Bc[i]))
        else:
            for k in reversed(range(Bnt)):
                la = alpha if k == Bnt - 1 else 1.0
                spawn(BLAS.trsv, "L", trans, diag, In(Ac[k, k]), InOut(Bc[k]))
                for i in range(k):
                    spawn(BLAS.gemv, trans, -1.0, In(Ac[k, i]), In(Bc[k]), la, InOut(Bc[i]))
    return b


def ldiv_(F, B, via="native", options=None):
    """`ldiv!(C::Cholesky, B)` (src/array/linalg.jl:352-405), B overwritten.
    An LU factorization dispatches to `ldiv!(::LU, B)` (lu.py)."""
    if not isinstance(F, Cholesky):
        from .lu import LU as _LU, ldiv_ as _lu_ldiv_
        if isinstance(F, _LU):
            return _lu_ldiv_(F, B, options=options)
        raise TypeError(f"ldiv! of a {type(F).__name__} is not supported")
    options = options or DEFAULT_OPTIONS
    pA = F.factors
    bs = min(pA.partitioning.blocksize)
    sq = Blocks(bs, bs)
    iscomplex = np.issubdtype(pA.dtype, np.complexfloating)
    tc = "C" if iscomplex else "T"
    if B.ndim == 1 and iscomplex:
        # a complex vector solves as the n x 1 matrix over the same storage
        # (the vector kernels are real); its Blocks(bs) tiles are the column's
        if B.partitioning != Blocks(bs):
            Bv = B.reblock(Blocks(bs))
            ldiv_(F, Bv, via=via, options=options)
            Bv.copy_into(B)
            Bv.unsafe_free()
            return B
        ldiv_(F, DArray(B.dtype, (B.m, 1), sq, _buffer=B.buffer), via=via, options=options)
        return B
    if B.ndim == 1:
        partB = Blocks(bs)

        def vec(pa, pb):
            if F.uplo == "U":
                trsv_("U", "T", "N", 1.0, pa, pb)
                trsv_("U", "N", "N", 1.0, pa, pb)
            else:
                trsv_("L", "N", "N", 1.0, pa, pb)
                trsv_("L", "T", "N", 1.0, pa, pb)

        _maybe_copy_buffered([(pA, sq), (B, partB)], vec)
        return B
    partB = sq

    def mat(pa, pb):
        if via == "native":
            ta, tb = pa.tiled(), pb.tiled()
            o = options.c()
            call("dm_potrs_tiled", F.uplo.encode(), C.byref(ta), C.byref(tb), C.byref(o), None)
            call("dm_stream_sync", None)
        elif F.uplo == "U":
            _trsm_datadeps("L", "U", tc, "N", 1.0, pa.chunks, pb.chunks, options)
            _trsm_datadeps("L", "U", "N", "N", 1.0, pa.chunks, pb.chunks, options)
        else:
            _trsm_datadeps("L", "L", "N", "N", 1.0, pa.chunks, pb.chunks, options)
            _trsm_datadeps("L", "L", tc, "N", 1.0, pa.chunks, pb.chunks, options)

    _maybe_copy_buffered([(pA, sq), (B, partB)], mat)
    return B


def ldiv(F, B, via="native", options=None):
    """`F \\ B` for `F::Cholesky{T,<:DMatrix}`: a new DArray X with A X = B;
    B (a host array or a DArray) is left as it was.

    The stdlib `ldiv(F, B)` stages B into a plain array whose automatic
    blocking almost never matches the factor's, so every solve would re-block
    the whole factor (src/array/linalg.jl:208-220). As the reference's LU
    override does (src/array/linalg.jl:221-246), B is staged once into a
    DArray with the factor's square blocking, so `ldiv!`
    (src/array/linalg.jl:352-405) runs without any copy of the factor.
    An LU factorization dispatches to its own `\\` (lu.py)."""
    if not isinstance(F, Cholesky):
        from .lu import LU as _LU, ldiv as _lu_ldiv
        if isinstance(F, _LU):
            return _lu_ldiv(F, B, options=options)
        raise TypeError(f"\\ of a {type(F).__name__} is not supported")
    A = F.factors
    if A.m != A.n:
        raise ValueError("the factor must be square")
    if not isinstance(B, DArray):
        B = np.asarray(B)
    B_shape = B.shape
    if len(B_shape) not in (1, 2):
        raise ValueError("B must be a vector or a matrix")
    if B_shape[0] != A.m:
        raise ValueError("arguments must have the same number of rows")
    if np.result_type(B.dtype, A.dtype) != A.dtype:
        raise NotImplementedError(
            f"promoting a {A.dtype} factor to the right-hand side's {B.dtype} is not supported")
    bs = min(A.partitioning.blocksize)
    part = Blocks(bs) if len(B_shape) == 1 else Blocks(bs, bs)
    if isinstance(B, DArray):
        if B.dtype != A.dtype:
            raise NotImplementedError("a DArray right-hand side must have the factor's type")
        X = B.copy() if B.partitioning == part else B.reblock(part)
    else:
        X = DArray.from_array(np.asarray(B, dtype=A.dtype), part)
    call("dm_stream_sync", None)
    return ldiv_(F, X, via=via, options=options)


def ldiv_triangular_(uplo, diag, A, B, options=None):
    """`ldiv!(T(A), B)` for T in {Upper,UnitUpper,Lower,UnitLower}Triangular
    of a DMatrix (src/array/linalg.jl:310-335): B overwritten by T(A) \\ B,
    through `trsv!` for a vector and `trsm!('L', uplo, 'N', diag, 1, ...)` for
    a matrix, both on A and B re-blocked to A's smallest block size."""
    if uplo not in ("U", "L"):
        raise ValueError("uplo must be 'U' or 'L'")
    if diag not in ("N", "U"):
        raise ValueError("diag must be 'N' or 'U'")
    if A.m != A.n or B.m != A.m:
        raise ValueError("A must be square with as many rows as B")
    bsa = min(A.partitioning.blocksize)
    if B.ndim == 1:
        bs = min(bsa, B.partitioning.blocksize[0])
        _maybe_copy_buffered([(A, Blocks(bs, bs)), (B, Blocks(bs))],
                             lambda pa, pb: trsv_(uplo, "N", diag, 1.0, pa, pb, options))
        return B
    _maybe_copy_buffered([(A, Blocks(bsa, bsa)), (B, Blocks(bsa, bsa))],
                         lambda pa, pb: trsm_("L", uplo, "N", diag, 1.0, pa, pb, options=options))
    return B


def mul_(Cm, A, B, alpha=1.0, beta=0.0, transA="N", transB="N", options=None):
    """`mul!(C, op(A), op(B), alpha, beta)` on DMatrices: Dagger's
    generic_matmatmul! (src/array/mul.jl:11-43) and gemm_dagger!'s tile loop
    (src/array/mul.jl:107-210: for each C tile, k-ordered tile products, beta
    applied by the first), planned and run by the native executor.

    Operands are re-blocked (maybe_copy_buffered, src/array/copy.jl:3-34) to
    square blocks of the smallest block dimension of A and B, the size
    _repartition_matmatmul falls back to (src/array/mul.jl:44-93); C is
    copied back into its own blocking afterwards. For A is B with one
    operand transposed the reference takes syrk_dagger! (upper tiles, then
    copytri!); here the full product is formed by the same tile products,
    which equals it for symmetric C or beta = 0."""
    ta = "T" if transA in ("T", "C") else transA
    tb = "T" if transB in ("T", "C") else transB
    if ta not in ("N", "T") or tb not in ("N", "T"):
        raise ValueError("trans must be 'N', 'T' or 'C'")
    if A.ndim == 2 and B.ndim == 1 and Cm.ndim == 1:
        return _mul_vec(Cm, A, B, alpha, beta, ta, options)
    if A.ndim != 2 or B.ndim != 2 or Cm.ndim != 2:
        raise ValueError("mul_ takes three matrices, or C and B vectors")
    if not (Cm.dtype == A.dtype == B.dtype):
        raise ValueError("A, B and C must share one element type")
    am, ak = (A.n, A.m) if ta == "T" else (A.m, A.n)
    bk, bn = (B.n, B.m) if tb == "T" else (B.m, B.n)
    if ak != bk or am != Cm.m or bn != Cm.n:
        raise ValueError("shapes do not conform")
    sz = min(*A.partitioning.blocksize, *B.partitioning.blocksize)
    sq = Blocks(sz, sz)
    options = options or DEFAULT_OPTIONS

    def run(c, a, b):
        tc, tx, ty = c.tiled(), a.tiled(), b.tiled()
        o = options.c()
        call("dm_gemm_tiled", ta.encode(), tb.encode(), float(alpha), C.byref(tx), C.byref(ty),
             float(beta), C.byref(tc), C.byref(o), None)
        call("dm_stream_sync", None)

    if A is B:
        _maybe_copy_buffered([(Cm, sq), (A, sq)], lambda c, a: run(c, a, a))
    else:
        _maybe_copy_buffered([(Cm, sq), (A, sq), (B, sq)], run)
    return Cm


def _mul_vec(y, A, x, alpha, beta, ta, options):
    """`mul!(y, op(A), x, alpha, beta)` for DVectors x and y: Dagger's
    generic_matvecmul! (src/array/mul.jl:437-449) with its repartitioning
    (_repartition_matvecmul, :450-472: x and A's inner blocking fall back to
    the smallest block dimension when they differ, y takes op(A)'s row
    blocking) and gemv_dagger!'s loop (:474-541: for each y block, the
    k-ordered tile GEMVs, beta applied by the first), run by the native
    executor's GEMV tile kernel (FP64 or FP32)."""
    if not (y.dtype == A.dtype == x.dtype):
        raise ValueError("y, A and x must share one element type")
    am, ak = (A.n, A.m) if ta == "T" else (A.m, A.n)
    if ak != x.m or am != y.m:
        raise ValueError("shapes do not conform")
    pa = A.partitioning.blocksize
    pb = x.partitioning.blocksize[0]
    dim_a, dim_inner = (pa[1], pa[0]) if ta == "T" else (pa[0], pa[1])
    part_a, part_x = pa, pb
    if dim_inner != pb:
        sz = min(pa[0], pa[1], pb)
        part_a = (sz, pa[1]) if ta == "T" else (pa[0], sz)
        part_x = sz
    options = options or PlanOptions(num_streams=1, batch=False, lookahead=0)

    def run(c, a, b):
        Ac, Bc, Cc = a.chunks, b.chunks, c.chunks
        K = a.mt if ta == "T" else a.nt
        with spawn_datadeps(options=options):
            for m in range(len(Cc)):
                for k in range(K):
                    tile = Ac[k, m] if ta == "T" else Ac[m, k]
                    spawn(BLAS.gemv, ta, float(alpha), In(tile), In(Bc[k]),
                          float(beta) if k == 0 else 1.0, InOut(Cc[m]))

    _maybe_copy_buffered([(y, Blocks(dim_a)), (A, Blocks(*part_a)), (x, Blocks(part_x))], run)
    return y


def mul(A, B, transA="N", transB="N", options=None):
    """`op(A) * op(B)` on DMatrices (`A * B`, `A' * B`, ... of
    src/array/mul.jl): a new C in HBM, square blocks of the smallest block
    dimension of A and B, filled by `mul_` with alpha = 1, beta = 0."""
    m = A.n if transA in ("T", "C") else A.m
    if B.ndim == 1:
        pa = A.partitioning.blocksize
        y = zeros(Blocks(pa[1] if transA in ("T", "C") else pa[0]), A.dtype, m)
        return _mul_vec(y, A, B, 1.0, 0.0, "T" if transA in ("T", "C") else "N", options)
    n = B.m if transB in ("T", "C") else B.n
    sz = min(*A.partitioning.blocksize, *B.partitioning.blocksize)
    Cm = zeros(Blocks(sz, sz), A.dtype, m, n)
    return mul_(Cm, A, B, 1.0, 0.0, transA, transB, options)
