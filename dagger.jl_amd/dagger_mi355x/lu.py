"""lu / lu! / ldiv! for LU factorizations of resident DArrays.

Host mirror of Dagger's tiled LU (src/array/lu.jl):
  * `lu_(A, RowMaximum())`  lu!(A, RowMaximum()) (:289-371): per block
    column the panel with partial pivoting, the row swaps of the other block
    columns, trsm!('L','L','N','U') across the row and the GEMM update --
    the task DAG recorded natively (dm_getrf_tiled) and run by the executor;
  * `lu_(A, NoPivot())`     lu!(A, NoPivot()) (:10-40);
  * `lu(A, pivot)`          the out-of-place form (a copy first, :1-9);
  * `ldiv_(F, B)`           ldiv!(::LU, B) (src/array/linalg.jl:248-310):
    the pivots applied to B's rows (dm_tiled_laswp), then the unit lower and
    the upper tiled triangular solves;
  * `ldiv(F, B)`            F \\ B with B staged in the factor's blocking
    (src/array/linalg.jl:208-246).
Operands with non-square blocks are re-blocked to the smaller block size
(maybe_copy_buffered), as the reference does.
"""
import ctypes as C

import numpy as np

from ._lib import call
from .darray import DArray, Blocks, DeviceBuffer, DeviceInt
from .linalg import (DEFAULT_OPTIONS, _maybe_copy_buffered, chkfinite, ldiv_triangular_)


class RowMaximum:
    """LinearAlgebra.RowMaximum(): partial pivoting by maximum magnitude."""


class NoPivot:
    """LinearAlgebra.NoPivot()."""


class SingularException(Exception):
    """LinearAlgebra.SingularException(info)."""

    def __init__(self, info):
        super().__init__(f"matrix is singular (info={info})")
        self.info = info


class LU:
    """`LU{T,DMatrix{T},DVector{Int}}`: packed factors (unit lower L below
    the diagonal, U on and above), ipiv (global 1-based pivot rows, numpy)
    and info; `L`, `U`, `P` are host arrays with P A = L U."""

    def __init__(self, factors, ipiv, info, ipiv_dev):
        self.factors, self.ipiv, self.info = factors, ipiv, int(info)
        self._ipiv_dev = ipiv_dev

    def _parts(self):
        F = self.factors.collect()
        m, n = F.shape
        r = min(m, n)
        L = np.tril(F[:, :r], -1) + np.eye(m, r, dtype=F.dtype)
        U = np.triu(F[:r, :])
        perm = np.arange(m)
        for p, g in enumerate(self.ipiv):
            q = int(g) - 1
            perm[[p, q]] = perm[[q, p]]
        return L, U, np.eye(m, dtype=F.dtype)[perm]

    @property
    def L(self):
        return self._parts()[0]

    @property
    def U(self):
        return self._parts()[1]

    @property
    def P(self):
        return self._parts()[2]

    def issuccess(self):
        return self.info == 0


def lu_(A, pivot=None, check=True, allowsingular=False, options=None):
    """`lu!(A, pivot)` on a DMatrix, in place (default RowMaximum()).

    As the reference, check=True first rejects non-finite input
    (LAPACK.chkfinite) and afterwards raises SingularException(info) for an
    exactly zero pivot unless allowsingular; the factorization itself always
    completes (LAPACK getrf semantics)."""
    pivot = pivot if pivot is not None else RowMaximum()
    if not isinstance(pivot, (RowMaximum, NoPivot)):
        raise TypeError("pivot must be RowMaximum() or NoPivot()")
    if A.ndim != 2:
        raise ValueError("lu! needs a matrix")
    if check:
        chkfinite(A)
    options = options or DEFAULT_OPTIONS
    bs = min(A.partitioning.blocksize)
    r = min(A.m, A.n)
    ipiv_dev = DeviceInt(max(r, 1))
    info = DeviceInt(1)

    def run(Ab):
        t = Ab.tiled()
        o = options.c()
        call("dm_getrf_tiled", 1 if isinstance(pivot, RowMaximum) else 0, C.byref(t),
             C.c_void_p(ipiv_dev.ptr), C.c_void_p(info.ptr), C.byref(o), None)
        call("dm_stream_sync", None)

    _maybe_copy_buffered([(A, Blocks(bs, bs))], run)
    ipiv = ipiv_dev.get()[:r].astype(np.int64)
    inf = int(info.get()[0])
    info.free()
    if check and inf > 0 and not allowsingular:
        ipiv_dev.free()
        raise SingularException(inf)
    return LU(A, ipiv, inf, ipiv_dev)


def lu(A, pivot=None, check=True, allowsingular=False, options=None):
    """Out-of-place `lu(A, pivot)`: a copy of A factored by `lu_`."""
    return lu_(A.copy(), pivot=pivot, check=check, allowsingular=allowsingular, options=options)


def _apply_ipiv_rows(F, B):
    """_apply_ipiv_rows! (src/array/linalg.jl:266-310): B's rows permuted by
    F's pivots in order, on the device."""
    Bm = B if B.ndim == 2 else DArray(B.dtype, (B.m, 1), Blocks(B.mb, 1), _buffer=B.buffer)
    t = Bm.tiled()
    call("dm_tiled_laswp", C.byref(t), C.c_void_p(F._ipiv_dev.ptr), len(F.ipiv), None)


def ldiv_(F, B, options=None):
    """`ldiv!(F::LU, B)` (src/array/linalg.jl:248-252): pivots, then
    UnitLowerTriangular and UpperTriangular solves of the packed factors;
    B (a DVector or DMatrix) is overwritten with the solution."""
    A = F.factors
    if A.m != A.n:
        raise ValueError("ldiv! needs a square factorization")
    if B.m != A.m:
        raise ValueError("DimensionMismatch: arguments must have the same number of rows")
    bs = min(A.partitioning.blocksize)
    part = Blocks(bs) if B.ndim == 1 else Blocks(bs, B.partitioning.blocksize[1])
    _maybe_copy_buffered([(B, part)], lambda b: _apply_ipiv_rows(F, b))
    call("dm_stream_sync", None)
    ldiv_triangular_("L", "U", A, B, options=options)
    ldiv_triangular_("U", "N", A, B, options=options)
    return B


def ldiv(F, B, options=None):
    """`F \\ B` for `F::LU{T,<:DMatrix}` (src/array/linalg.jl:221-246): B (a
    host array or a DArray) staged once into a DArray with the factor's
    blocking, solved by `ldiv_`; B is left unchanged."""
    A = F.factors
    if A.m != A.n:
        raise ValueError("non-square systems are not supported")
    bs = A.partitioning.blocksize[0]
    if isinstance(B, DArray):
        part = Blocks(bs) if B.ndim == 1 else Blocks(bs, bs)
        X = B.copy() if B.partitioning == part else B.reblock(part)
    else:
        Bh = np.asarray(B, dtype=A.dtype)
        if Bh.shape[0] != A.m:
            raise ValueError("DimensionMismatch: arguments must have the same number of rows")
        X = DArray.from_array(Bh, Blocks(bs) if Bh.ndim == 1 else Blocks(bs, bs))
    call("dm_stream_sync", None)
    return ldiv_(F, X, options=options)
