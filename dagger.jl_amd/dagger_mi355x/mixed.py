"""Precision form: low-precision tiled Cholesky + FP64 iterative refinement.

Beside the plain FP64 drop-in.  The factorization runs the same tile DAG
(src/array/cholesky.jl:38-53) on an FP32 copy of A: TRSM and POTRF in FP32,
the SYRK/GEMM trailing updates on the BF16 matrix cores with 3-term operand
splitting (FP32-accurate); the residual r = b - A x, its norms and the update x += d stay in
FP64 on the untouched FP64 A (LAPACK dsposv's structure, on resident tiles):

    x0 = F32^-1 b
    repeat: r = b - A x                       (FP64 tiled GEMM, HBM-bound)
            berr = max_j ||r_j|| / (||A|| ||x_j|| + ||b_j||)   (inf-norms)
            stop if berr <= tol
            d = F32^-1 r ; x += d              (FP32 solve, FP64 update)

If refinement stalls (kappa(A) * u32 near 1), `solve` switches to GMRES-IR:
FP64 GMRES on A with the FP32 factor as a left preconditioner (Carson &
Higham, "Accelerating the solution of linear systems by iterative
refinement in three precisions", SISC 2018), one right-hand side at a time.

Every device call takes an explicit stream, so independent solves can
overlap (the serving loop runs one per worker stream).
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import call
from .darray import DArray, Blocks, DeviceBuffer, DeviceInt, Tile, _Temp
from .datadeps import PlanOptions, Dag, make_task, BLAS, In, Out, InOut
from .linalg import PosDefException

# Trailing updates of the FP32 factor on the BF16 matrix cores with 3-term
# operand splitting (FP32-accurate, several times the FP32 MFMA rate).
_OPTS = PlanOptions(num_streams=2, batch=True, lookahead=1, use_graph=True, f32_update=6,
                    f64_diag=True)


def _square(t):
    """The leading rows x rows part of a tile (ld unchanged)."""
    return Tile(t.ptr, t.rows, t.rows, t.ld, t.dtype, owner=None, index=None)


class _DiagInverseSolve:
    """F32^-1 R for the FP32 factor F = U (U^T U = A32), in place, as two
    blocked substitutions whose diagonal steps are products with the
    inverted diagonal tiles U_kk^-1: every launch is a GEMM, batched per step
    by the executor, so a solve with few right-hand sides is a short chain of
    bandwidth-bound launches instead of 2 * mt dependent leaf solves.
    (The refinement loop's own solver; the plain drop-in's ldiv! keeps
    Dagger's trsm! loop nest.)  Rounding differs from substitution by the
    inverse's conditioning, which refinement absorbs like the FP32 factor's
    own error."""

    def __init__(self, F, options):
        self.F, self.options = F, options
        nb = F.mb
        self.Dinv = DArray(np.float32, (F.m, nb), Blocks(nb, nb))
        tasks = [make_task(BLAS.trsm, ("L", "U", "N", "N", 1.0, In(F.tile(k, k)),
                                        InOut(_square(self.Dinv.tile(k, 0)))))
                 for k in range(F.mt)]
        self._inv = Dag(tasks).plan(options.replace(use_graph=False))
        self._solves = {}

    def invert(self, stream):
        """Dinv_k := U_kk^-1 for every diagonal tile (after each refactor)."""
        t = self.Dinv.tiled()
        call("dm_tiled_identity", C.byref(t), stream)
        self._inv.run(stream, sync=False)

    def _dag(self, R, Y):
        key = (R.buffer.ptr, Y.buffer.ptr)
        d = self._solves.get(key)
        if d is None:
            F, mt = self.F, self.F.mt
            D = [_square(self.Dinv.tile(k, 0)) for k in range(mt)]
            ts = []
            for k in range(mt):  # U^T Y = R: Y_k = U_kk^-T R_k, R_j -= U_kj^T Y_k
                ts.append(make_task(BLAS.gemm, ("T", "N", 1.0, In(D[k]), In(R.tile(k, 0)), 0.0,
                                                Out(Y.tile(k, 0)))))
                for j in range(k + 1, mt):
                    ts.append(make_task(BLAS.gemm, ("T", "N", -1.0, In(F.tile(k, j)),
                                                    In(Y.tile(k, 0)), 1.0, InOut(R.tile(j, 0)))))
            for k in reversed(range(mt)):  # U X = Y: X_k = U_kk^-1 Y_k, Y_i -= U_ik X_k
                ts.append(make_task(BLAS.gemm, ("N", "N", 1.0, In(D[k]), In(Y.tile(k, 0)), 0.0,
                                                Out(R.tile(k, 0)))))
                for i in range(k):
                    ts.append(make_task(BLAS.gemm, ("N", "N", -1.0, In(F.tile(i, k)),
                                                    In(R.tile(k, 0)), 1.0, InOut(Y.tile(i, 0)))))
            d = Dag(ts).plan(self.options)
            self._solves[key] = d
        return d

    def solve(self, R, Y, stream):
        """R := F32^-1 R on `stream` (Y: scratch of R's shape)."""
        self._dag(R, Y).run(stream, sync=False)


def _convert(dst, src, stream=None):
    td, ts = dst.tiled(), src.tiled()
    call("dm_tiled_convert", C.byref(td), C.byref(ts), stream)


def _colnorms(X, buf, stream=None):
    t = X.tiled()
    call("dm_tiled_col_norm_inf", C.byref(t), C.c_void_p(buf.ptr), stream)
    out = np.empty(X.n)
    call("dm_memcpy_d2h", out.ctypes.data_as(C.c_void_p), C.c_void_p(buf.ptr),
         C.c_size_t(8 * X.n), stream)
    return out


def release_factor_buffers(F):
    """Free the refinement buffers a MixedCholesky attached to its FP32
    factor array F (info flags, norm scratch, per-right-hand-side buffers,
    inverted diagonal tiles).  They reference F, so without this they are
    reclaimed only by Python's cyclic collector, whose hipFree calls then
    wait for the whole device at an arbitrary later time."""
    aux = getattr(F, "_mixed_aux", None)
    if aux is None:
        return
    for key in ("info", "ginfo", "work", "norm"):
        aux[key].free()
    for b in aux["bufs"].values():
        for name in ("X", "R", "D", "Y"):
            b[name].unsafe_free()
        b["nrm"].free()
    iso = aux.get("isolve")
    if iso is not None:
        iso.Dinv.unsafe_free()
        iso._solves.clear()
        iso._inv = None
    aux.clear()
    F._mixed_aux = None


class MixedCholesky:
    """cholesky(A) in FP32 tiles, kept with the FP64 A for refinement.

    A64 must stay resident and unchanged while the factor is used.  `F`
    may be passed to reuse a preallocated FP32 DArray of A's shape."""

    def __init__(self, A64, options=None, stream=None, F=None):
        if A64.dtype != np.float64:
            raise TypeError("A must be an FP64 DArray")
        self.A = A64
        self.options = options or _OPTS
        self.stream = stream
        bs = min(A64.partitioning.blocksize)
        if A64.partitioning != Blocks(bs, bs):
            raise ValueError("mixed Cholesky needs square Blocks(nb, nb)")
        self.nb = bs
        self.F = F if F is not None else DArray(np.float32, A64.shape, A64.partitioning)
        if getattr(self.F, "_mixed_aux", None) is None:
            self.F._mixed_aux = dict(info=DeviceInt(1), ginfo=DeviceInt(1),
                                     work=DeviceBuffer(8 * max(A64.m, 2) * (1 + -(-A64.n // bs))),
                                     norm=DeviceBuffer(16),
                                     bufs={})
        aux = self.F._mixed_aux
        self.info, self.ginfo = aux["info"], aux["ginfo"]
        self._work, self._norm, self._bufs = aux["work"], aux["norm"], aux["bufs"]
        if aux.get("isolve") is None:
            aux["isolve"] = _DiagInverseSolve(self.F, self.options.replace(f32_update=0))
        self._isolve = aux["isolve"]
        self.refactor()

    # Scaled diagonal shifts A + c u32 diag(A) tried in turn when the FP32
    # copy of A is not positive definite (kappa(A) u32 near or above 1);
    # c = 0 first.  The shifted factor is only a preconditioner: refinement
    # (GMRES-IR when classic refinement stalls) still solves A x = b
    # (Higham & Pranesh, "Exploiting lower precision arithmetic in solving
    # symmetric positive definite linear systems", SISC 2021).
    SHIFTS = (0.0, 4.0, 32.0, 256.0)

    def refactor(self):
        """(Re)build the FP32 factor from the current FP64 A; on a failed
        pivot retry with the next scaled diagonal shift (self.shift: the c
        that succeeded)."""
        s = self.stream
        ta = self.A.tiled()
        call("dm_tiled_norm_inf", C.byref(ta), C.c_void_p(self._norm.ptr),
             C.c_void_p(self._work.ptr), s)
        out = np.zeros(1)
        call("dm_memcpy_d2h", out.ctypes.data_as(C.c_void_p), C.c_void_p(self._norm.ptr),
             C.c_size_t(8), s)
        self.normA = float(out[0])
        t = self.F.tiled()
        o = self.options.c()
        u32 = 2.0 ** -24
        info = 0
        for c in self.SHIFTS:
            _convert(self.F, self.A, s)
            if c:
                call("dm_tiled_scale_diag", C.byref(t), 1.0 + c * u32, s)
            call("dm_cholesky_tiled", b"U", C.byref(t), C.c_void_p(self.info.ptr),
                 C.c_void_p(self.ginfo.ptr), C.byref(o), s)
            info = int(self.info.get(s)[0])
            if info == 0:
                self.shift = c
                break
        if info != 0:
            raise PosDefException(info)
        self._isolve.invert(s)

    # ---------------------------------------------------------------- pieces
    def _buffers(self, nrhs):
        # per right-hand-side count and per stream: callers sharing one
        # factor from different streams (the serving form) never share
        # scratch
        key = (nrhs, getattr(self.stream, "value", self.stream))
        b = self._bufs.get(key)
        if b is None:
            blk = Blocks(self.nb, self.nb)
            b = dict(X=DArray(np.float64, (self.A.m, nrhs), blk),
                     R=DArray(np.float64, (self.A.m, nrhs), blk),
                     D=DArray(np.float32, (self.A.m, nrhs), blk),
                     Y=DArray(np.float32, (self.A.m, nrhs), blk),
                     nrm=DeviceBuffer(8 * max(nrhs, 1)))
            self._bufs[key] = b
        return b

    def _solve32(self, D):
        """D := F32^-1 D (FP32; D one of the per-nrhs buffers)."""
        buf = self._buffers(D.n)
        self._isolve.solve(D, buf["Y"], self.stream)

    def residual(self, B, X, R):
        """R = B - A X in FP64."""
        call("dm_memcpy_d2d", C.c_void_p(R.buffer.ptr), C.c_void_p(B.buffer.ptr),
             C.c_size_t(B.m * B.n * 8), self.stream)
        ta, tx, tr = self.A.tiled(), X.tiled(), R.tiled()
        o = self.options.c()
        call("dm_gemm_tiled", b"N", b"N", -1.0, C.byref(ta), C.byref(tx), 1.0, C.byref(tr),
             C.byref(o), self.stream)

    # ---------------------------------------------------------------- solve
    def solve(self, B, tol=1e-13, max_iter=30, X=None, force_gmres=False):
        """Solve A X = B (B an FP64 DArray N x nrhs, Blocks(nb, nb)).

        Returns (X, info) with info = dict(iterations, berr (per column),
        method, history, escalated, shift).  Classic refinement runs until
        tol, max_iter or a stall; then GMRES-IR if tol is not met
        (force_gmres: straight after the first solve).  A factor built with
        the 2-term BF16 update that does not reach tol is rebuilt with the
        3-term update (escalated = True)."""
        if B.dtype != np.float64 or B.m != self.A.m:
            raise ValueError("B must be an FP64 DArray with A's rows")
        s = self.stream
        nrhs = B.n
        buf = self._buffers(nrhs)
        Xd = X if X is not None else buf["X"]
        R, D, nrm = buf["R"], buf["D"], buf["nrm"]
        _convert(D, B, s)
        self._solve32(D)
        _convert(Xd, D, s)
        nb_ = _colnorms(B, nrm, s)
        hist = []
        it = 0
        while True:
            self.residual(B, Xd, R)
            nr = _colnorms(R, nrm, s)
            nx = _colnorms(Xd, nrm, s)
            berr = nr / (self.normA * nx + nb_)
            hist.append(float(berr.max()))
            if berr.max() <= tol or it >= max_iter or force_gmres:
                break
            if it >= 3 and hist[-1] > 0.5 * hist[-3]:
                break  # stalled: kappa * u32 too close to 1
            _convert(D, R, s)
            self._solve32(D)
            td, tx = D.tiled(), Xd.tiled()
            call("dm_tiled_axpy_s2d", C.byref(tx), C.byref(td), s)
            it += 1
        method = "ir"
        if berr.max() > tol:
            method = "gmres-ir"
            it += self._gmres_ir(B, Xd, tol, max_iter)
            self.residual(B, Xd, R)
            berr = _colnorms(R, nrm, s) / (self.normA * _colnorms(Xd, nrm, s) + nb_)
        call("dm_stream_sync", s)
        if berr.max() > tol and self.options.f32_update == 3:
            # The 2-term BF16 update (about 16 significant bits) is too coarse
            # a preconditioner for this matrix (kappa * 2^-16 near or above
            # 1): refactor with the 3-term update (FP32-accurate) and solve
            # again from scratch.
            self.options = self.options.replace(f32_update=6)
            self.refactor()
            Xd, info = self.solve(B, tol=tol, max_iter=max_iter, X=X, force_gmres=force_gmres)
            info["iterations"] += it
            info["escalated"] = True
            return Xd, info
        return Xd, dict(iterations=it, berr=berr, method=method, history=hist, escalated=False,
                        shift=self.shift)

    def _gmres_ir(self, B, X, tol, max_outer, restart=30, inner_tol=1e-4):
        """GMRES-IR over all right-hand sides at once: outer FP64 refinement
        X += d, each correction solving (M^-1 A) d = M^-1 r by restarted
        GMRES in FP64 with M = F32^T F32 (one GMRES per column, advanced in
        lockstep).  The Krylov vectors stay on the device as N x nrhs blocks:
        A and M^-1 are applied to all columns in one product / one solve,
        modified Gram-Schmidt runs on column dot products and scaled updates
        (dm_tiled_col_dot / dm_tiled_col_axpby, coefficients on the device);
        only each column's (restart+1) x restart Hessenberg, its Givens
        rotations and the small triangular solve are on the host.  Returns
        the number of outer iterations."""
        n, r, s = self.A.m, B.n, self.stream
        blk = Blocks(self.nb, self.nb)
        buf = self._buffers(r)
        R, D, nrm = buf["R"], buf["D"], buf["nrm"]
        V = [DArray(np.float64, (n, r), blk) for _ in range(restart + 1)]
        W = DArray(np.float64, (n, r), blk)
        hbuf = DeviceBuffer(8 * (restart + 2) * r)
        cbuf = [DeviceBuffer(8 * r) for _ in range(restart + 1)]
        tV = [v.tiled() for v in V]
        tW, tX, ta = W.tiled(), X.tiled(), self.A.tiled()
        o = self.options.c()
        keep = []  # host coefficient arrays until the next synchronizing copy

        def coef(vals, slot=0):
            arr = np.ascontiguousarray(vals, dtype=np.float64)
            keep.append(arr)
            call("dm_memcpy_h2d", C.c_void_p(cbuf[slot].ptr), arr.ctypes.data_as(C.c_void_p),
                 C.c_size_t(8 * r), s)
            return C.c_void_p(cbuf[slot].ptr)

        def hdev(i):
            return C.c_void_p(hbuf.ptr + 8 * i * r)

        def host(count):  # rows 0..count-1 of the device coefficients (synchronizes)
            out = np.empty(count * r)
            call("dm_memcpy_d2h", out.ctypes.data_as(C.c_void_p), C.c_void_p(hbuf.ptr),
                 C.c_size_t(8 * count * r), s)
            keep.clear()
            return out.reshape(count, r)

        def precond_into(dst_t):  # dst = M^-1 R (R holds the operand)
            _convert(D, R, s)
            self._solve32(D)
            td = D.tiled()
            call("dm_tiled_convert", C.byref(dst_t), C.byref(td), s)

        nb_ = _colnorms(B, nrm, s)
        total = 0
        try:
            for _ in range(max_outer):
                self.residual(B, X, R)
                berr = _colnorms(R, nrm, s) / (self.normA * _colnorms(X, nrm, s) + nb_)
                if berr.max() <= tol:
                    break
                precond_into(tV[0])
                call("dm_tiled_col_dot", C.byref(tV[0]), C.byref(tV[0]), hdev(0), s)
                beta = np.sqrt(np.maximum(host(1)[0], 0.0))
                live = (beta > 0) & (berr > tol)
                safe = np.where(beta > 0, beta, 1.0)
                call("dm_tiled_col_axpby", coef(np.where(live, 1.0 / safe, 0.0)), 1.0,
                     C.byref(tV[0]), 0.0, C.byref(tV[0]), s)
                H = np.zeros((r, restart + 1, restart))
                cs, sn = np.zeros((r, restart)), np.zeros((r, restart))
                g = np.zeros((r, restart + 1))
                g[:, 0] = np.where(live, beta, 0.0)
                used = 0
                for k in range(restart):
                    # W = M^-1 A V_k for every column at once
                    call("dm_gemm_tiled", b"N", b"N", 1.0, C.byref(ta), C.byref(tV[k]), 0.0,
                         C.byref(tW), C.byref(o), s)
                    call("dm_memcpy_d2d", C.c_void_p(R.buffer.ptr), C.c_void_p(W.buffer.ptr),
                         C.c_size_t(8 * n * r), s)
                    precond_into(tW)
                    # modified Gram-Schmidt: h_i = V_i . W, W -= h_i V_i (per column)
                    for i in range(k + 1):
                        call("dm_tiled_col_dot", C.byref(tV[i]), C.byref(tW), hdev(i), s)
                        call("dm_tiled_col_axpby", hdev(i), -1.0, C.byref(tV[i]), 1.0,
                             C.byref(tW), s)
                    call("dm_tiled_col_dot", C.byref(tW), C.byref(tW), hdev(k + 1), s)
                    h = host(k + 2)
                    hk1 = np.sqrt(np.maximum(h[k + 1], 0.0))
                    call("dm_tiled_col_axpby",
                         coef(np.where(live & (hk1 > 0), 1.0 / np.where(hk1 > 0, hk1, 1.0), 0.0)),
                         1.0, C.byref(tW), 0.0, C.byref(tV[k + 1]), s)
                    for j in np.nonzero(live)[0]:
                        col = np.append(h[:k + 1, j], hk1[j])
                        for i in range(k):
                            t = cs[j, i] * col[i] + sn[j, i] * col[i + 1]
                            col[i + 1] = -sn[j, i] * col[i] + cs[j, i] * col[i + 1]
                            col[i] = t
                        rho = np.hypot(col[k], col[k + 1])
                        cs[j, k], sn[j, k] = (col[k] / rho, col[k + 1] / rho) if rho > 0 else (1.0, 0.0)
                        col[k], col[k + 1] = rho, 0.0
                        H[j, :k + 2, k] = col
                        g[j, k + 1] = -sn[j, k] * g[j, k]
                        g[j, k] = cs[j, k] * g[j, k]
                    used = k + 1
                    if np.all(~live | (np.abs(g[:, k + 1]) <= inner_tol * safe)):
                        break
                # y_j = H_j^-1 g_j, then X += sum_i V_i diag(y_i)
                Y = np.zeros((used, r))
                for j in np.nonzero(live)[0]:
                    Hj = np.triu(H[j, :used, :used])
                    if np.all(np.diag(Hj) != 0):
                        Y[:, j] = np.linalg.solve(Hj, g[j, :used])
                    else:
                        Y[:, j] = np.linalg.lstsq(Hj, g[j, :used], rcond=None)[0]
                for i in range(used):
                    call("dm_tiled_col_axpby", coef(Y[i], slot=i), 1.0, C.byref(tV[i]), 1.0,
                         C.byref(tX), s)
                total += 1
        finally:
            call("dm_stream_sync", s)
            keep.clear()
            for v in V:
                v.unsafe_free()
            W.unsafe_free()
            hbuf.free()
            for c in cbuf:
                c.free()
        return total


class HouseholderSpec:
    """Host algebra of A = H diag(s) H^T (kappa(A) = kappa exactly), H a
    product of `nrefl` reflectors from seeded Gaussian vectors (device
    Philox stream `seed`), s log-spaced from 1 to 1/kappa.  Compact WY
    (dlarft, forward columnwise): H = I - V T V^T, so
      A = S - Q V^T - V Q^T = S - U W^T,  U = [Q V], W = [V Q],
      Y = S V T^T, Z = T V^T Y, Q = Y - V Z / 2."""

    def __init__(self, n, kappa, seed, nrefl=64, stream=None, scratch=None):
        self.n, self.kappa, self.r = n, float(kappa), nrefl
        V = _philox_randn(n * nrefl, seed, stream, scratch).reshape(nrefl, n).T.copy()
        # T from the UT transform (Joffrain, Low, Quintana-Orti, van de Geijn,
        # Van Zee, "Accumulating Householder transformations, revisited",
        # ACM TOMS 32(2), 2006): with tau_i = 2 / v_i^T v_i,
        # T^-1 = striu(V^T V) + diag(V^T V) / 2 -- the same T as dlarft's
        # column recurrence, from one product and one inverse (LU of an upper
        # triangular matrix needs no row exchange, so T comes out upper
        # triangular).
        G = V.T @ V
        Tinv = np.triu(G, 1) + np.diag(0.5 * np.diag(G))
        T = np.triu(np.linalg.inv(Tinv))
        self.V, self.T = V, T
        self.s = self.kappa ** (-np.arange(n) / max(n - 1, 1))
        Y = self.s[:, None] * (V @ T.T)
        Z = T @ (V.T @ Y)
        Q = Y - 0.5 * (V @ Z)
        self.U = np.hstack([Q, V])
        self.W = np.hstack([V, Q])

    def matvec(self, x):
        x2 = x.reshape(self.n, -1)
        return (self.s[:, None] * x2 - self.U @ (self.W.T @ x2)).reshape(x.shape)


def householder_fill(D, spec, stream=None, scratch=None):
    """Write A = H diag(s) H^T into the resident DArray D (n_pad >= n):
    D = [A 0; 0 I] (zero rows appended to U, W and ones to s), via one FP64
    GEMM n_pad x n_pad x 2r, the diagonal add and the tile pack.  `scratch`
    (a Scratch) provides the device temporaries."""
    n, npad = spec.n, D.m
    r2 = spec.U.shape[1]
    Up = np.zeros((npad, r2), order="F")
    Wp = np.zeros((npad, r2), order="F")
    Up[:n] = spec.U
    Wp[:n] = spec.W
    sp = np.ones(npad)
    sp[:n] = spec.s
    with _Temp(scratch, "hh_u", npad * r2 * 8) as U, _Temp(scratch, "hh_w", npad * r2 * 8) as W, \
            _Temp(scratch, "hh_s", npad * 8) as sd, _Temp(scratch, "hh_full", npad * npad * 8) as Fb:
        call("dm_memcpy_h2d", C.c_void_p(U.ptr), Up.ctypes.data_as(C.c_void_p),
             C.c_size_t(Up.nbytes), stream)
        call("dm_memcpy_h2d", C.c_void_p(W.ptr), Wp.ctypes.data_as(C.c_void_p),
             C.c_size_t(Wp.nbytes), stream)
        call("dm_memcpy_h2d", C.c_void_p(sd.ptr), sp.ctypes.data_as(C.c_void_p),
             C.c_size_t(8 * npad), stream)
        call("dm_dgemm", b"N", b"T", npad, npad, r2, -1.0, C.c_void_p(U.ptr), npad,
             C.c_void_p(W.ptr), npad, 0.0, C.c_void_p(Fb.ptr), npad, stream)
        call("dm_add_diag_d", C.c_void_p(Fb.ptr), npad, npad, C.c_void_p(sd.ptr), stream)
        t = D.tiled()
        call("dm_tiled_pack", C.byref(t), C.c_void_p(Fb.ptr), npad, 0.0, stream)
        call("dm_stream_sync", stream)
    return D


def householder_generate(n, blocks, kappa, seed, nrefl=64, stream=None):
    """A = H diag(s) H^T as a new resident DArray; returns (DArray, spec)."""
    spec = HouseholderSpec(n, kappa, seed, nrefl, stream)
    D = DArray(np.float64, (n, n), blocks)
    householder_fill(D, spec, stream)
    return D, spec


def _philox_randn(count, seed, stream=None, scratch=None):
    """The device Philox/Box-Muller stream (count values, offset 0)."""
    out = np.empty(count)
    with _Temp(scratch, "randn", 8 * count) as buf:
        call("dm_fill_randn_d", C.c_void_p(buf.ptr), C.c_size_t(count), C.c_uint64(seed),
             C.c_uint64(0), stream)
        call("dm_memcpy_d2h", out.ctypes.data_as(C.c_void_p), C.c_void_p(buf.ptr),
             C.c_size_t(8 * count), stream)
    return out
