"""Every entry point of the library at small shapes, run on the CPU emulation
build (tests/cpu_emu) under AddressSanitizer: an out-of-bounds access in any
kernel or host path aborts with a report, a barrier or wave op that not all
work-items reach aborts as a deadlock.  Results are also checked against
numpy so that the emulated MFMA lane layouts stay honest.

Run by tests/test_cpu_emulation.py as
    LD_PRELOAD=<asan runtime> DM_TEST_EMU_LIB=<emu lib> python scenarios.py [group ...]
"""
import ctypes as C
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "dagger.jl_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import dagger_mi355x as dm  # noqa: E402
from oracle import dagger_tiled as O  # noqa: E402
from oracle import gen  # noqa: E402

dm._lib.load(os.environ["DM_TEST_EMU_LIB"])
dm.init(0)
call = dm._lib.call


class Dev:
    def __init__(self, a):
        self.a = np.asfortranarray(a)
        self.buf = dm.DeviceBuffer(max(self.a.nbytes, 8))
        call("dm_memcpy_h2d", C.c_void_p(self.buf.ptr), self.a.ctypes.data_as(C.c_void_p),
             C.c_size_t(self.a.nbytes), None)

    @property
    def p(self):
        return C.c_void_p(self.buf.ptr)

    def get(self):
        out = np.empty_like(self.a, order="F")
        call("dm_memcpy_d2h", out.ctypes.data_as(C.c_void_p), C.c_void_p(self.buf.ptr),
             C.c_size_t(out.nbytes), None)
        return out


def relerr(x, ref):
    return float(np.linalg.norm(x - ref) / max(np.linalg.norm(ref), 1e-300))


def check(cond, what):
    if not cond:
        raise AssertionError(what)


# ---------------------------------------------------------------- tile ops
def prefetch_gemms(rng, cases):
    """dm_dgemm on products of at most 64 columns with K >= 512 in launches
    of >= 64 tiles of 64 x 64 (the 64 x 64 tile with two-deep prefetch)."""
    for (ta, tb, m, n, k) in cases:
        A = rng.standard_normal((k, m) if ta == "T" else (m, k))
        B = rng.standard_normal((n, k) if tb == "T" else (k, n))
        Cm = rng.standard_normal((m, n))
        dA, dB, dC = Dev(A), Dev(B), Dev(Cm)
        call("dm_dgemm", ta.encode(), tb.encode(), m, n, k, -1.0, dA.p, A.shape[0], dB.p,
             B.shape[0], 1.0, dC.p, m, None)
        ref = Cm - (A.T if ta == "T" else A) @ (B.T if tb == "T" else B)
        check(relerr(dC.get(), ref) < 1e-13, ("dgemm 64-column prefetch", ta, tb, m, n, k))


def tiles():
    rng = np.random.default_rng(1)
    prefetch_gemms(rng, (("N", "N", 4096, 64, 544), ("N", "T", 4100, 50, 1000),
                         ("T", "T", 4096, 64, 512), ("T", "N", 4100, 32, 600)))
    for ta in "NT":
        for tb in "NT":
            for (m, n, k) in [(200, 136, 77), (1, 1, 1), (64, 33, 0)]:
                A = rng.standard_normal((k, m) if ta == "T" else (m, k))
                B = rng.standard_normal((n, k) if tb == "T" else (k, n))
                Cm = rng.standard_normal((m, n))
                dA, dB, dC = Dev(A), Dev(B), Dev(Cm)
                call("dm_dgemm", ta.encode(), tb.encode(), m, n, k, -1.0, dA.p,
                     max(A.shape[0], 1), dB.p, max(B.shape[0], 1), 0.5, dC.p, m, None)
                ref = 0.5 * Cm - (A.T if ta == "T" else A) @ (B.T if tb == "T" else B)
                check(relerr(dC.get(), ref) < 1e-13, ("dgemm", ta, tb, m, n, k))
    # 128x128 workgroup tiles (>= 192 of them), odd leading dimensions
    m, n, k = 2048, 1536, 4
    A, B = rng.standard_normal((m, k)), rng.standard_normal((k, n))
    dA, dB, dC = Dev(A), Dev(B), Dev(np.zeros((m, n)))
    call("dm_dgemm", b"N", b"N", m, n, k, 1.0, dA.p, m, dB.p, k, 0.0, dC.p, m, None)
    check(relerr(dC.get(), A @ B) < 1e-13, "dgemm big tiles")
    # 256x128 workgroup tiles (>= 256 of them), plain and as a SYRK whose
    # rectangular tiles outside the triangle are skipped
    m, n, k = 4096, 2048, 4
    A, B = rng.standard_normal((k, m)), rng.standard_normal((k, n))
    dA, dB, dC = Dev(A), Dev(B), Dev(np.zeros((m, n)))
    call("dm_dgemm", b"T", b"N", m, n, k, 1.0, dA.p, k, dB.p, k, 0.0, dC.p, m, None)
    check(relerr(dC.get(), A.T @ B) < 1e-13, "dgemm 256x128 tiles")
    for uplo in "UL":
        n = 4096
        A = rng.standard_normal((k, n))
        Cm = rng.standard_normal((n, n))
        dA, dC = Dev(A), Dev(Cm)
        call("dm_dsyrk", uplo.encode(), b"T", n, k, -1.0, dA.p, k, 1.0, dC.p, n, None)
        got = dC.get()
        tri = np.triu if uplo == "U" else np.tril
        other = np.tril if uplo == "U" else np.triu
        off = -1 if uplo == "U" else 1
        check(relerr(tri(got), tri(Cm - A.T @ A)) < 1e-13 and
              np.array_equal(other(got, off), other(Cm, off)), ("dsyrk 256x128 tiles", uplo))
    m, n, k, ld = 101, 67, 55, 103
    A, B = rng.standard_normal((ld, k)), rng.standard_normal((ld, n))
    dA, dB, dC = Dev(A), Dev(B), Dev(np.zeros((ld, n)))
    call("dm_dgemm", b"N", b"N", m, n, k, 1.0, dA.p, ld, dB.p, ld, 0.0, dC.p, ld, None)
    got = dC.get()
    check(relerr(got[:m], A[:m] @ B[:k]) < 1e-13 and np.all(got[m:] == 0), "dgemm odd ld")
    for uplo in "UL":
        for trans in "NT":
            for (n, k) in [(130, 70), (5, 3)]:
                A = rng.standard_normal((k, n) if trans == "T" else (n, k))
                Cm = rng.standard_normal((n, n))
                dA, dC = Dev(A), Dev(Cm)
                call("dm_dsyrk", uplo.encode(), trans.encode(), n, k, -1.0, dA.p, A.shape[0],
                     1.0, dC.p, n, None)
                ref = Cm.copy(order="F")
                O.tile_syrk(uplo, trans, -1.0, np.asfortranarray(A), 1.0, ref)
                tri = np.triu if uplo == "U" else np.tril
                check(relerr(tri(dC.get()), tri(ref)) < 1e-13, ("dsyrk", uplo, trans, n, k))
    for side in "LR":
        for uplo in "UL":
            for trans in "NT":
                for diag in "NU":
                    m, n = (70, 30) if side == "L" else (30, 70)
                    ka = m if side == "L" else n
                    A = 0.05 * rng.standard_normal((ka, ka)) + 2 * np.eye(ka)
                    B = rng.standard_normal((m, n))
                    dA, dB = Dev(A), Dev(B)
                    call("dm_dtrsm", side.encode(), uplo.encode(), trans.encode(),
                         diag.encode(), m, n, 0.75, dA.p, ka, dB.p, m, None)
                    ref = B.copy(order="F")
                    O.tile_trsm(side, uplo, trans, diag, 0.75, np.asfortranarray(A), ref)
                    check(relerr(dB.get(), ref) < 1e-12, ("dtrsm", side, uplo, trans, diag))
    # solves longer than the LDS-resident limit (host recursion)
    for side, uplo, trans in [("L", "U", "T"), ("R", "L", "T")]:
        ka, nv = 1100, 3
        F, _ = O.cholesky(gen.spd(ka, 5), ka, uplo)
        m, n = (ka, nv) if side == "L" else (nv, ka)
        B = rng.standard_normal((m, n))
        dA, dB = Dev(F), Dev(B)
        call("dm_dtrsm", side.encode(), uplo.encode(), trans.encode(), b"N", m, n, 1.0, dA.p,
             ka, dB.p, m, None)
        ref = B.copy(order="F")
        O.tile_trsm(side, uplo, trans, "N", 1.0, np.asfortranarray(F), ref)
        check(relerr(dB.get(), ref) < 1e-12, ("dtrsm long", side, uplo, trans))
    for uplo in "UL":
        for n in (1, 64, 100, 200):
            A = gen.spd(n, 900 + n)
            dA = Dev(A)
            info = dm.DeviceInt(1, value=-7)
            call("dm_dpotrf", uplo.encode(), n, dA.p, n, C.c_void_p(info.ptr), None)
            ref, rinfo = O.cholesky(A, n, uplo)
            check(int(info.get()[0]) == 0 and O.rel_factor_diff(dA.get(), ref, uplo) < 1e-12,
                  ("dpotrf", uplo, n))
    A = gen.spd(150, 5)
    A[130, 130] = -1.0
    dA, info = Dev(A), dm.DeviceInt(1)
    call("dm_dpotrf", b"U", 150, dA.p, 150, C.c_void_p(info.ptr), None)
    check(int(info.get()[0]) == 131, "dpotrf info")
    for trans in "NT":
        m, n = 300, 100
        A = rng.standard_normal((m, n))
        x = rng.standard_normal(n if trans == "N" else m)
        y = rng.standard_normal(m if trans == "N" else n)
        dA, dx, dy = Dev(A), Dev(x), Dev(y)
        call("dm_dgemv", trans.encode(), m, n, -1.0, dA.p, m, dx.p, 1, 2.0, dy.p, 1, None)
        ref = 2.0 * y - (A.T @ x if trans == "T" else A @ x)
        check(relerr(dy.get(), ref) < 1e-13, ("dgemv", trans))
    for uplo in "UL":
        for trans in "NT":
            n = 150
            F, _ = O.cholesky(gen.spd(n, 3), n, uplo)
            x = gen.randn(n, 4)
            dA, dx = Dev(F), Dev(x)
            call("dm_dtrsv", uplo.encode(), trans.encode(), b"N", n, dA.p, n, dx.p, 1, None)
            ref = x.copy()
            O.tile_trsv(uplo, trans, "N", np.asfortranarray(F), ref)
            check(relerr(dx.get(), ref) < 1e-12, ("dtrsv", uplo, trans))
    for ta, tb in [("T", "N"), ("N", "T")]:
        m = n = k = 96
        A = rng.standard_normal((m, k)).astype(np.float32)
        B = rng.standard_normal((k, n)).astype(np.float32)
        Cm = rng.standard_normal((m, n)).astype(np.float32)
        dA, dB, dC = Dev(A), Dev(B), Dev(Cm)
        call("dm_sgemm", ta.encode(), tb.encode(), m, n, k, C.c_float(1.0), dA.p, m, dB.p, k,
             C.c_float(1.0), dC.p, m, None)
        opA = A.T if ta == "T" else A
        opB = B.T if tb == "T" else B
        ref = opA.astype(np.float64) @ opB.astype(np.float64) + Cm
        check(relerr(dC.get().astype(np.float64), ref) < 1e-5, ("sgemm", ta, tb))
    for terms, tol in [(6, 2e-6), (3, 1e-4)]:
        for (m, n, k, ld) in [(200, 136, 77, 80), (129, 1, 33, 36), (64, 64, 64, 64)]:
            A = np.zeros((ld, m), np.float32)
            B = np.zeros((ld, n), np.float32)
            A[:k] = rng.standard_normal((k, m))
            B[:k] = rng.standard_normal((k, n))
            Cm = rng.standard_normal((m, n)).astype(np.float32)
            dA, dB, dC = Dev(A), Dev(B), Dev(Cm)
            call("dm_sgemm_bf16x", terms, b"T", b"N", m, n, k, C.c_float(-1.0), dA.p, ld, dB.p,
                 ld, C.c_float(1.0), dC.p, m, None)
            ref = Cm.astype(np.float64) - A[:k].T.astype(np.float64) @ B[:k].astype(np.float64)
            check(relerr(dC.get().astype(np.float64), ref) < tol, ("bf16x", terms, m, n, k))
    n = 5000
    buf = dm.DeviceBuffer(8 * n)
    call("dm_fill_randn_d", C.c_void_p(buf.ptr), C.c_size_t(n), C.c_uint64(42), C.c_uint64(5),
         None)
    out = np.empty(n)
    call("dm_memcpy_d2h", out.ctypes.data_as(C.c_void_p), C.c_void_p(buf.ptr),
         C.c_size_t(8 * n), None)
    check(np.max(np.abs(out - gen.randn(n, 42, offset=5))) < 1e-13, "randn")


# ---------------------------------------------------------------- tiled path
def tiled():
    n, nb = 300, 64  # ragged: 4 x 64 + 44
    A = gen.spd(n, 6)
    outs = {}
    for uplo in "UL":
        ref, rinfo = O.cholesky(A, nb, uplo)
        for opts in (dm.SEQUENTIAL,
                     dm.PlanOptions(num_streams=2, batch=True, lookahead=1, use_graph=False),
                     dm.PlanOptions(num_streams=3, batch=True, lookahead=2, use_graph=False),
                     dm.PlanOptions(num_streams=2, batch=True, lookahead=1, use_graph=True)):
            DA = dm.DArray.from_array(A, dm.Blocks(nb, nb))
            dm.cholesky_(DA, uplo=uplo, options=opts)
            f = DA.collect()
            check(O.rel_factor_diff(f, ref, uplo) < 1e-12, ("cholesky", uplo))
            outs.setdefault(uplo, f)
            # Bit-identical across schedules except for GEMMs on tiles of at
            # most 48 columns (here the ragged 44-column edge): alone they run
            # the K-split skinny kernel, batched with 64-column tiles the
            # regular one, so the two agree to rounding.
            check(np.abs(f - outs[uplo]).max() <= 4 * np.finfo(float).eps * np.abs(f).max(),
                  ("schedules differ", uplo))
        DA = dm.DArray.from_array(A, dm.Blocks(nb, nb))
        Fc = dm.cholesky_(DA, uplo=uplo)
        for nrhs in (5, 1):
            b = gen.randn(n * nrhs, 8).reshape(nrhs, n).T.copy()
            DB = dm.DArray.from_array(b, dm.Blocks(nb, nb))
            dm.ldiv_(Fc, DB)
            check(O.backward_error(A, DB.collect(), b) < 10 * n * np.finfo(float).eps,
                  ("ldiv", uplo, nrhs))
        bv = gen.randn(n, 9)
        Dv = dm.DArray.from_array(bv, dm.Blocks(nb))
        dm.ldiv_(Fc, Dv)
        check(O.backward_error(A, Dv.collect().reshape(-1, 1), bv.reshape(-1, 1)) <
              10 * n * np.finfo(float).eps, ("ldiv vector", uplo))
    # tiles longer than the leaf solve: the executor's recursive TRSM
    n2, nb2 = 600, 288
    A2 = gen.spd(n2, 16)
    for uplo in "UL":
        ref, _ = O.cholesky(A2, nb2, uplo)
        DA = dm.DArray.from_array(A2, dm.Blocks(nb2, nb2))
        Fc = dm.cholesky_(DA, uplo=uplo)
        check(O.rel_factor_diff(DA.collect(), ref, uplo) < 1e-12, ("cholesky long tiles", uplo))
        b = gen.randn(n2 * 3, 17).reshape(3, n2).T.copy()
        DB = dm.DArray.from_array(b, dm.Blocks(nb2, nb2))
        dm.ldiv_(Fc, DB)
        check(O.backward_error(A2, DB.collect(), b) < 10 * n2 * np.finfo(float).eps,
              ("ldiv long tiles", uplo))
    # non-PD: tile-local info
    Abad = A.copy()
    Abad[200, 200] = -1.0
    DA = dm.DArray.from_array(Abad, dm.Blocks(nb, nb))
    try:
        dm.cholesky_(DA)
        raise AssertionError("non-PD not reported")
    except dm.PosDefException:
        pass
    # tiled trsm branches and tiled gemm
    F, _ = O.cholesky(A, nb, "U")
    F = np.triu(F)
    rng = np.random.default_rng(3)
    for side, uplo, trans in [("L", "U", "T"), ("L", "U", "N"), ("L", "L", "N"), ("L", "L", "T"),
                              ("R", "U", "N"), ("R", "L", "T")]:
        T = F if uplo == "U" else F.T.copy()
        shape = (n, 7) if side == "L" else (7, n)
        B = rng.standard_normal(shape)
        DT = dm.DArray.from_array(T, dm.Blocks(nb, nb))
        DB = dm.DArray.from_array(B, dm.Blocks(nb, nb))
        dm.trsm_(side, uplo, trans, "N", 0.5, DT, DB)
        # Dagger's right-transposed branches apply alpha at every block
        # step (src/array/trsm.jl:157-163,177-183): compare with the tiled
        # restatement, which follows the reference loop nest.
        ref = O.trsm(side, uplo, trans, "N", 0.5, T, B, nb)
        check(relerr(DB.collect(), ref) < 1e-11, ("trsm tiled", side, uplo, trans))
    X = rng.standard_normal((n, 9))
    Y = rng.standard_normal((n, 9))
    DA = dm.DArray.from_array(A, dm.Blocks(nb, nb))
    DX = dm.DArray.from_array(X, dm.Blocks(nb, nb))
    DY = dm.DArray.from_array(Y, dm.Blocks(nb, nb))
    ta, tx, ty = DA.tiled(), DX.tiled(), DY.tiled()
    o = dm.PlanOptions().c()
    call("dm_gemm_tiled", b"N", b"N", -1.0, C.byref(ta), C.byref(tx), 1.0, C.byref(ty),
         C.byref(o), None)
    check(relerr(DY.collect(), Y - A @ X) < 1e-13, "gemm tiled")
    check(dm.ishermitian(dm.DArray.from_array(A, dm.Blocks(nb, nb))), "ishermitian")
    Ab = A.copy()
    Ab[5, 250] += 1e-3
    check(not dm.ishermitian(dm.DArray.from_array(Ab, dm.Blocks(nb, nb))), "not hermitian")
    # re-blocking (Blocks mismatch) and FP32 tiles with every update arithmetic
    DA = dm.DArray.from_array(A, dm.Blocks(96, 64))
    dm.cholesky_(DA)
    check(O.rel_factor_diff(DA.collect(), O.cholesky(A, nb, "U")[0], "U") < 1e-12, "reblock")
    for upd in (0, 6, 3):
        F32 = dm.DArray.from_array(A.astype(np.float32), dm.Blocks(nb, nb))
        info = dm.DeviceInt(2)
        t = F32.tiled()
        opts = dm.PlanOptions(f32_update=upd).c()
        call("dm_cholesky_tiled", b"U", C.byref(t), C.c_void_p(info.ptr),
             C.c_void_p(info.ptr + 4), C.byref(opts), None)
        f = np.triu(F32.collect().astype(np.float64))
        U = np.linalg.cholesky(A).T
        check(relerr(f, U) < (1e-5 if upd == 3 else 1e-6), ("fp32 cholesky", upd))


# ---------------------------------------------------------------- mixed
def mixed():
    from dagger_mi355x import mixed as M
    n, nb = 256, 64
    for kappa, upd in [(1e4, 0), (1e4, 6), (1e6, 3)]:
        D, spec = M.householder_generate(n, dm.Blocks(nb, nb), kappa, seed=11, nrefl=16)
        host = gen.HouseholderSPD(n, kappa, 11, nrefl=16)
        check(np.allclose(D.collect(), host.dense(), rtol=0, atol=1e-12), "householder")
        F = M.MixedCholesky(D, options=M._OPTS.replace(f32_update=upd))
        xt = gen.randn(n * 3, 12).reshape(3, n).T.copy()
        b = host.matvec(xt)
        X, info = F.solve(dm.DArray.from_array(b, dm.Blocks(nb, nb)), tol=1e-13, max_iter=30)
        # the solve's own criterion: per column (dsposv), not the matrix norm
        check(O.backward_error_columns(host.dense(), X.collect(), b) <= 1e-13,
              ("mixed", kappa, upd))


# ---------------------------------------------------------------- batched
def batched():
    from dagger_mi355x import batched as Bt
    for dtype in (np.float64, np.float32):
        sizes, nrhs = Bt.log_uniform_sizes(8, 32, 120, seed=5)
        sizes[:3] = [32, 33, 120]
        bs = Bt.BatchedSystems(sizes, nrhs, dtype).generate(seed=100)
        A0s = [bs.get_system(m) for m in range(bs.count)]
        bs.posv(panels={})
        check(np.all(bs.infos() == 0), "batched info")
        for m in range(bs.count):
            A0, B0 = A0s[m]
            F, X = bs.get_system(m)
            if B0.shape[1]:
                be = O.backward_error(A0.astype(np.float64), X.astype(np.float64),
                                      B0.astype(np.float64))
                check(be <= 10 * A0.shape[0] * np.finfo(dtype).eps, ("batched", dtype, m))
    sizes, nrhs = Bt.log_uniform_sizes(6, 32, 100, seed=6)
    bs = Bt.BatchedSystems(sizes, nrhs, np.float64).generate(seed=7)
    snap = bs.snapshot()
    for P in Bt.PANELS:
        bs.restore(snap)
        bs.posv(panels={b: P for b in range(16)})
    bs.restore(snap)
    Bt.autotune(bs, sample=3, reps=1)
    try:  # order entries outside the item table are rejected
        order = np.array([0, 99], np.int32)
        it = bs.items({})
        call("dm_batched_posv", dm._lib.DM_F64, bs.count, C.c_void_p(bs.A.ptr),
             C.c_void_p(bs.B.ptr), C.cast(it, C.c_void_p), 2, order.ctypes.data_as(C.c_void_p),
             C.c_void_p(bs.info.ptr), 0, 0, None)
        raise AssertionError("bad order accepted")
    except dm.DMError:
        pass


def batched_tiled():
    """Tiled level-synchronous path of the batching form (outer tiles of
    128 columns over 64-column blocks, so the larger systems take several of
    each, with ragged last blocks), mixed with
    the persistent kernel by BatchedSystems.solve; a non-positive pivot in a
    later tile of one system gives its LAPACK info and leaves its B as it
    was, while the other systems are solved."""
    from dagger_mi355x import batched as Bt
    for dtype in (np.float64, np.float32):
        sizes = np.array([200, 64, 130, 40, 65, 129, 96])
        nrhs = np.array([2, 1, 0, 3, 1, 8, 2])
        bs = Bt.BatchedSystems(sizes, nrhs, dtype).generate(seed=11)
        if dtype == np.float64:  # system 5: pivot 100 (second tile) not positive
            A5, _ = bs.get_system(5)
            A5[99, 99] = -1.0
            bs.put_system(5, A=A5)
        A0s = [bs.get_system(m) for m in range(bs.count)]
        # per-bucket outer tiles: 64 for n in [64, 128), 192 for [128, 256)
        tile_nb = {6: 64, 7: 192} if dtype == np.float64 else None
        bs.solve(panels={}, tiled_min=64, nb=128, tile_nb=tile_nb)
        infos = bs.infos()
        tol = 10 * np.finfo(dtype).eps
        for m in range(bs.count):
            A0, B0 = A0s[m]
            F, X = bs.get_system(m)
            ref = np.array(A0, dtype=np.float64, order="F")
            want = O.tile_potrf("U", ref)
            check(infos[m] == want, ("batched tiled info", dtype, m, infos[m], want))
            if want:
                check(np.array_equal(X, B0), ("failed system's B changed", dtype, m))
                continue
            rtol = 1e-12 if dtype == np.float64 else 5e-5
            check(O.rel_factor_diff(F.astype(np.float64), ref, "U") <= rtol,
                  ("batched tiled factor", dtype, m))
            if B0.shape[1]:
                be = O.backward_error(A0.astype(np.float64), X.astype(np.float64),
                                      B0.astype(np.float64))
                check(be <= tol * A0.shape[0], ("batched tiled solve", dtype, m, be))
        # a second run of the cached plan on restored data gives the same bits
        if dtype == np.float64:
            F1 = [bs.get_system(m) for m in range(bs.count)]
            for m in range(bs.count):
                bs.put_system(m, A=A0s[m][0], B=A0s[m][1])
            bs.solve(panels={}, tiled_min=64, nb=128, tile_nb=tile_nb)
            for m in range(bs.count):
                F2, X2 = bs.get_system(m)
                check(np.array_equal(F2, F1[m][0]) and np.array_equal(X2, F1[m][1]),
                      ("batched tiled rerun", m))
        bs.close()
    # edge sizes in one plan: 1, 2, ragged around the 64-wide block, and an
    # empty system (its info must read 0); ld padded to a multiple of 8
    sizes = np.array([1, 2, 63, 64, 65, 0, 130])
    nrhs = np.array([1, 2, 3, 0, 8, 1, 2])
    bs = Bt.BatchedSystems(sizes, nrhs, np.float64, ld_pad=8).generate(seed=13)
    A0s = [bs.get_system(m) for m in range(bs.count)]
    call("dm_fill_i32", C.c_void_p(bs.info.ptr), 7, C.c_size_t(bs.count), None)
    bs.solve(panels={}, nb=64)
    infos = bs.infos()
    check(np.all(infos == 0), ("edge sizes info", infos))
    for m in range(bs.count):
        A0, B0 = A0s[m]
        if sizes[m] == 0:
            continue
        F, X = bs.get_system(m)
        ref = np.array(A0, order="F")
        O.tile_potrf("U", ref)
        check(O.rel_factor_diff(F, ref, "U") <= 1e-12, ("edge sizes factor", m))
        if B0.shape[1]:
            check(O.backward_error(A0, X, B0) <= 10 * max(sizes[m], 1) * np.finfo(float).eps,
                  ("edge sizes solve", m))
    bs.close()
    for bad in (32, 96 + 1, 2048):  # nb not a multiple of 64 in [64, 1024]
        try:
            bs = Bt.BatchedSystems([70], [1]).generate(seed=1)
            Bt.TiledPlan(bs, [0], nb=bad)
            raise AssertionError(("bad tiled plan accepted", bad))
        except dm.DMError:
            pass
    try:  # a per-item tile that is not a multiple of 64
        bs = Bt.BatchedSystems([70], [1]).generate(seed=1)
        Bt.TiledPlan(bs, [0], nb=128, tile_nb={6: 100})
        raise AssertionError("bad item tile accepted")
    except dm.DMError:
        pass


# ---------------------------------------------------------------- serving
def serving():
    from dagger_mi355x import serving as S
    trace = S.make_trace(rate=400.0, count=6, seed=3, n_lo=64, n_hi=200, k_lo=1e2, k_hi=1e4,
                         reuse=0.3, tols=(1e-8, 1e-13), nrhs_max=4)
    srv = S.Server(workers=2, budget_bytes=1e8)
    stats = srv.run(trace)
    srv.close()
    check(stats["completed"] == len(trace), "serving completed")
    for r in trace:
        check(r.berr <= r.tol, ("serving tol", r.rid))


def capped():
    """Bulk launches capped below their tile count (EMU_CUS=8 in the
    environment makes the executor cap bulk launches at 7 CUs' worth), so
    GEMM and TRSM workgroups loop over several tiles each."""
    n, nb = 600, 288
    A = gen.spd(n, 26)
    for uplo in "UL":
        ref, _ = O.cholesky(A, nb, uplo)
        DA = dm.DArray.from_array(A, dm.Blocks(nb, nb))
        opts = dm.PlanOptions(2, True, 1, use_graph=False)
        F = dm.cholesky_(DA, uplo=uplo, options=opts)
        check(O.rel_factor_diff(DA.collect(), ref, uplo) < 1e-12, ("capped cholesky", uplo))
        DB = dm.DArray.from_array(gen.randn(n * 70, 27).reshape(70, n).T.copy(), dm.Blocks(nb, nb))
        b = DB.collect()
        dm.ldiv_(F, DB, options=opts)
        check(O.backward_error_columns(A, DB.collect(), b) < 10 * n * np.finfo(float).eps,
              ("capped ldiv", uplo))


def quick():
    """A few seconds of every kernel family: the CPU suite's default group."""
    rng = np.random.default_rng(2)
    # an FP32 plan with FP64 diagonal POTRF: the untouched first diagonal
    # tile is the FP64 Cholesky of the FP32 tile, rounded once to FP32
    from dagger_mi355x.datadeps import PlanOptions
    n, nb = 160, 64
    G = rng.standard_normal((n, n))
    A32 = (G @ G.T / n + np.eye(n)).astype(np.float32)
    for uplo in ("U", "L"):
        D = dm.DArray.from_array(A32, dm.Blocks(nb, nb))
        info, ginfo = Dev(np.zeros(1, np.int32)), Dev(np.zeros(1, np.int32))
        o = PlanOptions(num_streams=2, batch=True, lookahead=1, f64_diag=True).c()
        t = D.tiled()
        call("dm_cholesky_tiled", uplo.encode(), C.byref(t), info.p, ginfo.p, C.byref(o), None)
        check(int(info.get()[0]) == 0, ("f64 diag info", uplo))
        F = D.collect().astype(np.float64)
        L0 = np.linalg.cholesky(A32[:nb, :nb].astype(np.float64))
        ref0 = L0.T if uplo == "U" else L0
        got0 = np.triu(F[:nb, :nb]) if uplo == "U" else np.tril(F[:nb, :nb])
        check(np.all(np.abs(got0 - ref0) <= (2.0 ** -24 + 1e-12) * np.abs(ref0) + 1e-30),
              ("f64 diag tile", uplo))
        ref = np.array(A32, dtype=np.float64, order="F")
        O.tile_potrf(uplo, ref)
        check(O.rel_factor_diff(F, ref, uplo) <= 5e-5, ("f64 diag factor", uplo))
    # ldiv! twice through one cached plan (same factor and right-hand-side
    # buffers, new values): the inverse tiles' other triangle, cleared once
    # when the plan is built, stays zero
    nl, nbl = 300, 64
    Al = gen.spd(nl, 31)
    DAl = dm.DArray.from_array(Al, dm.Blocks(nbl, nbl))
    Fl = dm.cholesky_(DAl)
    DBl = dm.DArray.from_array(np.zeros((nl, 4)), dm.Blocks(nbl, nbl))
    for sd in (41, 42):
        bl = gen.randn(nl * 4, sd).reshape(4, nl).T.copy()
        DBl.copy_from_host(bl)
        dm.ldiv_(Fl, DBl)
        check(O.backward_error(Al, DBl.collect(), bl) < 10 * nl * np.finfo(float).eps,
              ("ldiv through a cached plan", sd))
    # a product of 64 columns with K >= 512 in a launch of >= 64 tiles (the
    # two-deep prefetch k-loop): ragged M, an odd count of 32-deep k-steps,
    # a ragged last k-step (the other transpositions: group "tiles")
    prefetch_gemms(rng, (("T", "N", 4100, 64, 528),))
    # column dot products and scaled column updates (blocked GMRES-IR) on a
    # ragged tiling: 300 x 5 in Blocks(64, 2)
    Xh, Yh = rng.standard_normal((300, 5)), rng.standard_normal((300, 5))
    DX = dm.DArray.from_array(Xh, dm.Blocks(64, 2))
    DY = dm.DArray.from_array(Yh, dm.Blocks(64, 2))
    tx, ty = DX.tiled(), DY.tiled()
    dots = Dev(np.zeros(5))
    call("dm_tiled_col_dot", C.byref(tx), C.byref(ty), dots.p, None)
    check(np.allclose(dots.get(), (Xh * Yh).sum(0), rtol=1e-13, atol=0), "col dot")
    coefs = Dev(np.array([0.5, -2.0, 0.0, 3.0, 1.25]))
    call("dm_tiled_col_axpby", coefs.p, -1.0, C.byref(tx), 1.0, C.byref(ty), None)
    Yh = Yh - Xh * np.array([0.5, -2.0, 0.0, 3.0, 1.25])
    check(np.allclose(DY.collect(), Yh, rtol=1e-15, atol=1e-15), "col axpby")
    call("dm_tiled_col_axpby", coefs.p, 2.0, C.byref(tx), 0.0, C.byref(tx), None)  # in place
    check(np.array_equal(DX.collect(), Xh * (2.0 * np.array([0.5, -2.0, 0.0, 3.0, 1.25]))),
          "col axpby in place")
    DZ = dm.DArray.from_array(np.full((300, 5), np.nan), dm.Blocks(64, 2))
    tz = DZ.tiled()
    call("dm_tiled_col_axpby", None, 1.0, C.byref(ty), 0.0, C.byref(tz), None)
    check(np.array_equal(DZ.collect(), DY.collect()), "col axpby beta 0 ignores Y")
    # ishermitian with tiles wider than one 64 x 64 sub-block (ragged
    # 150 = 2 x 64 + 22): the pair check walks several sub-blocks per pair
    Ah = gen.spd(300, 6)
    check(dm.ishermitian(dm.DArray.from_array(Ah, dm.Blocks(150, 150))), "ishermitian wide")
    Ab = Ah.copy()
    Ab[290, 140] += 1e-3  # tile pair (1, 0), last sub-block row, ragged column
    check(not dm.ishermitian(dm.DArray.from_array(Ab, dm.Blocks(150, 150))),
          "not hermitian wide")
    Ab = Ah.copy()
    Ab[70, 71] += 1.0  # diagonal tile 0, sub-block (1, 1): exact comparison
    check(not dm.ishermitian(dm.DArray.from_array(Ab, dm.Blocks(150, 150))),
          "not hermitian wide diagonal")
    # the allocation-free TRSM: static scratch, bitwise equal to dm_dtrsm,
    # two calls in a row (the second smaller), a solve longer than one leaf
    for (side, uplo, tr, m, n) in (("L", "U", "T", 300, 20), ("R", "L", "N", 17, 70)):
        ka = m if side == "L" else n
        A = 0.1 * rng.standard_normal((ka, ka)) + np.diag(np.full(ka, 2.0))
        B = rng.standard_normal((m, n))
        dA, d1, d2 = Dev(A), Dev(B), Dev(B)
        call("dm_dtrsm", side.encode(), uplo.encode(), tr.encode(), b"N", m, n, 0.5, dA.p, ka,
             d1.p, m, None)
        call("dm_dtrsm_noalloc", side.encode(), uplo.encode(), tr.encode(), b"N", m, n, 0.5,
             dA.p, ka, d2.p, m, None)
        check(np.array_equal(d1.get(), d2.get()), ("dtrsm_noalloc", side, uplo, tr))
    # norm2 / chkfinite pass over ragged tiles, whole and triangular parts
    for dt in (np.float64, np.float32):
        A = rng.standard_normal((75, 75)).astype(dt)
        DA = dm.DArray.from_array(A, dm.Blocks(16, 16))
        A64 = A.astype(np.float64)
        for part, ref in (("full", A64), ("upper", np.triu(A64)), ("lower", np.tril(A64))):
            want = np.linalg.norm(ref)
            check(abs(dm.norm2(DA, part) - want) <= 1e-13 * want, ("norm2", part, dt))
        check(dm.chkfinite(DA), "chkfinite")
        # large tiles: the four-way unrolled loop of the whole-tile path
        L = rng.standard_normal((600, 500)).astype(dt)
        DL = dm.DArray.from_array(L, dm.Blocks(300, 300))
        want = np.linalg.norm(L.astype(np.float64))
        check(abs(dm.norm2(DL) - want) <= 1e-13 * want, ("norm2 large tiles", dt))
        L[599, 499] = np.inf
        DL.copy_from_host(L)
        try:
            dm.chkfinite(DL)
            check(False, "chkfinite missed an Inf")
        except ValueError:
            pass
        A[74, 74] = np.nan
        DA.copy_from_host(A)
        try:
            dm.chkfinite(DA)
            check(False, "chkfinite missed a NaN")
        except ValueError:
            pass
    # F \\ B staged in the factor's blocking; triangular ldiv! (vector and
    # matrix right-hand sides, unit diagonal, B re-blocked and copied back)
    A = gen.spd(70, 9)
    DA = dm.DArray.from_array(A, dm.Blocks(32, 32))
    F = dm.cholesky_(DA, uplo="L")
    for B in (rng.standard_normal(70), rng.standard_normal((70, 3))):
        X = dm.ldiv(F, B).collect()
        check(relerr(X, np.linalg.solve(A, B)) < 1e-11, ("backslash", B.ndim))
    T = 0.1 * rng.standard_normal((70, 70)) + np.diag(np.full(70, 2.0))
    DT = dm.DArray.from_array(T, dm.Blocks(32, 32))
    for uplo in ("U", "L"):
        Tu = np.triu(T, 1) if uplo == "U" else np.tril(T, -1)
        Tu = Tu + np.eye(70)
        b = rng.standard_normal(70)
        Db = dm.DArray.from_array(b, dm.Blocks(32))
        dm.ldiv_triangular_(uplo, "U", DT, Db)
        check(relerr(Db.collect(), np.linalg.solve(Tu, b)) < 1e-12, ("trsv unit", uplo))
        Bm = rng.standard_normal((70, 5))
        DBm = dm.DArray.from_array(Bm, dm.Blocks(20, 2))
        dm.ldiv_triangular_(uplo, "U", DT, DBm)
        check(relerr(DBm.collect(), np.linalg.solve(Tu, Bm)) < 1e-12, ("trsm unit", uplo))
    # mul! on DMatrices of unequal blockings (re-blocked to the smallest)
    Am, Bm, Cm = (rng.standard_normal(s) for s in ((50, 37), (37, 29), (50, 29)))
    DCm = dm.DArray.from_array(Cm, dm.Blocks(24, 24))
    dm.mul_(DCm, dm.DArray.from_array(Am, dm.Blocks(16, 16)),
            dm.DArray.from_array(Bm, dm.Blocks(20, 20)), 2.0, -1.0)
    check(relerr(DCm.collect(), 2.0 * Am @ Bm - Cm) < 1e-13, "mul_")
    # matrix-vector mul! (GEMV tile tasks), x blocked unlike A's columns
    Av, xv, yv = rng.standard_normal((45, 38)), rng.standard_normal(38), rng.standard_normal(45)
    Dy = dm.DArray.from_array(yv, dm.Blocks(16))
    dm.mul_(Dy, dm.DArray.from_array(Av, dm.Blocks(16, 12)), dm.DArray.from_array(xv, dm.Blocks(10)),
            1.5, 0.5)
    check(relerr(Dy.collect(), 1.5 * Av @ xv + 0.5 * yv) < 1e-13, "mul_ vector")
    Dz = dm.mul(dm.DArray.from_array(Av, dm.Blocks(16, 12)), dm.DArray.from_array(yv, dm.Blocks(9)),
                "T")
    check(relerr(Dz.collect(), Av.T @ yv) < 1e-13, "mul vector T")
    for ta, tb, (m, n, k) in [("T", "N", (70, 40, 33)), ("N", "T", (5, 9, 3))]:
        A = rng.standard_normal((k, m) if ta == "T" else (m, k))
        B = rng.standard_normal((n, k) if tb == "T" else (k, n))
        Cm = rng.standard_normal((m, n))
        dA, dB, dC = Dev(A), Dev(B), Dev(Cm)
        call("dm_dgemm", ta.encode(), tb.encode(), m, n, k, -1.0, dA.p, A.shape[0], dB.p,
             B.shape[0], 0.5, dC.p, m, None)
        ref = 0.5 * Cm - (A.T if ta == "T" else A) @ (B.T if tb == "T" else B)
        check(relerr(dC.get(), ref) < 1e-13, ("dgemm", ta, tb))
    F, _ = O.cholesky(gen.spd(70, 5), 70, "U")
    B = rng.standard_normal((70, 20))
    dA, dB = Dev(F), Dev(B)
    call("dm_dtrsm", b"L", b"U", b"T", b"N", 70, 20, 1.0, dA.p, 70, dB.p, 70, None)
    ref = B.copy(order="F")
    O.tile_trsm("L", "U", "T", "N", 1.0, np.asfortranarray(F), ref)
    check(relerr(dB.get(), ref) < 1e-12, "dtrsm")
    for uplo in "UL":
        n = 130
        A = gen.spd(n, 900 + n)
        dA = Dev(A)
        info = dm.DeviceInt(1, value=-7)
        call("dm_dpotrf", uplo.encode(), n, dA.p, n, C.c_void_p(info.ptr), None)
        ref, _ = O.cholesky(A, n, uplo)
        check(int(info.get()[0]) == 0 and O.rel_factor_diff(dA.get(), ref, uplo) < 1e-12,
              ("dpotrf", uplo))
    A = gen.spd(100, 5)
    A[70, 70] = -1.0
    dA, info = Dev(A), dm.DeviceInt(1)
    call("dm_dpotrf", b"L", 100, dA.p, 100, C.c_void_p(info.ptr), None)
    check(int(info.get()[0]) == 71, "dpotrf info")
    A = np.zeros((36, 64), np.float32)
    Bm = np.zeros((36, 48), np.float32)
    A[:33] = rng.standard_normal((33, 64))
    Bm[:33] = rng.standard_normal((33, 48))
    Cm = rng.standard_normal((64, 48)).astype(np.float32)
    dA, dB, dC = Dev(A), Dev(Bm), Dev(Cm)
    call("dm_sgemm_bf16x", 6, b"T", b"N", 64, 48, 33, C.c_float(-1.0), dA.p, 36, dB.p, 36,
         C.c_float(1.0), dC.p, 64, None)
    ref = Cm.astype(np.float64) - A[:33].T.astype(np.float64) @ Bm[:33].astype(np.float64)
    check(relerr(dC.get().astype(np.float64), ref) < 2e-6, "bf16x6")
    n, nb = 200, 64  # ragged tiles
    A = gen.spd(n, 6)
    ref, _ = O.cholesky(A, nb, "U")
    DA = dm.DArray.from_array(A, dm.Blocks(nb, nb))
    Fc = dm.cholesky_(DA, options=dm.PlanOptions(num_streams=2, batch=True, lookahead=1,
                                                  use_graph=False))
    check(O.rel_factor_diff(DA.collect(), ref, "U") < 1e-12, "tiled cholesky")
    b = gen.randn(n * 3, 8).reshape(3, n).T.copy()
    DB = dm.DArray.from_array(b, dm.Blocks(nb, nb))
    dm.ldiv_(Fc, DB)
    check(O.backward_error(A, DB.collect(), b) < 10 * n * np.finfo(float).eps, "ldiv")
    for (m, nb, nrhs) in [(300, 128, 9), (256, 128, 40)]:  # streaming residual kernel
        Am = rng.standard_normal((m, m))
        X = rng.standard_normal((m, nrhs))
        Y = rng.standard_normal((m, nrhs))
        DA = dm.DArray.from_array(Am, dm.Blocks(nb, nb))
        DX = dm.DArray.from_array(X, dm.Blocks(nb, nb))
        DY = dm.DArray.from_array(Y, dm.Blocks(nb, nb))
        ta, tx, ty = DA.tiled(), DX.tiled(), DY.tiled()
        o = dm.PlanOptions().c()
        call("dm_gemm_tiled", b"N", b"N", -1.0, C.byref(ta), C.byref(tx), 0.5, C.byref(ty),
             C.byref(o), None)
        check(relerr(DY.collect(), 0.5 * Y - Am @ X) < 1e-13, ("gemm block column", m, nrhs))
    from dagger_mi355x import batched as Bt
    bs = Bt.BatchedSystems([40, 70, 33], [2, 1, 3], np.float64).generate(seed=100)
    A0s = [bs.get_system(m) for m in range(bs.count)]
    bs.posv(panels={})
    check(np.all(bs.infos() == 0), "batched info")
    for m in range(bs.count):
        A0, B0 = A0s[m]
        _, X = bs.get_system(m)
        check(O.backward_error(A0, X, B0) <= 10 * A0.shape[0] * np.finfo(float).eps,
              ("batched", m))


def complex_ops():
    """Complex tile ops (ComplexF64 / ComplexF32) on the emulation build:
    every op/side/uplo/trans form at shapes that cross the 64-wide leaves,
    against numpy; then a small tiled complex Cholesky and its solve."""
    rng = np.random.default_rng(9)

    def crand(*shape, dt=np.complex128):
        return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dt)

    def op(X, t):
        return X if t == "N" else (X.T if t == "T" else X.conj().T)

    for dt, pre, tol in ((np.complex128, "z", 1e-12), (np.complex64, "c", 1e-4)):
        R = np.float64 if dt == np.complex128 else np.float32
        al, be = np.array([0.5, -0.25], R), np.array([1.5, 0.5], R)
        a, b = complex(0.5, -0.25), complex(1.5, 0.5)
        for ta, tb in (("N", "N"), ("C", "N"), ("N", "C"), ("T", "C")):
            m, n, k = 70, 45, 37
            A = crand(*((m, k) if ta == "N" else (k, m)), dt=dt)
            B = crand(*((k, n) if tb == "N" else (n, k)), dt=dt)
            Cm = crand(m, n, dt=dt)
            dA, dB, dC = Dev(A), Dev(B), Dev(Cm)
            call(f"dm_{pre}gemm", ta.encode(), tb.encode(), m, n, k, al.ctypes.data_as(C.c_void_p),
                 dA.p, A.shape[0], dB.p, B.shape[0], be.ctypes.data_as(C.c_void_p), dC.p, m, None)
            ref = a * op(A, ta).astype(np.complex128) @ op(B, tb) + b * Cm
            check(relerr(dC.get(), ref) <= tol, (pre, "gemm", ta, tb))
        for uplo, tr in (("U", "C"), ("L", "N")):
            n, k = 90, 33
            A = crand(*((k, n) if tr == "C" else (n, k)), dt=dt)
            Cm = crand(n, n, dt=dt)
            dA, dC = Dev(A), Dev(Cm)
            fn = C.c_double if dt == np.complex128 else C.c_float
            call(f"dm_{pre}herk", uplo.encode(), tr.encode(), n, k, fn(-1.0), dA.p, A.shape[0],
                 fn(1.0), dC.p, n, None)
            P = op(A, tr).astype(np.complex128)
            full = Cm - P @ P.conj().T
            tri = np.triu if uplo == "U" else np.tril
            got = dC.get()
            want = tri(full) + (np.tril(Cm, -1) if uplo == "U" else np.triu(Cm, 1))
            np.fill_diagonal(want, np.diag(want).real)  # herk: real diagonal
            check(relerr(got, want) <= tol, (pre, "herk", uplo, tr))
        for side in "LR":
            for uplo in "UL":
                for tr in "NTC":
                    m, n = (100, 7) if side == "L" else (9, 100)
                    ka = m if side == "L" else n
                    T = crand(ka, ka, dt=dt) * 0.05 + np.diag(np.full(ka, 2.0 + 0.5j)).astype(dt)
                    T = np.triu(T) if uplo == "U" else np.tril(T)
                    B = crand(m, n, dt=dt)
                    dT, dB = Dev(T), Dev(B)
                    call(f"dm_{pre}trsm", side.encode(), uplo.encode(), tr.encode(), b"N", m, n,
                         al.ctypes.data_as(C.c_void_p), dT.p, ka, dB.p, m, None)
                    X = dB.get().astype(np.complex128)
                    Tm = op(T, tr).astype(np.complex128)
                    lhs = Tm @ X if side == "L" else X @ Tm
                    check(relerr(lhs, a * B) <= tol, (pre, "trsm", side, uplo, tr))
        for uplo in "UL":
            n = 100
            G = crand(n, n, dt=np.complex128)
            A = (G @ G.conj().T / n + np.eye(n)).astype(dt)
            dA = Dev(A)
            info = dm.DeviceInt(1)
            call(f"dm_{pre}potrf", uplo.encode(), n, dA.p, n, C.c_void_p(info.ptr), None)
            check(int(info.get()[0]) == 0, (pre, "potrf info"))
            L = np.linalg.cholesky(A.astype(np.complex128))
            ref = L.conj().T if uplo == "U" else L
            tri = np.triu if uplo == "U" else np.tril
            check(relerr(tri(dA.get()), ref) <= tol, (pre, "potrf", uplo))
            Bad = A.copy()
            Bad[70, 70] = -1.0  # first failure at column 71
            dB = Dev(Bad)
            call(f"dm_{pre}potrf", uplo.encode(), n, dB.p, n, C.c_void_p(info.ptr), None)
            check(int(info.get()[0]) == 71, (pre, "potrf failed column", int(info.get()[0])))
    # tiled complex Cholesky (ragged tiles) and ldiv!, native and datadeps
    n, nb = 150, 48
    G = crand(n, n)
    A = G @ G.conj().T / n + np.eye(n)
    A = 0.5 * (A + A.conj().T)  # exactly Hermitian (a GEMM product is not)
    X0 = crand(n, 3)
    for uplo in "UL":
        for via in ("native", "datadeps"):
            DA = dm.DArray.from_array(A, dm.Blocks(nb, nb))
            check(dm.ishermitian(DA), "complex ishermitian")
            F = dm.cholesky_(DA, uplo, via=via)
            Lref = np.linalg.cholesky(A)
            got = F.U if uplo == "U" else F.L
            ref = Lref.conj().T if uplo == "U" else Lref
            check(relerr(got, ref) <= 1e-12, ("complex chol", uplo, via))
            DB = dm.DArray.from_array(A @ X0, dm.Blocks(nb, nb))
            dm.ldiv_(F, DB, via=via)
            check(relerr(DB.collect(), X0) <= 1e-10, ("complex ldiv", uplo, via))
            xv = dm.ldiv(F, (A @ X0)[:, 0]).collect()
            check(relerr(xv, X0[:, 0]) <= 1e-10, ("complex vector ldiv", uplo, via))


GROUPS = dict(quick=quick, complex_ops=complex_ops, batched_tiled=batched_tiled, tiles=tiles, tiled=tiled, mixed=mixed, batched=batched, serving=serving,
              capped=capped)

if __name__ == "__main__":
    names = sys.argv[1:] or list(GROUPS)
    for name in names:
        t = time.time()
        GROUPS[name]()
        print(f"{name}: ok ({time.time() - t:.1f}s)", flush=True)
    print("ALL OK", flush=True)
