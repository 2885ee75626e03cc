"""GPU parity of the BLAS/LAPACK-shaped tile entry points (through the C ABI)
against the CPU tile routines of the oracle (scipy OpenBLAS)."""
import ctypes as C

import numpy as np
import pytest

from oracle import dagger_tiled as O
from oracle import gen

pytestmark = pytest.mark.gpu


class Dev:
    """A host array mirrored in device memory (column-major)."""

    def __init__(self, dm, a):
        self.a = np.asfortranarray(a)
        self.buf = dm.DeviceBuffer(max(self.a.nbytes, 8))
        dm._lib.call("dm_memcpy_h2d", C.c_void_p(self.buf.ptr), self.a.ctypes.data_as(C.c_void_p),
                     C.c_size_t(self.a.nbytes), None)

    @property
    def p(self):
        return C.c_void_p(self.buf.ptr)

    def get(self, dm):
        out = np.empty_like(self.a, order="F")
        dm._lib.call("dm_memcpy_d2h", out.ctypes.data_as(C.c_void_p), C.c_void_p(self.buf.ptr),
                     C.c_size_t(out.nbytes), None)
        return out


def relerr(x, ref):
    return float(np.linalg.norm(x - ref) / max(np.linalg.norm(ref), 1e-300))


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
@pytest.mark.parametrize("m,n,k", [(1024, 1024, 1024), (200, 136, 77), (1, 1, 1), (64, 33, 0),
                                   (300, 520, 130)])
def test_dgemm(dm, rng, ta, tb, m, n, k):
    A = rng.standard_normal((k, m) if ta == "T" else (m, k))
    B = rng.standard_normal((n, k) if tb == "T" else (k, n))
    Cm = rng.standard_normal((m, n))
    dA, dB, dC = Dev(dm, A), Dev(dm, B), Dev(dm, Cm)
    dm._lib.call("dm_dgemm", ta.encode(), tb.encode(), m, n, k, -1.0, dA.p, max(A.shape[0], 1),
                 dB.p, max(B.shape[0], 1), 0.5, dC.p, m, None)
    got = dC.get(dm)
    ref = Cm.copy(order="F")
    if k:
        O.tile_gemm(ta, tb, -1.0, np.asfortranarray(A), np.asfortranarray(B), 0.5, ref)
    else:
        ref *= 0.5
    assert relerr(got, ref) < 1e-14


def test_dgemm_beta_zero_ignores_nan(dm, rng):
    m = n = k = 96
    A, B = rng.standard_normal((m, k)), rng.standard_normal((k, n))
    Cm = np.full((m, n), np.nan)
    dA, dB, dC = Dev(dm, A), Dev(dm, B), Dev(dm, Cm)
    dm._lib.call("dm_dgemm", b"N", b"N", m, n, k, 1.0, dA.p, m, dB.p, k, 0.0, dC.p, m, None)
    assert relerr(dC.get(dm), A @ B) < 1e-14


def test_dgemm_odd_leading_dimensions(dm, rng):
    # forces the scalar-load path (odd ld, odd extents)
    m, n, k, ld = 101, 67, 55, 103
    A = rng.standard_normal((ld, k))
    B = rng.standard_normal((ld, n))
    Cm = np.zeros((ld, n))
    dA, dB, dC = Dev(dm, A), Dev(dm, B), Dev(dm, Cm)
    dm._lib.call("dm_dgemm", b"N", b"N", m, n, k, 1.0, dA.p, ld, dB.p, ld, 0.0, dC.p, ld, None)
    got = dC.get(dm)
    assert relerr(got[:m], A[:m] @ B[:k]) < 1e-14
    assert np.all(got[m:] == 0)


@pytest.mark.parametrize("uplo", ["U", "L"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("n,k", [(1024, 1024), (130, 70), (5, 3)])
def test_dsyrk(dm, rng, uplo, trans, n, k):
    A = rng.standard_normal((k, n) if trans == "T" else (n, k))
    Cm = rng.standard_normal((n, n))
    dA, dC = Dev(dm, A), Dev(dm, Cm)
    dm._lib.call("dm_dsyrk", uplo.encode(), trans.encode(), n, k, -1.0, dA.p, A.shape[0], 1.0,
                 dC.p, n, None)
    got = dC.get(dm)
    ref = Cm.copy(order="F")
    O.tile_syrk(uplo, trans, -1.0, np.asfortranarray(A), 1.0, ref)
    tri = np.triu if uplo == "U" else np.tril
    other = np.tril if uplo == "U" else np.triu
    assert relerr(tri(got), tri(ref)) < 1e-14
    # the other triangle is not touched
    assert np.array_equal(other(got, -1 if uplo == "U" else 1), other(Cm, -1 if uplo == "U" else 1))


@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("uplo", ["U", "L"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("diag", ["N", "U"])
def test_dtrsm_all_variants(dm, side, uplo, trans, diag):
    rng = np.random.default_rng(sum(map(ord, side + uplo + trans + diag)))
    m, n = (300, 70) if side == "L" else (70, 300)
    ka = m if side == "L" else n
    A = 0.05 * rng.standard_normal((ka, ka)) + np.diag(np.full(ka, 2.0))
    B = rng.standard_normal((m, n))
    dA, dB = Dev(dm, A), Dev(dm, B)
    dm._lib.call("dm_dtrsm", side.encode(), uplo.encode(), trans.encode(), diag.encode(), m, n,
                 0.75, dA.p, ka, dB.p, m, None)
    ref = B.copy(order="F")
    O.tile_trsm(side, uplo, trans, diag, 0.75, np.asfortranarray(A), ref)
    assert relerr(dB.get(dm), ref) < 1e-13


@pytest.mark.parametrize("side,uplo,trans", [("L", "U", "T"), ("R", "L", "T"), ("L", "L", "N"),
                                             ("L", "U", "N")])
@pytest.mark.parametrize("ka,nv", [(1024, 1024), (2500, 40), (33, 5)])
def test_dtrsm_cholesky_shapes(dm, side, uplo, trans, ka, nv):
    """The trsm shapes of the tiled Cholesky / ldiv!, on a real factor,
    including lengths above the LDS-resident limit (host recursion)."""
    A = gen.spd(ka, ka)
    F, info = O.cholesky(A, max(ka, 1), uplo)
    assert info == 0
    rng = np.random.default_rng(ka + nv)
    m, n = (ka, nv) if side == "L" else (nv, ka)
    B = rng.standard_normal((m, n))
    dA, dB = Dev(dm, F), Dev(dm, B)
    dm._lib.call("dm_dtrsm", side.encode(), uplo.encode(), trans.encode(), b"N", m, n, 1.0, dA.p,
                 ka, dB.p, m, None)
    ref = B.copy(order="F")
    O.tile_trsm(side, uplo, trans, "N", 1.0, np.asfortranarray(F), ref)
    assert relerr(dB.get(dm), ref) < 1e-13


@pytest.mark.parametrize("uplo", ["U", "L"])
@pytest.mark.parametrize("n", [1, 64, 100, 512, 1024])
def test_dpotrf(dm, uplo, n):
    A = gen.spd(n, 900 + n)
    dA = Dev(dm, A)
    info = dm.DeviceInt(1, value=-7)
    dm._lib.call("dm_dpotrf", uplo.encode(), n, dA.p, n, C.c_void_p(info.ptr), None)
    got = dA.get(dm)
    assert int(info.get()[0]) == 0
    ref, rinfo = O.cholesky(A, n, uplo)  # one tile = dense LAPACK potrf
    assert rinfo == 0
    assert O.rel_factor_diff(got, ref, uplo) < 1e-13
    other = np.tril if uplo == "U" else np.triu
    off = -1 if uplo == "U" else 1
    assert np.array_equal(other(got, off), other(A, off))


@pytest.mark.parametrize("n,bad", [(64, 0), (300, 130), (1024, 1000)])
def test_dpotrf_info_on_non_pd(dm, n, bad):
    A = gen.spd(n, 5)
    A[bad, bad] = -1.0
    dA = Dev(dm, A)
    info = dm.DeviceInt(1)
    dm._lib.call("dm_dpotrf", b"U", n, dA.p, n, C.c_void_p(info.ptr), None)
    _, rinfo = O.cholesky(A, n, "U")
    assert rinfo == bad + 1
    assert int(info.get()[0]) == rinfo


def test_dpotrf_rejects_bad_arguments(dm):
    info = dm.DeviceInt(1)
    with pytest.raises(dm.DMError) as e:
        dm._lib.call("dm_dpotrf", b"X", 4, C.c_void_p(0), 4, C.c_void_p(info.ptr), None)
    assert e.value.status == -1
    with pytest.raises(dm.DMError) as e:
        dm._lib.call("dm_dpotrf", b"U", 8, C.c_void_p(0), 4, C.c_void_p(info.ptr), None)
    assert e.value.status == -4


@pytest.mark.parametrize("trans", ["N", "T"])
def test_dgemv(dm, rng, trans):
    m, n = 1000, 300
    A = rng.standard_normal((m, n))
    x = rng.standard_normal(n if trans == "N" else m)
    y = rng.standard_normal(m if trans == "N" else n)
    dA, dx, dy = Dev(dm, A), Dev(dm, x), Dev(dm, y)
    dm._lib.call("dm_dgemv", trans.encode(), m, n, -1.0, dA.p, m, dx.p, 1, 2.0, dy.p, 1, None)
    ref = y.copy()
    O.tile_gemv(trans, -1.0, A, x, 2.0, ref)
    assert relerr(dy.get(dm), ref) < 1e-14


@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("beta", [0.0, -1.5])
def test_dgemv_empty_inner_dimension_scales_y(dm, rng, trans, beta):
    """BLAS gemv with an empty inner dimension (n = 0 for 'N', m = 0 for
    'T') still sets y := beta y; beta = 0 writes zeros without reading y
    (a NaN there must not survive)."""
    m, n = (300, 0) if trans == "N" else (0, 300)
    y = rng.standard_normal(300)
    y[7] = np.nan
    A = np.zeros((max(m, 1), 1))
    dA, dx, dy = Dev(dm, A), Dev(dm, np.zeros(1)), Dev(dm, y)
    dm._lib.call("dm_dgemv", trans.encode(), m, n, 2.0, dA.p, max(m, 1), dx.p, 1, beta, dy.p, 1,
                 None)
    got = dy.get(dm)
    if beta == 0.0:
        assert np.array_equal(got, np.zeros(300))
    else:
        ref = beta * y
        assert np.array_equal(got[~np.isnan(ref)], ref[~np.isnan(ref)])
        assert np.isnan(got[7])


@pytest.mark.parametrize("uplo", ["U", "L"])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_dtrsv(dm, uplo, trans):
    n = 700
    A = gen.spd(n, 3)
    F, _ = O.cholesky(A, n, uplo)
    x = gen.randn(n, 4)
    dA, dx = Dev(dm, F), Dev(dm, x)
    dm._lib.call("dm_dtrsv", uplo.encode(), trans.encode(), b"N", n, dA.p, n, dx.p, 1, None)
    ref = x.copy()
    O.tile_trsv(uplo, trans, "N", np.asfortranarray(F), ref)
    assert relerr(dx.get(dm), ref) < 1e-13


@pytest.mark.parametrize("ta,tb", [("T", "N"), ("N", "T")])
def test_sgemm(dm, rng, ta, tb):
    m = n = k = 512
    A = rng.standard_normal((m, k)).astype(np.float32)
    B = rng.standard_normal((k, n)).astype(np.float32)
    Cm = rng.standard_normal((m, n)).astype(np.float32)
    dA, dB, dC = Dev(dm, A), Dev(dm, B), Dev(dm, Cm)
    dm._lib.call("dm_sgemm", ta.encode(), tb.encode(), m, n, k, C.c_float(1.0), dA.p, m, dB.p, k,
                 C.c_float(1.0), dC.p, m, None)
    opA = A.T if ta == "T" else A
    opB = B.T if tb == "T" else B
    ref = opA.astype(np.float64) @ opB.astype(np.float64) + Cm
    assert relerr(dC.get(dm).astype(np.float64), ref) < 1e-5


@pytest.mark.parametrize("terms,tol", [(6, 2e-6), (3, 1e-4)])
@pytest.mark.parametrize("m,n,k,ld", [(512, 512, 512, 512), (200, 136, 77, 80), (129, 1, 33, 36),
                                      (1024, 384, 1024, 1024)])
def test_sgemm_bf16x(dm, rng, terms, tol, m, n, k, ld):
    """FP32 GEMM on the BF16 matrix cores (operand splitting) against the
    FP64 product of the same FP32 inputs: 3-term splitting (6 products) is
    FP32-accurate, 2-term splitting (3 products) carries ~16 bits."""
    A = np.zeros((ld, m), np.float32)
    B = np.zeros((ld, n), np.float32)
    A[:k] = rng.standard_normal((k, m))
    B[:k] = rng.standard_normal((k, n))
    Cm = rng.standard_normal((m, n)).astype(np.float32)
    dA, dB, dC = Dev(dm, A), Dev(dm, B), Dev(dm, Cm)
    dm._lib.call("dm_sgemm_bf16x", terms, b"T", b"N", m, n, k, C.c_float(-1.0), dA.p, ld, dB.p,
                 ld, C.c_float(1.0), dC.p, m, None)
    ref = Cm.astype(np.float64) - A[:k].T.astype(np.float64) @ B[:k].astype(np.float64)
    assert relerr(dC.get(dm).astype(np.float64), ref) < tol * np.sqrt(k / 512 + 1)


def test_sgemm_bf16x_rejects_other_shapes(dm):
    A = Dev(dm, np.zeros((64, 64), np.float32))
    with pytest.raises(dm.DMError):
        dm._lib.call("dm_sgemm_bf16x", 6, b"N", b"N", 64, 64, 64, C.c_float(1.0), A.p, 64, A.p,
                     64, C.c_float(0.0), A.p, 64, None)
    with pytest.raises(dm.DMError):
        dm._lib.call("dm_sgemm_bf16x", 4, b"T", b"N", 64, 64, 64, C.c_float(1.0), A.p, 64, A.p,
                     64, C.c_float(0.0), A.p, 64, None)
    with pytest.raises(dm.DMError):  # lda not a multiple of 4
        dm._lib.call("dm_sgemm_bf16x", 6, b"T", b"N", 63, 63, 63, C.c_float(1.0), A.p, 63, A.p,
                     63, C.c_float(0.0), A.p, 63, None)


def test_randn_matches_host_generator(dm):
    n = 100000
    buf = dm.DeviceBuffer(8 * n)
    dm._lib.call("dm_fill_randn_d", C.c_void_p(buf.ptr), C.c_size_t(n), C.c_uint64(42),
                 C.c_uint64(5), None)
    out = np.empty(n)
    dm._lib.call("dm_memcpy_d2h", out.ctypes.data_as(C.c_void_p), C.c_void_p(buf.ptr),
                 C.c_size_t(8 * n), None)
    ref = gen.randn(n, 42, offset=5)
    # same Philox integers; log/cos may differ in the last ulps between libms
    assert np.max(np.abs(out - ref)) < 1e-13
