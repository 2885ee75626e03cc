"""Complex element types (ComplexF64, ComplexF32) on the GPU: the tile ops
against the oracle's BLAS/LAPACK (zpotrf/ztrsm/zherk/zgemm), the tiled
Cholesky with trans 'C' and herk! (src/array/cholesky.jl:19-108 for
T <: Complex) against the oracle's restatement, its solves, and the
reference's Cholesky testset for the complex types
(test/array/linalg/cholesky.jl:1-48)."""
import ctypes as C

import numpy as np
import pytest

from oracle import dagger_tiled as O

pytestmark = pytest.mark.gpu

EPS = {np.complex128: np.finfo(np.float64).eps, np.complex64: np.finfo(np.float32).eps}
PRE = {np.complex128: "z", np.complex64: "c"}
REAL = {np.complex128: np.float64, np.complex64: np.float32}


def crand(rng, *shape, dt=np.complex128):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dt)


def hermitian_pd(rng, n, dt=np.complex128, shift=None):
    G = crand(rng, n, n)
    A = G @ G.conj().T / n + (np.eye(n) if shift is None else shift * np.eye(n))
    return (0.5 * (A + A.conj().T)).astype(dt)  # exactly Hermitian


def isapprox(x, y, rtol):
    return np.linalg.norm(x - y) <= rtol * max(np.linalg.norm(x), np.linalg.norm(y))


def relerr(x, ref):
    return float(np.linalg.norm(x - ref) / np.linalg.norm(ref))


class Dev:
    def __init__(self, dm, a):
        self.dm = dm
        self.a = np.asfortranarray(a)
        self.buf = dm.DeviceBuffer(max(self.a.nbytes, 16))
        dm._lib.call("dm_memcpy_h2d", C.c_void_p(self.buf.ptr),
                     self.a.ctypes.data_as(C.c_void_p), C.c_size_t(self.a.nbytes), None)

    @property
    def p(self):
        return C.c_void_p(self.buf.ptr)

    def get(self):
        out = np.empty_like(self.a, order="F")
        self.dm._lib.call("dm_memcpy_d2h", out.ctypes.data_as(C.c_void_p), C.c_void_p(self.buf.ptr),
                          C.c_size_t(out.nbytes), None)
        self.dm._lib.call("dm_stream_sync", None)
        return out


@pytest.mark.parametrize("T", [np.complex64, np.complex128])
def test_reference_cholesky_testset_complex(dm, T):
    """test/array/linalg/cholesky.jl for ComplexF32/ComplexF64: rand(Blocks(4,4),
    T, 32, 32) is not Hermitian; A*A' + n I on Blocks(32, 32) is; out of place
    leaves A; in place writes the factor into DA; L and U `≈` the dense
    factor; A[1,1] = -100 keeps A Hermitian and fails the factorization."""
    rng = np.random.default_rng(42)
    rtol = np.sqrt(EPS[T])
    D = dm.DArray.from_array((rng.random((32, 32)) + 1j * rng.random((32, 32))).astype(T),
                             dm.Blocks(4, 4))
    assert not dm.ishermitian(D)
    A = (rng.random((128, 128)) + 1j * rng.random((128, 128))).astype(np.complex128)
    A = A @ A.conj().T
    A = 0.5 * (A + A.conj().T)  # Julia's A * A' is exactly Hermitian (herk)
    A[np.diag_indices(128)] += 128
    A = A.astype(T)
    DA = dm.DArray.from_array(A, dm.Blocks(32, 32))
    assert dm.ishermitian(DA)
    U = np.linalg.cholesky(A.astype(np.complex128)).conj().T
    F = dm.cholesky(DA)
    assert np.array_equal(DA.collect(), A)
    assert isapprox(F.U, U, rtol) and isapprox(F.L, U.conj().T, rtol)
    F2 = dm.cholesky_(DA)
    assert isapprox(np.triu(DA.collect()), U, rtol) and isapprox(F2.U, U, rtol)
    A[0, 0] = -100
    DN = dm.DArray.from_array(A, dm.Blocks(32, 32))
    assert dm.ishermitian(DN)
    with pytest.raises(dm.PosDefException) as e:
        dm.cholesky(DN)
    assert e.value.info == 1


@pytest.mark.parametrize("T", [np.complex128, np.complex64])
def test_complex_tile_ops_match_oracle(dm, T):
    """dm_{z,c}gemm / herk / trsm / potrf against scipy's zgemm/zherk/ztrsm/
    zpotrf (cgemm ... for ComplexF32) on the same inputs, every op form."""
    rng = np.random.default_rng(7)
    pre, R = PRE[T], REAL[T]
    tol = 1e-12 if T == np.complex128 else 2e-5
    call = dm._lib.call
    al, be = np.array([0.75, -0.5], R), np.array([0.5, 1.25], R)
    a, b = complex(0.75, -0.5), complex(0.5, 1.25)
    for ta in "NTC":
        for tb in "NTC":
            m, n, k = 200, 130, 150
            A = crand(rng, *((m, k) if ta == "N" else (k, m)), dt=T)
            B = crand(rng, *((k, n) if tb == "N" else (n, k)), dt=T)
            Cm = crand(rng, m, n, dt=T)
            dA, dB, dC = Dev(dm, A), Dev(dm, B), Dev(dm, Cm)
            call(f"dm_{pre}gemm", ta.encode(), tb.encode(), m, n, k, al.ctypes.data_as(C.c_void_p),
                 dA.p, A.shape[0], dB.p, B.shape[0], be.ctypes.data_as(C.c_void_p), dC.p, m, None)
            ref = Cm.copy(order="F")
            O.tile_gemm(ta, tb, a, A, B, b, ref)
            assert relerr(dC.get(), ref) <= tol, ("gemm", ta, tb)
    fn = C.c_double if T == np.complex128 else C.c_float
    for uplo in "UL":
        for tr in "NC":
            n, k = 190, 70
            A = crand(rng, *((n, k) if tr == "N" else (k, n)), dt=T)
            Cm = crand(rng, n, n, dt=T)
            dA, dC = Dev(dm, A), Dev(dm, Cm)
            call(f"dm_{pre}herk", uplo.encode(), tr.encode(), n, k, fn(-1.0), dA.p, A.shape[0],
                 fn(0.5), dC.p, n, None)
            ref = Cm.copy(order="F")
            O.tile_herk(uplo, tr, -1.0, A, 0.5, ref)
            got = dC.get()
            tri = np.triu if uplo == "U" else np.tril
            assert relerr(tri(got), tri(ref)) <= tol, ("herk", uplo, tr)
            assert np.all(np.diag(got).imag == 0)  # herk zeroes the diagonal's imaginary part
            off = np.tril(got, -1) if uplo == "U" else np.triu(got, 1)
            assert np.array_equal(off, np.tril(Cm, -1) if uplo == "U" else np.triu(Cm, 1))
    for side in "LR":
        for uplo in "UL":
            for tr in "NTC":
                for diag in "NU":
                    m, n = (260, 33) if side == "L" else (41, 260)
                    ka = m if side == "L" else n
                    Tm = 0.02 * crand(rng, ka, ka, dt=T) + np.diag(
                        rng.uniform(1, 2, ka) + 0.5j * rng.uniform(-1, 1, ka)).astype(T)
                    B = crand(rng, m, n, dt=T)
                    dT, dB = Dev(dm, Tm), Dev(dm, B)
                    call(f"dm_{pre}trsm", side.encode(), uplo.encode(), tr.encode(),
                         diag.encode(), m, n, al.ctypes.data_as(C.c_void_p), dT.p, ka, dB.p, m,
                         None)
                    ref = B.copy(order="F")
                    O.tile_trsm(side, uplo, tr, diag, a, Tm, ref)
                    assert relerr(dB.get(), ref) <= 50 * tol, ("trsm", side, uplo, tr, diag)
    for uplo in "UL":
        for n in (1, 63, 64, 65, 300):
            A = hermitian_pd(rng, n, T)
            dA = Dev(dm, A)
            info = dm.DeviceInt(1)
            call(f"dm_{pre}potrf", uplo.encode(), n, dA.p, n, C.c_void_p(info.ptr), None)
            ref = A.copy(order="F")
            assert O.tile_potrf(uplo, ref) == 0
            assert int(info.get()[0]) == 0
            tri = np.triu if uplo == "U" else np.tril
            got = dA.get()
            assert relerr(tri(got), tri(ref)) <= tol, ("potrf", uplo, n)
            other = np.tril(got, -1) if uplo == "U" else np.triu(got, 1)
            assert np.array_equal(other, np.tril(A, -1) if uplo == "U" else np.triu(A, 1))
        A = hermitian_pd(rng, 200, T)
        A[150, 150] = -1.0  # LAPACK info: the first failed column, 1-based
        ref = A.copy(order="F")
        want = O.tile_potrf(uplo, ref)
        dA = Dev(dm, A)
        info = dm.DeviceInt(1)
        call(f"dm_{pre}potrf", uplo.encode(), 200, dA.p, 200, C.c_void_p(info.ptr), None)
        assert int(info.get()[0]) == want == 151


@pytest.mark.parametrize("T", [np.complex128, np.complex64])
@pytest.mark.parametrize("uplo", ["U", "L"])
@pytest.mark.parametrize("n,nb", [(1000, 128), (2048, 256)])
def test_complex_tiled_cholesky_matches_oracle(dm, T, uplo, n, nb):
    """The tiled complex Cholesky (trans 'C', herk!) against the oracle's
    restatement of src/array/cholesky.jl on the same input (ragged tiles at
    n = 1000), backward error against the dense product."""
    rng = np.random.default_rng(n + nb + (0 if uplo == "U" else 1))
    A = hermitian_pd(rng, n, T)
    DA = dm.DArray.from_array(A, dm.Blocks(nb, nb))
    F = dm.cholesky_(DA, uplo)
    assert F.info == 0
    ref, info = O.cholesky(A, nb, uplo)
    assert info == 0
    got = DA.collect()
    tol = 1e-12 if T == np.complex128 else 1e-5
    assert O.rel_factor_diff(got, ref, uplo) <= tol
    Uf = F.U.astype(np.complex128)
    berr = np.linalg.norm(Uf.conj().T @ Uf - A) / np.linalg.norm(A)
    assert berr <= 10 * n * EPS[T]


@pytest.mark.parametrize("T", [np.complex128, np.complex64])
@pytest.mark.parametrize("uplo", ["U", "L"])
def test_complex_ldiv_matches_oracle(dm, T, uplo):
    """ldiv!(C::Cholesky, B) for complex factors, matrix and vector right-hand
    sides, native and datadeps forms (the two trsm! regions with 'C',
    src/array/linalg.jl:385-401) against the oracle and a dense solve."""
    n, nb = 1536, 256
    rng = np.random.default_rng(5 if uplo == "U" else 6)
    A = hermitian_pd(rng, n, T)
    DA = dm.DArray.from_array(A, dm.Blocks(nb, nb))
    F = dm.cholesky_(DA, uplo)
    Fh = DA.collect()
    B = crand(rng, n, 24, dt=T)
    tol = 1e-11 if T == np.complex128 else 2e-4
    outs = []
    for via in ("native", "datadeps"):
        DB = dm.DArray.from_array(B, dm.Blocks(nb, nb))
        dm.ldiv_(F, DB, via=via)
        outs.append(DB.collect())
    assert np.array_equal(outs[0], outs[1])  # the same tile tasks either way
    Xo = O.ldiv(Fh, uplo, B, nb)
    assert relerr(outs[0], Xo) <= tol
    Xd = np.linalg.solve(A.astype(np.complex128), B.astype(np.complex128))
    assert relerr(outs[0], Xd) <= 100 * tol
    b = B[:, 3].copy()
    xv = dm.ldiv(F, b).collect()
    assert relerr(xv, Xd[:, 3]) <= 100 * tol


@pytest.mark.parametrize("uplo", ["U", "L"])
def test_complex_schedules_are_bit_identical(dm, uplo):
    """Parallel equals sequential execution: one stream without batching
    and the two-stream look-ahead schedule give the same bits."""
    n, nb = 1280, 256
    A = hermitian_pd(np.random.default_rng(3), n)
    outs = []
    for opts in (dm.SEQUENTIAL, dm.DEFAULT_OPTIONS.replace(use_graph=False), dm.DEFAULT_OPTIONS):
        DA = dm.DArray.from_array(A, dm.Blocks(nb, nb))
        dm.cholesky_(DA, uplo, options=opts)
        outs.append(DA.collect())
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


@pytest.mark.parametrize("uplo", ["U", "L"])
def test_complex_non_pd_reports_tile_local_info(dm, uplo):
    """potrf_checked!: the failing tile's LAPACK info (tile-local, 1-based),
    as the reference fills info_arr (src/array/cholesky.jl:2-18)."""
    n, nb = 512, 128
    A = hermitian_pd(np.random.default_rng(11), n)
    A[300, 300] = -5.0  # tile 2, local column 45
    DA = dm.DArray.from_array(A, dm.Blocks(nb, nb))
    with pytest.raises(dm.PosDefException) as e:
        dm.cholesky_(DA, uplo)
    _, rinfo = O.cholesky(A, nb, uplo)
    assert e.value.info == rinfo == 300 - 2 * nb + 1
    DA = dm.DArray.from_array(A, dm.Blocks(nb, nb))
    F = dm.cholesky_(DA, uplo, check=False)
    assert F.info == rinfo and F.global_info == 301 and not F.issuccess()


def test_complex_ishermitian_rules(dm):
    """Exact on diagonal tiles (a non-zero imaginary diagonal fails), the
    conjugate transpose within sqrt(eps) on off-diagonal pairs."""
    rng = np.random.default_rng(2)
    A = hermitian_pd(rng, 96)
    assert dm.ishermitian(dm.DArray.from_array(A, dm.Blocks(32, 32)))
    B = A.copy()
    B[5, 5] += 1e-300j
    assert not dm.ishermitian(dm.DArray.from_array(B, dm.Blocks(32, 32)))
    Cm = A.copy()
    Cm[70, 10] += 1e-12 * abs(Cm[70, 10])  # within sqrt(eps) of A_ji^H
    assert dm.ishermitian(dm.DArray.from_array(Cm, dm.Blocks(32, 32)))
    Dm = A.copy()
    Dm[70, 10] = np.conj(Dm[10, 70]) + 1.0  # the pair differs
    assert not dm.ishermitian(dm.DArray.from_array(Dm, dm.Blocks(32, 32)))
    Tm = A.copy()
    Tm[70, 10] = Tm[10, 70]  # transpose without the conjugate
    assert not dm.ishermitian(dm.DArray.from_array(Tm, dm.Blocks(32, 32)))


def test_complex_lu_is_rejected(dm):
    A = hermitian_pd(np.random.default_rng(1), 128)
    with pytest.raises(dm.DMError):
        dm.lu(dm.DArray.from_array(A, dm.Blocks(64, 64)))
