"""GPU parity of `F \\ B` (staged in the factor's blocking) and of the
triangular `ldiv!` methods against the oracle's restatement of
src/array/linalg.jl:208-246, :310-335 and :352-405 (trsm!/trsv! regions of
src/array/trsm.jl).

Tolerances: the GPU and the oracle run the same tile DAG with different
FP64 kernels (different summation orders), so solutions agree to a relative
1e-12 in norm on these well-conditioned systems; backward errors stay below
10 n eps."""
import numpy as np
import pytest

from oracle import dagger_tiled as O
from oracle import gen

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps


def rel(x, ref):
    return np.linalg.norm(x - ref) / np.linalg.norm(ref)


def isapprox(x, y, rtol=np.sqrt(EPS)):
    """Julia's `x ≈ y` for arrays: norm(x - y) <= rtol * max(norm(x), norm(y)),
    rtol defaulting to sqrt(eps)."""
    return np.linalg.norm(x - y) <= rtol * max(np.linalg.norm(x), np.linalg.norm(y))


def factor(dm, n, nb, uplo, seed):
    A = gen.spd(n, seed)
    DA = dm.DArray.from_array(A, dm.Blocks(nb, nb))
    return A, DA, dm.cholesky_(DA, uplo=uplo)


@pytest.mark.parametrize("uplo", ["U", "L"])
@pytest.mark.parametrize("n,nb,nrhs", [(1000, 128, 0), (1000, 128, 7), (768, 256, 32)])
def test_backslash_host_rhs(dm, uplo, n, nb, nrhs):
    A, DA, F = factor(dm, n, nb, uplo, 40 + n)
    B = gen.randn(n * max(nrhs, 1), 41).reshape((n,) if nrhs == 0 else (nrhs, n)).T.copy()
    B0 = B.copy()
    X = dm.ldiv(F, B)
    assert X.partitioning == (dm.Blocks(nb) if nrhs == 0 else dm.Blocks(nb, nb))
    x = X.collect()
    assert np.array_equal(B, B0)
    ref = O.ldiv(DA.collect(), uplo, B, nb)
    assert rel(x, ref) <= 1e-12
    assert O.backward_error(A, x, B) <= 10 * n * EPS


def test_backslash_darray_rhs_other_blocking(dm):
    # B blocked differently from the factor: staged by one re-block of B, the
    # factor is not copied, and B itself is left unchanged
    n, nb = 900, 256
    A, DA, F = factor(dm, n, nb, "U", 50)
    B = gen.randn(n * 5, 51).reshape(5, n).T.copy()
    DB = dm.DArray.from_array(B, dm.Blocks(100, 3))
    base = F.factors.buffer.ptr
    X = dm.ldiv(F, DB)
    assert F.factors.buffer.ptr == base
    assert X.partitioning == dm.Blocks(nb, nb)
    assert np.array_equal(DB.collect(), B)
    assert rel(X.collect(), O.ldiv(DA.collect(), "U", B, nb)) <= 1e-12
    with pytest.raises(ValueError):
        dm.ldiv(F, np.zeros(n + 1))


def tri_system(n, uplo, diag, seed):
    rng = np.random.default_rng(seed)
    A = 0.5 * rng.standard_normal((n, n)) / np.sqrt(n) + np.diag(np.full(n, 2.0))
    if diag == "U":
        # stored diagonal deliberately not 1: a unit solve must ignore it
        A[np.diag_indices(n)] = 3.0
    T = np.triu(A) if uplo == "U" else np.tril(A)
    if diag == "U":
        T[np.diag_indices(n)] = 1.0
    return A, T


@pytest.mark.parametrize("uplo", ["U", "L"])
@pytest.mark.parametrize("diag", ["N", "U"])
@pytest.mark.parametrize("vector", [False, True])
def test_triangular_ldiv(dm, uplo, diag, vector):
    n, nb = 700, 128
    A, T = tri_system(n, uplo, diag, 60 + 2 * vector + (uplo == "L"))
    B = gen.randn(n if vector else n * 9, 61).reshape((n,) if vector else (9, n)).T.copy()
    DA = dm.DArray.from_array(A, dm.Blocks(nb, nb))
    DB = dm.DArray.from_array(B, dm.Blocks(nb) if vector else dm.Blocks(nb, nb))
    dm.ldiv_triangular_(uplo, diag, DA, DB)
    x = DB.collect()
    if vector:
        ref = O.trsv(uplo, "N", diag, A, B, nb)
    else:
        ref = O.trsm("L", uplo, "N", diag, 1.0, A, B, nb)
    assert rel(x, ref) <= 1e-12
    assert rel(x, np.linalg.solve(T, B)) <= 1e-11


def test_triangular_ldiv_reblocks_rhs(dm):
    # B blocked unlike A: re-blocked to A's block size and copied back
    n = 500
    A, T = tri_system(n, "L", "N", 70)
    B = gen.randn(n * 4, 71).reshape(4, n).T.copy()
    DA = dm.DArray.from_array(A, dm.Blocks(128, 128))
    DB = dm.DArray.from_array(B, dm.Blocks(77, 4))
    dm.ldiv_triangular_("L", "N", DA, DB)
    assert DB.partitioning == dm.Blocks(77, 4)
    assert rel(DB.collect(), O.trsm("L", "L", "N", "N", 1.0, A, B, 128)) <= 1e-12


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
@pytest.mark.parametrize("m,n,k,bs", [(700, 450, 380, (128, 128, 128)),
                                      (512, 512, 1024, (256, 256, 256)),
                                      (300, 200, 100, (96, 64, 128))])
def test_mul_matches_oracle(dm, ta, tb, m, n, k, bs):
    # bs = (A's, B's, C's block size); unequal sizes are re-blocked to the
    # smallest of A's and B's, C is copied back into its own blocking
    rng = np.random.default_rng(m + n + k)
    A = rng.standard_normal((k, m) if ta == "T" else (m, k))
    B = rng.standard_normal((n, k) if tb == "T" else (k, n))
    Cm = rng.standard_normal((m, n))
    DA = dm.DArray.from_array(A, dm.Blocks(bs[0], bs[0]))
    DB = dm.DArray.from_array(B, dm.Blocks(bs[1], bs[1]))
    DC = dm.DArray.from_array(Cm, dm.Blocks(bs[2], bs[2]))
    dm.mul_(DC, DA, DB, -0.75, 0.5, ta, tb)
    assert DC.partitioning == dm.Blocks(bs[2], bs[2])
    ref = O.gemm(ta, tb, -0.75, A, B, 0.5, Cm, min(bs[:2]))
    assert rel(DC.collect(), ref) <= 1e-13


def test_mul_same_operand_and_errors(dm):
    # A * A^T with A passed twice (the syrk_dagger! case of the reference)
    rng = np.random.default_rng(5)
    A = rng.standard_normal((400, 300))
    DA = dm.DArray.from_array(A, dm.Blocks(128, 128))
    DC = dm.zeros(dm.Blocks(128, 128), np.float64, 400, 400)
    dm.mul_(DC, DA, DA, 1.0, 0.0, "N", "T")
    got = DC.collect()
    assert rel(got, A @ A.T) <= 1e-13
    with pytest.raises(ValueError):
        dm.mul_(DC, DA, DA, 1.0, 0.0, "N", "N")


def test_mul_with_without_transpose(dm):
    # the reference's matmul testset: X'X, XX', XX on Blocks(10, 20), 40x40,
    # ||collect(Xk) - ref|| < 1e-12
    X = np.random.default_rng(40).random((40, 40))
    X1 = dm.DArray.from_array(X, dm.Blocks(10, 20))
    assert np.linalg.norm(dm.mul(X1, X1, "T", "N").collect() - X.T @ X) < 1e-12
    assert np.linalg.norm(dm.mul(X1, X1, "N", "T").collect() - X @ X.T) < 1e-12
    assert np.linalg.norm(dm.mul(X1, X1).collect() - X @ X) < 1e-12
    assert np.array_equal(X1.collect(), X)


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T")])
def test_mul_float32(dm, ta, tb):
    # FP32 tile products (sum of k = 600 products): relative 1e-5 in norm
    rng = np.random.default_rng(7)
    A = rng.standard_normal((600, 520)).astype(np.float32)
    B = rng.standard_normal((600, 480) if ta == "T" else (520, 480)).astype(np.float32)
    if tb == "T":
        B = np.ascontiguousarray(B.T)
    DA = dm.DArray.from_array(A, dm.Blocks(128, 128))
    DB = dm.DArray.from_array(B, dm.Blocks(128, 128))
    got = dm.mul(DA, DB, ta, tb).collect()
    A64, B64 = A.astype(np.float64), B.astype(np.float64)
    ref = (A64.T if ta == "T" else A64) @ (B64.T if tb == "T" else B64)
    assert got.dtype == np.float32
    assert rel(got.astype(np.float64), ref) <= 1e-5


@pytest.mark.parametrize("bs_i", [128, 64])
def test_reference_solve_testset_cholesky(dm, bs_i):
    # test/array/linalg/solve.jl's Cholesky part: N = 128, Blocks(bs_i, 128)
    # for A and B (non-square blocks for bs_i = 64), A = A'A + I from rand,
    # ldiv! with a DMatrix (`≈`) and a DVector (`≈ rtol=1e-10`) right-hand side
    N = 128
    rng = np.random.default_rng(bs_i)
    A = rng.random((N, N))
    A = A.T @ A + np.eye(N)
    B = rng.random((N, N))
    b = rng.random(N)
    DA = dm.DArray.from_array(A, dm.Blocks(bs_i, N))
    F = dm.cholesky(DA)
    assert np.array_equal(DA.collect(), A)
    DB = dm.DArray.from_array(B, dm.Blocks(bs_i, N))
    dm.ldiv_(F, DB)
    assert DB.partitioning == dm.Blocks(bs_i, N)
    X = DB.collect()
    assert isapprox(X, np.linalg.solve(A, B))
    Db = dm.DArray.from_array(b, dm.Blocks(bs_i))
    dm.ldiv_(F, Db)
    assert isapprox(Db.collect(), np.linalg.solve(A, b), rtol=1e-10)


@pytest.mark.parametrize("T,rtol", [(np.float32, np.sqrt(np.finfo(np.float32).eps)),
                                    (np.float64, np.sqrt(EPS))])
def test_reference_cholesky_testset(dm, T, rtol):
    # test/array/linalg/cholesky.jl for the real types: rand(Blocks(4,4), T,
    # 32, 32) is not Hermitian; A*A' + n I on Blocks(32, 32) is; out of place
    # leaves A; in place writes the factor into DA; L and U `≈` the dense
    # factor; a negative A[1,1] keeps A Hermitian and fails the factorization
    rng = np.random.default_rng(128)
    D = dm.DArray.from_array(rng.random((32, 32)).astype(T), dm.Blocks(4, 4))
    assert not dm.ishermitian(D)
    A = rng.random((128, 128)).astype(T)
    A = (A @ A.T).astype(T)
    A[np.diag_indices(128)] += 128
    DA = dm.DArray.from_array(A, dm.Blocks(32, 32))
    assert dm.ishermitian(DA)
    U = np.linalg.cholesky(A.astype(np.float64)).T
    F = dm.cholesky(DA)
    assert np.array_equal(DA.collect(), A)
    assert isapprox(F.U, U, rtol) and isapprox(F.L, U.T, rtol)
    F2 = dm.cholesky_(DA)
    assert isapprox(np.triu(DA.collect()), U, rtol) and isapprox(F2.U, U, rtol)
    A[0, 0] = -100
    DN = dm.DArray.from_array(A, dm.Blocks(32, 32))
    assert dm.ishermitian(DN)
    with pytest.raises(dm.PosDefException) as e:
        dm.cholesky(DN)
    assert e.value.info == 1


def _gemm_cases():
    sizes, parts = [], []
    for r, c in ((4, 4), (12, 12)):
        sizes += [((r, c), (c, r)), ((r // 2, c), (c, r // 2)), ((r, c // 2), (c // 2, r))]
        parts += [((r, c), (c, r)), ((r // 2, c), (c, r // 2)), ((r, c // 2), (c // 2, r))]
    return [(sa, sb, pa, pb) for sa, sb in sizes for pa, pb in parts]


@pytest.mark.parametrize("szA,szB,partA,partB", _gemm_cases())
def test_reference_gemm_testset(dm, szA, szB, partA, partB):
    # test/array/linalg/matmul.jl's test_gemm! for Float64: out-of-place and
    # in-place products with every transpose the testset uses, and the syrk
    # forms (A A', A' A) for square A; every result `≈` the dense product
    rng = np.random.default_rng(sum(szA) * 31 + sum(partA))
    A, B = rng.random(szA), rng.random(szB)
    DA = dm.DArray.from_array(A, dm.Blocks(*partA))
    DB = dm.DArray.from_array(B, dm.Blocks(*partB))
    partC = dm.Blocks(partA[0], partB[1])
    sq = szA == szB
    outs = [("N", "N", DA, DB, A @ B), ("T", "T", DA, DB, A.T @ B.T)]
    if sq:
        outs += [("N", "T", DA, DB, A @ B.T), ("T", "N", DA, DB, A.T @ B),
                 ("N", "T", DA, DA, A @ A.T), ("T", "N", DA, DA, A.T @ A)]
    for ta, tb, X, Y, ref in outs:
        assert isapprox(dm.mul(X, Y, ta, tb).collect(), ref), ("out of place", ta, tb)
        DC = dm.DArray.from_array(np.zeros(ref.shape), partC)
        dm.mul_(DC, X, Y, 1.0, 0.0, ta, tb)
        assert DC.partitioning == partC
        assert isapprox(DC.collect(), ref), ("in place", ta, tb)


@pytest.mark.parametrize("szA,partA,pb", [((4, 4), (4, 4), 4), ((2, 4), (2, 4), 4),
                                          ((4, 2), (4, 2), 2), ((12, 12), (6, 12), 12),
                                          ((12, 6), (12, 6), 6), ((12, 12), (12, 6), 4),
                                          ((700, 500), (128, 96), 128)])
def test_matvec(dm, szA, partA, pb):
    # matmul.jl's test_gemv! shapes: op(A) x out of place and in place, x
    # blocked like or unlike A's inner blocking (re-blocked to the smallest)
    rng = np.random.default_rng(sum(szA) + pb)
    A = rng.random(szA)
    DA = dm.DArray.from_array(A, dm.Blocks(*partA))
    for ta, n_in in (("N", szA[1]), ("T", szA[0])):
        x = rng.random(n_in)
        Dx = dm.DArray.from_array(x, dm.Blocks(pb))
        ref = (A.T if ta == "T" else A) @ x
        assert isapprox(dm.mul(DA, Dx, ta).collect(), ref), ("out of place", ta)
        y0 = rng.random(ref.shape[0])
        Dy = dm.DArray.from_array(y0, dm.Blocks(5))
        dm.mul_(Dy, DA, Dx, 2.0, -1.0, ta)
        assert Dy.partitioning == dm.Blocks(5)
        assert isapprox(Dy.collect(), 2.0 * ref - y0), ("in place", ta)
        assert np.array_equal(Dx.collect(), x)


@pytest.mark.parametrize("uplo,trans", [("U", "N"), ("U", "T"), ("L", "N"), ("L", "T")])
def test_trsv_alpha_matches_oracle(dm, uplo, trans):
    """Dagger.trsv! with alpha != 1 (src/array/trsm.jl:16-54): alpha is the
    first step's GEMV beta, as in the oracle (pinned by closed forms)."""
    n, nb = 600, 128
    rng = np.random.default_rng(sum(map(ord, "t" + uplo + trans)))
    A = 0.05 * rng.standard_normal((n, n)) + np.diag(np.full(n, 2.0))
    b = rng.standard_normal(n)
    DA = dm.DArray.from_array(A, dm.Blocks(nb, nb))
    Db = dm.DArray.from_array(b, dm.Blocks(nb))
    dm.trsv_(uplo, trans, "N", -2.5, DA, Db)
    ref = O.trsv(uplo, trans, "N", A, b, nb, alpha=-2.5)
    assert np.allclose(Db.collect(), ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("uplo,trans", [("U", "N"), ("U", "T"), ("L", "N"), ("L", "T")])
def test_trsm_right_side_datadeps_equals_native(dm, uplo, trans):
    """The datadeps mirror of trsm!'s side 'R' branches
    (src/array/trsm.jl:143-189) records the same tasks as the native
    builder, so both results are bitwise equal and match the oracle."""
    n, nb, m = 384, 128, 200
    rng = np.random.default_rng(sum(map(ord, "R" + uplo + trans)))
    A = 0.05 * rng.standard_normal((n, n)) + np.diag(np.full(n, 2.0))
    B = rng.standard_normal((m, n))
    DA = dm.DArray.from_array(A, dm.Blocks(nb, nb))
    outs = []
    for via in ("native", "datadeps"):
        DB = dm.DArray.from_array(B, dm.Blocks(nb, nb))
        dm.trsm_("R", uplo, trans, "N", 0.75, DA, DB, via=via)
        outs.append(DB.collect())
    assert np.array_equal(outs[0], outs[1])
    ref = O.trsm("R", uplo, trans, "N", 0.75, A, B, nb)
    assert np.allclose(outs[0], ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("uplo", ["U", "L"])
def test_float32_vector_ldiv(dm, uplo):
    """ldiv! of an FP32 Cholesky with a DVector right-hand side runs the
    trsv!/gemv! tile tasks in FP32 (GEMV tile kernel for FP32)."""
    n, nb = 640, 128
    A = gen.spd(n, 90)
    b = gen.randn(n, 91)
    DA = dm.DArray.from_array(A.astype(np.float32), dm.Blocks(nb, nb))
    F = dm.cholesky_(DA, uplo=uplo)
    Db = dm.DArray.from_array(b.astype(np.float32), dm.Blocks(nb))
    dm.ldiv_(F, Db)
    x = Db.collect().astype(np.float64)
    xref = np.linalg.solve(A, b)
    assert np.linalg.norm(x - xref) / np.linalg.norm(xref) < 1e-4
    assert O.backward_error(A, x, b) < 50 * n * np.finfo(np.float32).eps
