"""GPU LU (lu! / lu / ldiv! for LU of a DMatrix) against the oracle's
restatement of Dagger's tiled LU (src/array/lu.jl) and after the
reference's own LU testset (test/array/linalg/lu.jl, Float64 and Float32;
complex element types are not built)."""
import numpy as np
import pytest

from oracle import dagger_tiled as O

pytestmark = pytest.mark.gpu


def _check_factorization(F, A, rtol):
    """The reference testset's checks: P A = L U, U upper, L lower, P a
    permutation with det +-1."""
    L, U, P = F.L, F.U, F.P
    assert np.allclose(P @ A, L @ U, rtol=rtol, atol=rtol * np.abs(A).max())
    assert np.array_equal(U, np.triu(U)) and np.array_equal(L, np.tril(L))
    assert np.allclose(P.T @ P, np.eye(P.shape[0]))
    assert set(np.unique(P)) <= {0.0, 1.0}
    assert abs(abs(np.linalg.det(P)) - 1.0) < 1e-12


@pytest.mark.parametrize("pivot", ["nopivot", "rowmax"])
@pytest.mark.parametrize("T", [np.float64, np.float32])
def test_reference_lu_testset(dm, pivot, T):
    """test/array/linalg/lu.jl:17-101: 128 x 128 Blocks(64, 64), out of place
    and in place, non-square blocks Blocks(64, 32), solves against dense."""
    rng = np.random.default_rng(17)
    pv = dm.NoPivot() if pivot == "nopivot" else dm.RowMaximum()
    rtol = 1e-2 if T == np.float32 else (1e-7 if pivot == "nopivot" else 1e-12)
    A = rng.random((128, 128)).astype(T)
    B = A.copy()
    DA = dm.DArray.from_array(A, dm.Blocks(64, 64))
    F = dm.lu(DA, pv)                                   # out of place
    assert isinstance(F, dm.LU)
    if not (T == np.float32 and pivot == "nopivot"):   # the reference skips this case
        _check_factorization(F, B.astype(np.float64), rtol)
    b = rng.random(128).astype(T)
    x = dm.ldiv(F, b).collect()
    tol = 1e-2 if T == np.float32 else 1e-6
    assert np.allclose(x, np.linalg.solve(B.astype(np.float64), b), rtol=tol, atol=tol)
    assert np.array_equal(DA.collect(), B)             # lu left A alone
    DA2 = dm.DArray.from_array(A, dm.Blocks(64, 64))
    F2 = dm.lu_(DA2, pv)                                # in place
    if not (T == np.float32 and pivot == "nopivot"):
        _check_factorization(F2, B.astype(np.float64), rtol)
    assert not np.allclose(DA2.collect(), B)            # changes reached A
    An = rng.random((128, 128)).astype(T)               # non-square blocks
    Fn = dm.lu(dm.DArray.from_array(An, dm.Blocks(64, 32)), pv)
    assert Fn.factors.partitioning == dm.Blocks(64, 32)
    if not (T == np.float32 and pivot == "nopivot"):
        _check_factorization(Fn, An.astype(np.float64),
                             1e-2 if T == np.float32 else np.sqrt(np.finfo(float).eps))


def test_reference_lu_singular(dm):
    """lu(ones(Blocks(64,64), 128, 128)) throws SingularException; with
    allowsingular the factorization returns info 2 (the second pivot)."""
    with pytest.raises(dm.SingularException) as e:
        dm.lu(dm.DArray.from_array(np.ones((128, 128)), dm.Blocks(64, 64)))
    assert e.value.info == 2
    F = dm.lu(dm.DArray.from_array(np.ones((128, 128)), dm.Blocks(64, 64)),
              allowsingular=True)
    assert F.info == 2


@pytest.mark.parametrize("m,n,nb", [(1024, 1024, 128), (1000, 1000, 128), (2048, 2048, 256),
                                    (768, 512, 128), (512, 768, 128), (4096, 4096, 512)])
def test_lu_matches_oracle(dm, m, n, nb):
    """The device LU against the oracle's restatement on the same input:
    identical pivots, factors within 1e-11 relative, the same info."""
    rng = np.random.default_rng(m + 7 * n + nb)
    A = rng.standard_normal((m, n))
    DA = dm.DArray.from_array(A, dm.Blocks(nb, nb))
    F = dm.lu_(DA)
    ref, ipiv, info = O.lu(A, nb, pivot=True)
    assert F.info == info == 0
    assert np.array_equal(F.ipiv, ipiv)
    got = DA.collect()
    assert np.abs(got - ref).max() <= 1e-11 * np.abs(ref).max()


def test_lu_nopivot_matches_oracle(dm):
    n, nb = 1024, 256
    rng = np.random.default_rng(5)
    A = rng.standard_normal((n, n)) + n * np.eye(n)
    DA = dm.DArray.from_array(A, dm.Blocks(nb, nb))
    F = dm.lu_(DA, dm.NoPivot())
    ref, ipiv, info = O.lu(A, nb, pivot=False)
    assert np.array_equal(F.ipiv, ipiv) and F.info == info == 0
    assert np.abs(DA.collect() - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("nrhs", [0, 1, 7, 300])
def test_lu_ldiv_matches_oracle(dm, nrhs):
    """ldiv!(::LU, B) (src/array/linalg.jl:248-310) for vector (nrhs 0) and
    matrix right-hand sides against the oracle and a dense solve."""
    n, nb = 1536, 256
    rng = np.random.default_rng(nrhs + 3)
    A = rng.standard_normal((n, n))
    DA = dm.DArray.from_array(A, dm.Blocks(nb, nb))
    F = dm.lu_(DA)
    B = rng.standard_normal(n) if nrhs == 0 else rng.standard_normal((n, nrhs))
    X = dm.ldiv(F, B).collect()
    Xo = O.lu_ldiv(DA.collect(), F.ipiv, B, nb)
    assert np.allclose(X, Xo, rtol=1e-10, atol=1e-10)
    Xd = np.linalg.solve(A, B)
    assert np.abs(X - Xd).max() <= 1e-9 * max(np.abs(Xd).max(), 1.0)


def test_lu_schedules_are_bit_identical(dm):
    """Parallel equals sequential execution: one stream without batching and
    the batched two-stream look-ahead schedule give the same bits."""
    n, nb = 1280, 256
    A = np.random.default_rng(8).standard_normal((n, n))
    outs = []
    for opts in (dm.SEQUENTIAL, dm.DEFAULT_OPTIONS.replace(use_graph=False)):
        DA = dm.DArray.from_array(A, dm.Blocks(nb, nb))
        F = dm.lu_(DA, options=opts)
        outs.append((DA.collect(), F.ipiv))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])
