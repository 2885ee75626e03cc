"""Pins the CPU oracle (oracle/dagger_tiled.py, oracle/gen.py) before it is
trusted as the checker of the GPU path.  CPU only."""
import math

import numpy as np
import pytest
import scipy.linalg as sla

from oracle import dagger_tiled as O
from oracle import gen

GOLD = np.load(__import__("os").path.join(__import__("os").path.dirname(__file__), "golden",
                                          "cholesky_golden.npz"))


def dense_potrf(A, uplo):
    F, info = sla.lapack.dpotrf(A, lower=1 if uplo == "L" else 0, clean=0)
    return F, info


# ------------------------------------------------------------ Cholesky
@pytest.mark.parametrize("uplo", ["U", "L"])
@pytest.mark.parametrize("n,nb", [(1, 1), (7, 3), (64, 16), (100, 32), (128, 32), (130, 64),
                                  (96, 96), (50, 64)])
def test_cholesky_matches_dense_lapack(n, nb, uplo):
    A = gen.spd(n, 100 + n)
    F, info = O.cholesky(A, nb, uplo)
    Fd, infod = dense_potrf(A, uplo)
    assert info == 0 and infod == 0
    assert O.rel_factor_diff(F, Fd, uplo) < 1e-13
    # the other triangle keeps the input, as the reference leaves it
    other = np.tril(np.ones((n, n)), -1) if uplo == "U" else np.triu(np.ones((n, n)), 1)
    assert np.array_equal(F * other, A * other)
    assert O.reconstruction_error(F, A, uplo) < 1e-14


def _py_cholesky_upper(A):
    """Textbook Cholesky-Banachiewicz (pure Python), U with A = U^T U."""
    n = len(A)
    U = [[0.0] * n for _ in range(n)]
    for j in range(n):
        s = A[j][j] - sum(U[k][j] ** 2 for k in range(j))
        U[j][j] = math.sqrt(s)
        for i in range(j + 1, n):
            U[j][i] = (A[j][i] - sum(U[k][j] * U[k][i] for k in range(j))) / U[j][j]
    return U


@pytest.mark.parametrize("n,nb", [(5, 2), (6, 4), (9, 3)])
def test_cholesky_brute_force_tiny(n, nb):
    A = gen.spd(n, 7 * n)
    U = np.array(_py_cholesky_upper(A.tolist()))
    F, info = O.cholesky(A, nb, "U")
    assert info == 0
    assert np.allclose(np.triu(F), U, rtol=0, atol=1e-14)


def test_cholesky_closed_forms():
    n = 96
    F, info = O.cholesky(np.eye(n), 32, "U")
    assert info == 0 and np.array_equal(np.triu(F), np.eye(n))
    d = np.linspace(1.0, 50.0, n)
    F, info = O.cholesky(np.diag(d), 32, "L")
    assert np.allclose(np.tril(F), np.diag(np.sqrt(d)), rtol=1e-15, atol=0)
    # constructed factor: A = U^T U with U upper, positive diagonal
    rng = np.random.default_rng(5)
    U = np.triu(rng.standard_normal((n, n))) + np.diag(np.full(n, 4.0))
    U[np.diag_indices(n)] = np.abs(U[np.diag_indices(n)])
    A = U.T @ U
    F, info = O.cholesky(A, 32, "U")
    assert info == 0
    assert np.allclose(np.triu(F), U, rtol=1e-12, atol=1e-12)


def test_cholesky_non_pd_info_is_tile_local():
    # reference construction: A[1,1] = -100 fails the first pivot
    A = GOLD["npd_first/A"]
    F, info = O.cholesky(A, 32, "U")
    assert info == 1
    # failure inside the third diagonal tile: tile-local LAPACK info
    A = GOLD["npd_tile3/A"]
    dense = int(GOLD["npd_tile3/dense_info"][0])
    F, info = O.cholesky(A, 32, "U")
    assert dense > 64 and dense <= 96
    assert info == dense - 64  # potrf_checked! reports the tile's own info


@pytest.mark.parametrize("key", ["n64_b16_U", "n64_b16_L", "n128_b32_U", "n128_b32_L",
                                 "n100_b32_U", "n1_b1_U", "n7_b3_L"])
def test_cholesky_and_ldiv_golden(key):
    n, nb, lower, seed, nrhs = (int(v) for v in GOLD[f"{key}/meta"])
    uplo = "L" if lower else "U"
    A, Fg, B, Xg = (GOLD[f"{key}/{s}"] for s in ("A", "F", "B", "X"))
    F, info = O.cholesky(A, nb, uplo)
    assert info == 0
    assert O.rel_factor_diff(F, Fg, uplo) < 1e-13
    X = O.ldiv(F, uplo, B, nb)
    assert np.allclose(X, Xg, rtol=1e-11, atol=1e-12)
    if nrhs == 1:
        x = O.ldiv(F, uplo, B[:, 0], nb)
        assert np.allclose(x, Xg[:, 0], rtol=1e-11, atol=1e-12)


def test_planted_mistake_is_caught(monkeypatch):
    """A plausible oracle bug (gemm transpose flags swapped in the trailing
    update) must fail the dense-LAPACK comparison used above."""
    A = gen.spd(96, 3)
    good, _ = O.cholesky(A, 32, "U")
    Fd, _ = dense_potrf(A, "U")
    assert O.rel_factor_diff(good, Fd) < 1e-13
    real = O.tile_gemm

    def wrong(ta, tb, alpha, X, Y, beta, Cm):
        return real("N" if ta == "T" else "T", tb, alpha, X, Y, beta, Cm)

    monkeypatch.setattr(O, "tile_gemm", wrong)
    bad, _ = O.cholesky(A, 32, "U")
    assert O.rel_factor_diff(bad, Fd) > 1e-3


# ------------------------------------------------------------ trsm! / ldiv!
def _tri(A, uplo, diag):
    T = np.triu(A) if uplo == "U" else np.tril(A)
    if diag == "U":
        T = T.copy()
        T[np.diag_indices(len(T))] = 1.0
    return T


@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("uplo", ["U", "L"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("diag", ["N", "U"])
def test_trsm_all_branches(side, uplo, trans, diag):
    n, nb = 96, 32
    rng = np.random.default_rng(sum(map(ord, side + uplo + trans + diag)))
    A = 0.05 * rng.standard_normal((n, n)) + np.diag(np.full(n, 2.0))
    B = rng.standard_normal((n, 64) if side == "L" else (64, n))
    alpha = 1.0 if side == "R" and trans == "T" else 0.5
    X = O.trsm(side, uplo, trans, diag, alpha, A, B, nb)
    T = _tri(A, uplo, diag)
    opT = T.T if trans == "T" else T
    if side == "L":
        ref = np.linalg.solve(opT, alpha * B)
    else:
        ref = np.linalg.solve(opT.T, alpha * B.T).T
    assert np.allclose(X, ref, rtol=1e-10, atol=1e-11)


@pytest.mark.parametrize("ta", ["N", "T"])
@pytest.mark.parametrize("tb", ["N", "T"])
def test_gemm_tile_loop(ta, tb):
    # ragged tiles on every dimension (m=70, n=45, k=38 over nb=16); beta
    # applied once, by the first k-step
    m, n, k, nb = 70, 45, 38, 16
    rng = np.random.default_rng(sum(map(ord, ta + tb)))
    A = rng.standard_normal((k, m) if ta == "T" else (m, k))
    B = rng.standard_normal((n, k) if tb == "T" else (k, n))
    Cm = rng.standard_normal((m, n))
    got = O.gemm(ta, tb, -0.5, A, B, 2.0, Cm, nb)
    ref = 2.0 * Cm - 0.5 * (A.T if ta == "T" else A) @ (B.T if tb == "T" else B)
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("uplo", ["U", "L"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("diag", ["N", "U"])
def test_trsv_all_branches(uplo, trans, diag):
    # ragged last tile (100 = 3*32 + 4); the stored diagonal is not 1, so a
    # unit-diagonal solve that read it would fail
    n, nb = 100, 32
    rng = np.random.default_rng(sum(map(ord, "v" + uplo + trans + diag)))
    A = 0.05 * rng.standard_normal((n, n)) + np.diag(np.full(n, 2.0))
    b = rng.standard_normal(n)
    x = O.trsv(uplo, trans, diag, A, b, nb)
    T = _tri(A, uplo, diag)
    ref = np.linalg.solve(T.T if trans == "T" else T, b)
    assert np.allclose(x, ref, rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("uplo,trans", [("U", "N"), ("U", "T"), ("L", "N"), ("L", "T")])
def test_trsv_alpha_enters_first_updates_only(uplo, trans):
    """trsv! with alpha (src/array/trsm.jl:16-54): alpha is only the GEMV
    beta of the first step, so with one block it has no effect, and with two
    blocks the first block solved is op(T_ff)^-1 b_f while the other is
    op(T_ss)^-1 (alpha b_s - op(T)_sf x_f) -- closed forms from dense
    solves.  Uniform scaling by alpha (a plausible restatement) fails."""
    n, nb = 64, 32
    rng = np.random.default_rng(sum(map(ord, "a" + uplo + trans)))
    A = 0.05 * rng.standard_normal((n, n)) + np.diag(np.full(n, 2.0))
    b = rng.standard_normal(n)
    alpha = -2.5
    # one block: alpha ignored
    x1 = O.trsv(uplo, trans, "N", A, b, n, alpha)
    T = _tri(A, uplo, "N")
    op = T.T if trans == "T" else T
    assert np.allclose(x1, np.linalg.solve(op, b), rtol=1e-12, atol=1e-13)
    # two blocks: which block is solved first follows op(T)'s triangle
    x2 = O.trsv(uplo, trans, "N", A, b, nb, alpha)
    lower_op = (uplo == "L") == (trans == "N")
    f, s_ = (slice(0, nb), slice(nb, n)) if lower_op else (slice(nb, n), slice(0, nb))
    xf = np.linalg.solve(op[f, f], b[f])
    xs = np.linalg.solve(op[s_, s_], alpha * b[s_] - op[s_, f] @ xf)
    ref = np.empty(n)
    ref[f], ref[s_] = xf, xs
    assert np.allclose(x2, ref, rtol=1e-12, atol=1e-13)
    assert not np.allclose(x2, alpha * np.linalg.solve(op, b), rtol=1e-6)


def test_ishermitian_restatement():
    """ishermitian (src/array/linalg.jl:38-48, 90-124): exact symmetry on
    diagonal chunks, the sqrt(eps)-relative Frobenius rule on chunk pairs.
    Each rule is pinned by a case that only it decides: a one-ulp asymmetry
    inside a diagonal chunk fails; the same asymmetry across chunks passes;
    a pair perturbed at half / twice the threshold passes / fails; NaN and
    non-square fail; a pair check that compared A_ij with A_ji (no
    transpose) would fail the plain symmetric case."""
    n, nb = 96, 32
    rng = np.random.default_rng(5)
    G = rng.standard_normal((n, n))
    A = G + G.T
    assert O.ishermitian(A, nb)
    assert not O.ishermitian(A[:, :64], nb)
    B = A.copy()
    B[3, 5] = np.nextafter(B[3, 5], np.inf)          # inside diagonal chunk 0
    assert not O.ishermitian(B, nb)
    B = A.copy()
    B[70, 5] = np.nextafter(B[70, 5], np.inf)        # chunk pair (2, 0)
    assert O.ishermitian(B, nb)
    # threshold of pair (2, 0): sqrt(eps) * max of the two chunk norms
    a1, a2 = A[64:96, 0:32], A[0:32, 64:96]
    thr = np.sqrt(np.finfo(float).eps) * max(np.linalg.norm(a1), np.linalg.norm(a2))
    for scale, want in ((0.5, True), (2.0, False)):
        B = A.copy()
        B[70, 5] += scale * thr                        # ||A_ij - A_ji^T||_F = scale * thr
        assert O.ishermitian(B, nb) == want
    B = A.copy()
    B[40, 40] = np.nan
    assert not O.ishermitian(B, nb)
    B = A.copy()
    B[70, 5] = B[5, 70] = np.nan
    assert not O.ishermitian(B, nb)
    # ragged last chunk (n = 100 over 32): the corner chunk is 4 x 4
    C2 = rng.standard_normal((100, 100))
    C2 = C2 + C2.T
    assert O.ishermitian(C2, 32)
    C2[98, 99] += 1.0
    assert not O.ishermitian(C2, 32)
    # FP32 uses sqrt(eps(Float32))
    A32 = A.astype(np.float32)
    B32 = A32.copy()
    thr32 = np.sqrt(np.finfo(np.float32).eps) * max(np.linalg.norm(A32[64:96, 0:32]),
                                                    np.linalg.norm(A32[0:32, 64:96]))
    B32[70, 5] += np.float32(0.25 * thr32)
    assert O.ishermitian(B32, nb)
    B32[70, 5] += np.float32(4.0 * thr32)
    assert not O.ishermitian(B32, nb)


@pytest.mark.parametrize("uplo", ["U", "L"])
@pytest.mark.parametrize("vector", [False, True])
def test_ldiv_solves_system(uplo, vector):
    n, nb = 160, 64
    A = gen.spd(n, 77)
    F, info = O.cholesky(A, nb, uplo)
    B = gen.randn(n if vector else n * 3, 78).reshape((n,) if vector else (3, n)).T.copy()
    X = O.ldiv(F, uplo, B, nb)
    assert np.allclose(X, np.linalg.solve(A, B), rtol=1e-11, atol=1e-12)
    assert O.backward_error(A, X, B) < 10 * n * np.finfo(float).eps


def test_backward_error_known_values():
    """Closed forms: an exact solution has zero backward error; with A = I,
    X = 0 the error is ||B|| / ||B|| = 1; the per-column form is the largest
    single-column value, and the matrix form sums rows across columns."""
    A = np.eye(3)
    B = np.array([[1.0, 2.0], [0.0, -4.0], [3.0, 0.5]])
    assert O.backward_error(A, B.copy(), B) == 0.0
    assert O.backward_error(A, np.zeros_like(B), B) == 1.0
    X = B.copy()
    X[0, 0] += 1.0  # column 0: ||r|| = 1, ||x|| = 3, ||b|| = 3 -> 1/6
    assert np.isclose(O.backward_error_columns(A, X, B), 1.0 / 6.0)
    # rows of |X| sum to 4, 4, 3.5 and rows of |B| to 3, 4, 3.5 -> 1 / (4 + 4)
    assert np.isclose(O.backward_error(A, X, B), 1.0 / 8.0)


def test_threaded_executor_equals_sequential():
    A = gen.spd(256, 8)
    with O.blas_single_thread():
        F1, _ = O.cholesky(A, 64, "U")
        F2, _ = O.cholesky(A, 64, "U", threaded=True, workers=4)
    assert np.array_equal(F1, F2)


# ------------------------------------------------------------ datadeps rules
def test_datadeps_edge_rules():
    a, b, c = np.zeros(1), np.zeros(1), np.zeros(1)
    R = O.Region()
    nop = lambda *x: None  # noqa: E731
    R.spawn("r1", nop, (), [(a, O.IN)])          # 0
    R.spawn("r2", nop, (), [(a, O.IN)])          # 1: R after R -> no edge
    R.spawn("w1", nop, (), [(a, O.INOUT)])       # 2: W after R,R -> 0->2, 1->2
    R.spawn("r3", nop, (), [(a, O.IN), (b, O.IN)])  # 3: R after W -> 2->3
    R.spawn("w2", nop, (), [(b, O.OUT)])         # 4: W after R(b) -> 3->4
    R.spawn("w3", nop, (), [(c, O.OUT)])         # 5: independent
    R.spawn("w4", nop, (), [(c, O.OUT)])         # 6: W after W -> 5->6
    assert O.datadeps_edges(R.tasks) == {(0, 2), (1, 2), (2, 3), (3, 4), (5, 6)}


# ------------------------------------------------------------ generator
def test_philox_known_answers():
    """Random123 published known-answer vectors for philox4x32_10."""
    kat = [
        ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
        ((0xffffffff,) * 4, (0xffffffff, 0xffffffff),
         (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
        ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
         (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
    ]
    for ctr, key, want in kat:
        got = gen.philox4x32_10_block(*[np.array([c]) for c in ctr], *key)
        assert tuple(int(g[0]) for g in got) == want


def test_randn_moments_and_determinism():
    z = gen.randn(400000, 42)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1.0) < 0.01
    assert np.array_equal(z[1000:2000], gen.randn(1000, 42, offset=1000))
    assert not np.array_equal(z[:1000], gen.randn(1000, 43))
    A = gen.spd(64, 1)
    assert np.array_equal(A, A.T)
    assert np.linalg.eigvalsh(A).min() > 0.9


@pytest.mark.parametrize("m,n,nb", [(300, 300, 64), (257, 129, 32), (64, 64, 64), (1, 1, 4)])
def test_norm2_closed_forms(m, n, nb):
    """The chunked norm2 restatement against the dense Frobenius norm, its
    triangular views against triu/tril, and exact values on constant and
    identity matrices (a dropped diagonal or chunk would fail these)."""
    rng = np.random.default_rng(m * 7 + n)
    A = rng.standard_normal((m, n))
    assert abs(O.norm2(A, nb) - np.linalg.norm(A)) <= 1e-13 * np.linalg.norm(A)
    assert O.norm2(np.ones((m, n)), nb) == pytest.approx(np.sqrt(m * n), rel=1e-15)
    if m == n:
        up, lo = O.norm2(A, nb, "upper"), O.norm2(A, nb, "lower")
        assert up == pytest.approx(np.linalg.norm(np.triu(A)), rel=1e-13)
        assert lo == pytest.approx(np.linalg.norm(np.tril(A)), rel=1e-13)
        d2 = float(np.sum(np.diag(A) ** 2))
        assert up ** 2 + lo ** 2 - d2 == pytest.approx(np.linalg.norm(A) ** 2, rel=1e-12)
        assert O.norm2(np.eye(n), nb, "upper") == pytest.approx(np.sqrt(n), rel=1e-15)


def test_chkfinite_restatement():
    A = np.ones((130, 70))
    assert O.chkfinite(A, 32)
    for bad in (np.nan, np.inf, -np.inf):
        B = A.copy()
        B[129, 69] = bad  # last element of the ragged corner chunk
        with pytest.raises(ValueError, match="Inf or NaN"):
            O.chkfinite(B, 32)


@pytest.mark.parametrize("m,n,nb", [(128, 128, 64), (200, 200, 64), (300, 256, 64),
                                    (200, 300, 64), (130, 130, 32), (64, 64, 64)])
def test_lu_rowmaximum_matches_dense_lapack(m, n, nb):
    """lu!(A, RowMaximum()) restatement (src/array/lu.jl:289-371): the tiled
    panel/swap/trsm/gemm nest gives dense LAPACK getrf's pivots exactly and
    its factors to rounding, P A = L U, and ldiv! solves the system."""
    import scipy.linalg as sl
    rng = np.random.default_rng(m * 3 + n)
    A = rng.standard_normal((m, n))
    F, ipiv, info = O.lu(A, nb, pivot=True)
    lu, piv = sl.lu_factor(A)
    assert info == 0
    assert np.array_equal(ipiv, piv + 1)
    assert np.abs(F - lu).max() <= 1e-11 * np.abs(lu).max()
    L, U, P = O.lu_unpack(F, ipiv)
    assert np.abs(P @ A - L @ U).max() <= 1e-13 * np.abs(A).max() * max(m, n)
    if m == n:
        b = rng.standard_normal((m, 3))
        assert np.allclose(O.lu_ldiv(F, ipiv, b, nb), np.linalg.solve(A, b), rtol=1e-9,
                           atol=1e-10)


def test_lu_nopivot_and_singular_closed_forms():
    """NoPivot (src/array/lu.jl:10-40) on a diagonally dominant matrix equals
    the brute-force Doolittle LU of the whole matrix; ipiv is 1..n; an
    all-ones matrix is singular with info 2 (LAPACK: the second pivot is
    exactly zero); a diagonal matrix factors to itself with identity pivots."""
    rng = np.random.default_rng(9)
    n, nb = 96, 32
    A = rng.standard_normal((n, n)) + n * np.eye(n)
    F, ipiv, info = O.lu(A, nb, pivot=False)
    D = A.copy()
    for j in range(n):                      # brute force, no pivoting
        for i in range(j + 1, n):
            D[i, j] /= D[j, j]
            D[i, j + 1:] -= D[i, j] * D[j, j + 1:]
    assert np.allclose(F, D, rtol=1e-12, atol=1e-12)
    assert np.array_equal(ipiv, np.arange(1, n + 1))
    _, _, info1 = O.lu(np.ones((128, 128)), 64, pivot=True)
    assert info1 == 2
    Dg = np.diag(np.arange(1.0, 65.0))
    F2, ip2, info2 = O.lu(Dg, 16, pivot=True)
    assert info2 == 0 and np.array_equal(ip2, np.arange(1, 65)) and np.array_equal(F2, Dg)


def test_lu_planted_mistake_is_caught():
    """Skipping the row swaps of the block columns left of the panel (a
    plausible restatement error: swaprows_trail! over 1:k-1 as well as
    k+1:nt) breaks P A = L U, which the dense comparison catches."""
    rng = np.random.default_rng(2)
    A = rng.standard_normal((192, 192))
    orig_lu = O.lu

    def lu_without_left_swaps(A, nb, pivot=True):
        T = O.TiledMatrix(np.array(A, copy=True), nb, nb)
        Ac = T.chunks
        mt, nt = T.mt, T.nt
        ipiv = np.arange(1, min(T.m, T.n) + 1)
        info = [0]
        for k in range(min(mt, nt)):
            chunk = ipiv[k * nb:k * nb + min(T.m - k * nb, T.cols[k])]
            O.lu_panel([Ac[i, k] for i in range(k, mt)], k, nb, chunk, info)
            for j in range(k + 1, nt):            # the left columns are skipped
                O.swaprows_trail(Ac[k, j], [Ac[i, j] for i in range(k, mt)], chunk, k, nb)
            for j in range(k + 1, nt):
                O.tile_trsm("L", "L", "N", "U", 1.0, Ac[k, k], Ac[k, j])
                for i in range(k + 1, mt):
                    O.tile_gemm("N", "N", -1.0, Ac[i, k], Ac[k, j], 1.0, Ac[i, j])
        return T.collect(), ipiv, info[0]
    F, ipiv, _ = lu_without_left_swaps(A, 64)
    L, U, P = O.lu_unpack(F, ipiv)
    assert np.abs(P @ A - L @ U).max() > 1e-3
    F, ipiv, _ = orig_lu(A, 64)
    L, U, P = O.lu_unpack(F, ipiv)
    assert np.abs(P @ A - L @ U).max() < 1e-12


def test_cholesky_leading_steps_sample():
    """The bounded timing sample (the first k-iterations only): after one
    step, block row 0 is the full factor's block row 0 and the trailing upper
    triangle is the Schur complement A22 - U01^T U01 (closed form)."""
    rng = np.random.default_rng(31)
    n, nb = 384, 128
    G = rng.standard_normal((n, n))
    A = G @ G.T / n + np.eye(n)
    F1, info = O.cholesky(A, nb, "U", steps=1)
    Ff, _ = O.cholesky(A, nb, "U")
    assert info == 0
    assert np.allclose(np.triu(F1[:nb, :]), np.triu(Ff[:nb, :]), rtol=0, atol=1e-13)
    U01 = F1[:nb, nb:]
    S = A[nb:, nb:] - U01.T @ U01
    assert np.allclose(np.triu(F1[nb:, nb:]), np.triu(S), rtol=0, atol=1e-12)
    assert np.array_equal(np.tril(F1, -1), np.tril(A, -1))  # lower part untouched


@pytest.mark.parametrize("T", [np.complex128, np.complex64])
@pytest.mark.parametrize("uplo", ["U", "L"])
def test_complex_cholesky_matches_dense_lapack(T, uplo):
    """The restatement for T <: Complex (trans 'C', herk!, ragged tiles)
    equals the dense factor (numpy's zpotrf route) and reconstructs A."""
    rng = np.random.default_rng(41)
    n, nb = 173, 48
    G = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A = G @ G.conj().T / n + np.eye(n)
    A = (0.5 * (A + A.conj().T)).astype(T)
    F, info = O.cholesky(A, nb, uplo)
    assert info == 0
    L = np.linalg.cholesky(A.astype(np.complex128))
    ref = L.conj().T if uplo == "U" else L
    tol = 1e-13 if T == np.complex128 else 2e-6
    tri = np.triu if uplo == "U" else np.tril
    assert np.abs(tri(F) - ref).max() <= tol * np.abs(ref).max()
    U = np.triu(F) if uplo == "U" else np.tril(F).conj().T
    assert np.abs(U.conj().T @ U - A).max() <= 10 * tol * np.abs(A).max()
    assert np.all(np.diag(F).imag == 0)


def test_complex_herk_zeroes_the_imaginary_diagonal_and_planted_transpose_fails():
    """herk! on a tile: C - A^H A on the triangle with a real diagonal; and
    the restatement with 'T' in place of 'C' (a plausible slip) no longer
    factors A, which the dense comparison catches."""
    rng = np.random.default_rng(43)
    A = rng.standard_normal((30, 12)) + 1j * rng.standard_normal((30, 12))
    Cm = rng.standard_normal((12, 12)) + 1j * rng.standard_normal((12, 12))
    got = np.asfortranarray(Cm.copy())
    O.tile_herk("U", "C", -1.0, A, 1.0, got)
    want = Cm - A.conj().T @ A
    assert np.allclose(np.triu(got, 1), np.triu(want, 1), atol=1e-13)
    assert np.allclose(np.diag(got).real, np.diag(want).real, atol=1e-13)
    assert np.all(np.diag(got).imag == 0)
    assert np.array_equal(np.tril(got, -1), np.tril(Cm, -1))
    n, nb = 96, 32
    G = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    H = G @ G.conj().T / n + np.eye(n)
    H = 0.5 * (H + H.conj().T)
    T = O.TiledMatrix(np.array(H, copy=True), nb, nb)
    info = np.zeros(1, dtype=np.int64)
    R = O.chol_region(T, "U", info)
    for t in R.tasks:  # the planted mistake: plain transposes of complex tiles
        if t.name in ("trsm", "gemm"):
            t.args = tuple("T" if isinstance(a, str) and a == "C" else a for a in t.args)
    for t in R.tasks:
        try:
            t.run()
        except O.PosDefException:
            break
    L = np.linalg.cholesky(H)
    assert np.abs(np.triu(T.collect()) - L.conj().T).max() > 1e-3
