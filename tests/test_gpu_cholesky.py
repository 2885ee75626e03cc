"""GPU parity of cholesky!/ldiv!/trsm! on resident DArrays against the CPU
oracle's restatement of Dagger's tiled path, plus executor determinism."""
import numpy as np
import pytest

from oracle import dagger_tiled as O
from oracle import gen

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps


def factor_gpu(dm, A, nb, uplo="U", **kw):
    DA = dm.DArray.from_array(A, dm.Blocks(nb, nb))
    F = dm.cholesky_(DA, uplo=uplo, **kw)
    return F, DA.collect()


@pytest.mark.parametrize("n,nb,uplo", [(256, 64, "U"), (256, 64, "L"), (1000, 128, "U"),
                                       (1000, 128, "L"), (2048, 512, "U"), (2048, 512, "L"),
                                       (4096, 512, "U"), (3000, 1024, "U")])
def test_cholesky_matches_oracle(dm, n, nb, uplo):
    A = gen.spd(n, 1000 + n)
    F, got = factor_gpu(dm, A, nb, uplo)
    assert F.info == 0
    ref, info = O.cholesky(A, nb, uplo)
    assert info == 0
    # parity bar: relative factor difference <= 1e-12
    assert O.rel_factor_diff(got, ref, uplo) <= 1e-12
    assert O.reconstruction_error(got, A, uplo) < 10 * n * EPS
    # the other triangle keeps the input, as in the reference
    other = np.tril if uplo == "U" else np.triu
    off = -1 if uplo == "U" else 1
    assert np.array_equal(other(got, off), other(A, off))


def test_cholesky_n8192_nb1024_parity(dm):
    """Parity size of the N=32768 FP64 configuration: N=8192, Blocks(1024,1024)."""
    n, nb = 8192, 1024
    DA = dm.spd_generate(n, dm.Blocks(nb, nb), seed=2024)
    A = DA.collect()
    assert np.array_equal(A, A.T)
    F = dm.cholesky_(DA)
    assert F.info == 0
    got = DA.collect()
    with O.blas_single_thread():
        ref, info = O.cholesky(A, nb, "U", threaded=True)
    assert info == 0
    assert O.rel_factor_diff(got, ref, "U") <= 1e-12
    # backward error of a solve with 64 right-hand sides
    B = gen.randn(n * 64, 7).reshape(64, n).T.copy()
    DB = dm.DArray.from_array(B, dm.Blocks(nb, nb))
    dm.ldiv_(F, DB)
    X = DB.collect()
    assert O.backward_error(A, X, B) <= 10 * n * EPS


def test_device_generator_matches_host(dm):
    n = 512
    DA = dm.spd_generate(n, dm.Blocks(128, 128), seed=9)
    A = DA.collect()
    assert np.array_equal(A, A.T)
    assert np.allclose(A, gen.spd(n, 9), rtol=1e-12, atol=1e-13)


def test_native_and_datadeps_paths_bit_identical(dm):
    n, nb = 1024, 256
    A = gen.spd(n, 5)
    _, a = factor_gpu(dm, A, nb, via="native")
    _, b = factor_gpu(dm, A, nb, via="datadeps")
    assert np.array_equal(a, b)


@pytest.mark.parametrize("uplo", ["U", "L"])
def test_schedules_are_bit_identical(dm, uplo):
    """Parallel execution equals sequential execution (src/datadeps/queue.jl:
    52-54): one stream / no batching vs batched multi-stream look-ahead vs a
    captured hipGraph replayed several times."""
    n, nb = 1536, 256
    A = gen.spd(n, 6)
    outs = []
    for opts in (dm.SEQUENTIAL,
                 dm.PlanOptions(num_streams=2, batch=True, lookahead=1, use_graph=False),
                 dm.PlanOptions(num_streams=4, batch=True, lookahead=0, use_graph=False),
                 dm.PlanOptions(num_streams=3, batch=True, lookahead=2, use_graph=False),
                 dm.PlanOptions(num_streams=2, batch=True, lookahead=2, use_graph=True)):
        for rep in range(2 if opts.use_graph else 1):
            _, f = factor_gpu(dm, A, nb, uplo, options=opts)
            outs.append(f)
    for f in outs[1:]:
        assert np.array_equal(f, outs[0])


def test_wide_plans_issue_eagerly(dm):
    """Graph capture is limited to plans with at most two streams; a wider
    plan that asks for one runs eagerly and reports no graph."""
    n, nb = 768, 256
    DA = dm.DArray.from_array(gen.spd(n, 7), dm.Blocks(nb, nb))
    info = dm.DeviceInt(2)
    dag = dm.Dag.cholesky(DA, "U", info)
    dag.plan(dm.PlanOptions(num_streams=3, batch=True, lookahead=2, use_graph=True))
    assert dag.stats()["graph"] == 0
    dag2 = dm.Dag.cholesky(DA, "U", info)
    dag2.plan(dm.PlanOptions(num_streams=2, batch=True, lookahead=1, use_graph=True))
    assert dag2.stats()["graph"] == 1


def test_graph_replays_on_same_matrix(dm):
    """Cached plan + hipGraph reused across calls on one resident matrix."""
    n, nb = 1024, 256
    A = gen.spd(n, 8)
    DA = dm.DArray.from_array(A, dm.Blocks(nb, nb))
    first = None
    for rep in range(3):
        DA.copy_from_host(A)
        dm.cholesky_(DA)
        f = DA.collect()
        if first is None:
            first = f
        assert np.array_equal(f, first)


@pytest.mark.parametrize("uplo", ["U", "L"])
def test_non_pd_reports_tile_local_info(dm, uplo):
    A = gen.spd(256, 10)
    A[140, 140] = -5.0
    DA = dm.DArray.from_array(A, dm.Blocks(64, 64))
    with pytest.raises(dm.PosDefException) as e:
        dm.cholesky_(DA, uplo=uplo)
    _, rinfo = O.cholesky(A, 64, uplo)
    assert e.value.info == rinfo == 140 - 128 + 1
    DA = dm.DArray.from_array(A, dm.Blocks(64, 64))
    F = dm.cholesky_(DA, uplo=uplo, check=False)
    assert F.info == rinfo and F.global_info == 141
    assert not F.issuccess()
    # reference construction: A[1,1] = -100
    A = gen.spd(128, 11)
    A[0, 0] = -100.0
    DA = dm.DArray.from_array(A, dm.Blocks(32, 32))
    F = dm.cholesky_(DA, check=False)
    assert F.info == 1


def test_out_of_place_cholesky_keeps_input(dm):
    A = gen.spd(256, 12)
    DA = dm.DArray.from_array(A, dm.Blocks(64, 64))
    F = dm.cholesky(DA)
    assert np.array_equal(DA.collect(), A)
    ref, _ = O.cholesky(A, 64, "U")
    assert O.rel_factor_diff(F.factors.collect(), ref) <= 1e-12
    assert np.allclose(F.U, np.linalg.cholesky(A).T, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("uplo", ["U", "L"])
@pytest.mark.parametrize("n,nb,nrhs", [(1024, 256, 64), (1000, 128, 16), (512, 512, 1)])
def test_ldiv_matrix_rhs(dm, uplo, n, nb, nrhs):
    A = gen.spd(n, 20 + n)
    B = gen.randn(n * nrhs, 21).reshape(nrhs, n).T.copy()
    DA = dm.DArray.from_array(A, dm.Blocks(nb, nb))
    F = dm.cholesky_(DA, uplo=uplo)
    DB = dm.DArray.from_array(B, dm.Blocks(nb, nb))
    dm.ldiv_(F, DB)
    X = DB.collect()
    Fh = DA.collect()
    Xref = O.ldiv(Fh, uplo, B, nb)
    assert np.allclose(X, Xref, rtol=1e-12, atol=1e-13)
    assert O.backward_error(A, X, B) <= 10 * n * EPS


@pytest.mark.parametrize("uplo", ["U", "L"])
def test_ldiv_vector_rhs(dm, uplo):
    n, nb = 640, 128
    A = gen.spd(n, 30)
    b = gen.randn(n, 31)
    DA = dm.DArray.from_array(A, dm.Blocks(nb, nb))
    F = dm.cholesky_(DA, uplo=uplo)
    Db = dm.DArray.from_array(b, dm.Blocks(nb))
    dm.ldiv_(F, Db)
    x = Db.collect()
    xref = O.ldiv(DA.collect(), uplo, b, nb)
    assert np.allclose(x, xref, rtol=1e-10)
    assert O.backward_error(A, x, b) <= 10 * n * EPS


@pytest.mark.parametrize("uplo", ["U", "L"])
def test_ldiv_datadeps_and_native(dm, uplo):
    """The datadeps mirror spawns the reference's two trsm! regions
    (src/array/trsm.jl:90-141); the native dm_dag_build_potrs DAG records
    the same tasks, so the two are bitwise equal.  The native ldiv! (inverse
    diagonal tiles, GEMM-only sweeps) reorders the arithmetic and agrees
    with both, and with the oracle, to the 1e-12 parity bar."""
    n, nb, nrhs = 768, 256, 8
    A = gen.spd(n, 40)
    B = gen.randn(n * nrhs, 41).reshape(nrhs, n).T.copy()
    DA = dm.DArray.from_array(A, dm.Blocks(nb, nb))
    F = dm.cholesky_(DA, uplo=uplo)
    outs = {}
    for via in ("native", "datadeps"):
        DB = dm.DArray.from_array(B, dm.Blocks(nb, nb))
        dm.ldiv_(F, DB, via=via)
        outs[via] = DB.collect()
    DB = dm.DArray.from_array(B, dm.Blocks(nb, nb))
    dag = dm.Dag.potrs(DA, DB, uplo).plan(dm.DEFAULT_OPTIONS.replace(use_graph=False))
    dag.run()
    assert np.array_equal(DB.collect(), outs["datadeps"])
    ref = O.ldiv(DA.collect(), uplo, B, nb)
    for X in outs.values():
        assert np.allclose(X, ref, rtol=1e-12, atol=1e-13)
        assert O.backward_error(A, X, B) <= 10 * n * EPS


def test_reblocking_non_square_blocks(dm):
    """Blocks(64,128) on 128x128 (test/array/linalg/solve.jl:4-7): the
    factorization runs on a Blocks(64,64) buffer and copies back."""
    n = 128
    A = gen.spd(n, 50)
    DA = dm.DArray.from_array(A, dm.Blocks(64, 128))
    F = dm.cholesky_(DA)
    got = DA.collect()
    ref, _ = O.cholesky(A, 64, "U")
    assert O.rel_factor_diff(got, ref) <= 1e-12
    assert F.factors.partitioning == dm.Blocks(64, 128)


@pytest.mark.parametrize("side,uplo,trans", [("L", "U", "T"), ("L", "U", "N"), ("L", "L", "N"),
                                             ("L", "L", "T"), ("R", "U", "N"), ("R", "L", "T"),
                                             ("R", "U", "T"), ("R", "L", "N")])
def test_tiled_trsm_branches(dm, side, uplo, trans):
    n, nb = 384, 128
    rng = np.random.default_rng(sum(map(ord, side + uplo + trans)))
    A = 0.05 * rng.standard_normal((n, n)) + np.diag(np.full(n, 2.0))
    B = rng.standard_normal((n, 200) if side == "L" else (200, n))
    DA = dm.DArray.from_array(A, dm.Blocks(nb, nb))
    DB = dm.DArray.from_array(B, dm.Blocks(nb, nb))
    dm.trsm_(side, uplo, trans, "N", 1.0, DA, DB)
    ref = O.trsm(side, uplo, trans, "N", 1.0, A, B, nb)
    assert np.allclose(DB.collect(), ref, rtol=1e-12, atol=1e-12)


def test_reference_datadeps_cholesky_test(dm):
    """test/datadeps.jl:946-973: a hand-written lower tiled Cholesky DAG
    (m = n = 1000, nb = 100) spawned with LAPACK.potrf!/BLAS.trsm!/syrk!/
    gemm!, compared with LAPACK.potrf!('L', dense)."""
    m = n = 1000
    nb = 100
    rng = np.random.default_rng(0)
    M = rng.random((m, n))
    M = M @ M.T
    expected, info = O.cholesky(M, m, "L")
    D = dm.DArray.from_array(M, dm.Blocks(nb, nb))
    Mc = D.chunks
    mt = nt = m // nb
    with dm.spawn_datadeps():
        for k in range(mt):
            dm.spawn(dm.LAPACK.potrf, "L", dm.InOut(Mc[k, k]))
            for _m in range(k + 1, mt):
                dm.spawn(dm.BLAS.trsm, "R", "L", "T", "N", 1.0, dm.In(Mc[k, k]), dm.InOut(Mc[_m, k]))
            for _n in range(k + 1, nt):
                dm.spawn(dm.BLAS.syrk, "L", "N", -1.0, dm.In(Mc[_n, k]), 1.0, dm.InOut(Mc[_n, _n]))
                for _m in range(_n + 1, mt):
                    dm.spawn(dm.BLAS.gemm, "N", "T", -1.0, dm.In(Mc[_m, k]), dm.In(Mc[_n, k]), 1.0,
                             dm.InOut(Mc[_m, _n]))
    got = D.collect()
    assert np.allclose(np.tril(got), np.tril(expected), rtol=1e-8, atol=1e-8)
    # isapprox (rtol = sqrt(eps)) as in the reference test; this input is ill-conditioned
    assert O.rel_factor_diff(got, expected, "L") < np.sqrt(np.finfo(float).eps)


def test_ishermitian_rules(dm):
    """src/array/linalg.jl:38-48,90-124: exact on diagonal tiles, sqrt(eps)
    Frobenius on off-diagonal pairs; cholesky! rejects with info = -1."""
    n, nb = 96, 32
    A = gen.spd(n, 60)
    assert dm.ishermitian(dm.DArray.from_array(A, dm.Blocks(nb, nb)))
    B = A.copy()
    B[40, 5] *= 1 + 1e-12          # off-diagonal pair, within sqrt(eps)
    assert dm.ishermitian(dm.DArray.from_array(B, dm.Blocks(nb, nb)))
    B[40, 5] *= 1 + 1e-3           # beyond sqrt(eps)
    assert not dm.ishermitian(dm.DArray.from_array(B, dm.Blocks(nb, nb)))
    C = A.copy()
    C[3, 7] += 1e-15 * C[3, 7] + 1e-300  # inside a diagonal tile: exact rule
    C[3, 7] = np.nextafter(C[3, 7], np.inf)
    assert not dm.ishermitian(dm.DArray.from_array(C, dm.Blocks(nb, nb)))
    # the reference's negative case: a random (non-symmetric) 32x32, Blocks(4,4)
    R = np.random.default_rng(0).random((32, 32))
    assert not dm.ishermitian(dm.DArray.from_array(R, dm.Blocks(4, 4)))
    with pytest.raises(dm.PosDefException) as e:
        dm.cholesky_(dm.DArray.from_array(R, dm.Blocks(4, 4)))
    assert e.value.info == -1
    F = dm.cholesky_(dm.DArray.from_array(R, dm.Blocks(4, 4)), check=False)
    assert F.info == -1
    # non-square blocks are re-blocked for the check (Blocks(32, 64) -> 32)
    assert dm.ishermitian(dm.DArray.from_array(A, dm.Blocks(32, 64)))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_ishermitian_matches_oracle_at_size(dm, dtype):
    """The device pre-check against the oracle's restatement
    (src/array/linalg.jl:38-48, 90-124) on an 8192 x 8192 matrix in
    Blocks(1024, 1024) (36 chunk pairs): symmetric, a far chunk pair
    perturbed at a quarter / four times its own threshold, one ulp off in
    the last diagonal chunk, and a NaN in a pair."""
    n, nb = 8192, 1024
    A = gen.spd(n, 61).astype(dtype)
    eps = np.finfo(dtype).eps
    i, j = 7 * nb + 100, 2 * nb + 300          # chunk pair (7, 2)
    a1 = A[7 * nb:8 * nb, 2 * nb:3 * nb].astype(np.float64)
    thr = np.sqrt(eps) * np.linalg.norm(a1)
    cases = [("symmetric", A, True)]
    for scale, want in ((0.25, True), (4.0, False)):
        B = A.copy()
        B[i, j] = dtype(float(B[i, j]) + scale * thr)
        cases.append((f"pair x{scale}", B, want))
    B = A.copy()
    B[n - 3, n - 1] = np.nextafter(B[n - 3, n - 1], np.inf)
    cases.append(("ulp in diagonal chunk", B, False))
    B = A.copy()
    B[i, j] = np.nan
    cases.append(("nan in pair", B, False))
    for name, M, want in cases:
        assert O.ishermitian(M, nb) == want, name
        D = dm.DArray.from_array(M, dm.Blocks(nb, nb))
        assert dm.ishermitian(D) == want, name
        D.unsafe_free()


@pytest.mark.parametrize("m,nb,nrhs", [(1024, 128, 1), (1000, 128, 17), (2048, 256, 32),
                                       (3000, 1024, 64)])
def test_tiled_gemm_block_column(dm, m, nb, nrhs):
    """C = beta C + alpha A X with X, C one block column (the refinement
    residual's streaming kernel), against numpy, ragged edges included."""
    import ctypes as C
    rng = np.random.default_rng(m + nrhs)
    A = rng.standard_normal((m, m))
    X = rng.standard_normal((m, nrhs))
    Y = rng.standard_normal((m, nrhs))
    DA = dm.DArray.from_array(A, dm.Blocks(nb, nb))
    DX = dm.DArray.from_array(X, dm.Blocks(nb, nb))
    DY = dm.DArray.from_array(Y, dm.Blocks(nb, nb))
    ta, tx, ty = DA.tiled(), DX.tiled(), DY.tiled()
    o = dm.PlanOptions().c()
    dm._lib.call("dm_gemm_tiled", b"N", b"N", -1.0, C.byref(ta), C.byref(tx), 0.5, C.byref(ty),
                 C.byref(o), None)
    ref = 0.5 * Y - A @ X
    err = np.abs(DY.collect() - ref).max() / np.abs(ref).max()
    assert err < 1e-13


def test_plans_on_user_streams_and_stream_release(dm):
    """Eager plans fork onto executor streams shared per caller stream and
    released with it: the same cached plan run from one user stream, then
    (after that stream is destroyed) from a new one, and with one stream
    directly on the caller's stream, gives the default stream's bits."""
    import ctypes as C
    from dagger_mi355x._lib import call
    n, nb = 1536, 256
    A = gen.spd(n, 21)
    _, want = factor_gpu(dm, A, nb, options=dm.PlanOptions(num_streams=2, batch=True,
                                                            lookahead=1, use_graph=False))
    DA = dm.DArray.from_array(A, dm.Blocks(nb, nb))
    info = dm.DeviceInt(2)
    for ns in (2, 3, 1):
        dag = dm.Dag.cholesky(DA, "U", info)
        dag.plan(dm.PlanOptions(num_streams=ns, batch=True, lookahead=1 if ns > 1 else 0,
                                use_graph=False))
        assert dag.stats()["graph"] == 0
        for rep in range(2):
            h = C.c_void_p()
            call("dm_stream_create", C.byref(h))
            DA.copy_from_host(A, h)
            dag.run(h)
            got = DA.collect(h)
            call("dm_stream_destroy", h)
            assert int(info.get()[0]) == 0
            assert np.array_equal(got, want), (ns, rep)
