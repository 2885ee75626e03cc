"""GPU checks of the precision form (FP32 tiled factor + FP64 refinement)
against closed forms: A = H diag(s) H^T with kappa(A) = kappa exactly and
A^-1 = H diag(1/s) H^T (oracle/gen.HouseholderSPD)."""
import numpy as np
import pytest

from oracle import dagger_tiled as O
from oracle import gen

pytestmark = pytest.mark.gpu


def test_householder_device_matches_host(dm):
    from dagger_mi355x import mixed
    n = 512
    D, spec = mixed.householder_generate(n, dm.Blocks(128, 128), 1e4, seed=3, nrefl=16)
    host = gen.HouseholderSPD(n, 1e4, 3, nrefl=16)
    A = D.collect()
    assert np.allclose(A, host.dense(), rtol=0, atol=1e-13)
    assert np.allclose(np.linalg.eigvalsh((A + A.T) / 2)[[0, -1]], [1e-4, 1.0], rtol=1e-8)


@pytest.mark.parametrize("kappa,update", [(1e2, 0), (1e4, 0), (1e6, 0), (1e2, 6), (1e4, 6),
                                          (1e6, 6), (1e2, 3), (1e4, 3), (1e6, 3)])
@pytest.mark.parametrize("nrhs", [1, 32])
def test_mixed_ir_reaches_fp64_backward_error(dm, kappa, update, nrhs):
    """update: FP32 trailing-update arithmetic (0 FP32 MFMA, 6/3 BF16 MFMA
    with 3-/2-term operand splitting); refinement recovers FP64 accuracy."""
    from dagger_mi355x import mixed
    n, nb = 2048, 512
    D, spec = mixed.householder_generate(n, dm.Blocks(nb, nb), kappa, seed=11, nrefl=64)
    host = gen.HouseholderSPD(n, kappa, 11, nrefl=64)
    F = mixed.MixedCholesky(D, options=mixed._OPTS.replace(f32_update=update))
    xt = gen.randn(n * nrhs, 12).reshape(nrhs, n).T.copy()
    b = host.matvec(xt)
    DB = dm.DArray.from_array(b, dm.Blocks(nb, nb))
    X, info = F.solve(DB, tol=1e-13, max_iter=30)
    x = X.collect()
    # per-column backward error from the closed form
    A = host.dense()
    berr = O.backward_error_columns(A, x, b)
    assert berr <= 1e-13, (berr, info)
    # forward error bounded by kappa * backward error (with slack)
    ferr = np.abs(x - xt).max() / np.abs(xt).max()
    # (kappa is the 2-norm condition number; the inf-norm one can be larger by
    # a factor up to sqrt(n), hence the slack)
    assert ferr <= 100 * kappa * 1e-13 + 1e-12
    assert info["iterations"] <= 60
    assert info["berr"].max() <= 1e-13
    if update == 6 or kappa <= 1e4:
        assert not info["escalated"]


def test_split_update_factor_close_to_fp32(dm):
    """The FP32 factor built with BF16x6 updates is as close to the FP64
    factor as the FP32-MFMA one (both ~ kappa * 2^-24 relative)."""
    from dagger_mi355x import mixed
    n, nb = 2048, 512
    A = gen.spd(n, 31)
    U = np.linalg.cholesky(A).T
    errs = {}
    for update in (0, 6, 3):
        DA = dm.DArray.from_array(A, dm.Blocks(nb, nb))
        F = mixed.MixedCholesky(DA, options=mixed._OPTS.replace(f32_update=update))
        f = np.triu(F.F.collect().astype(np.float64))
        errs[update] = np.linalg.norm(f - U) / np.linalg.norm(U)
    assert errs[0] < 1e-6 and errs[6] < 1e-6, errs
    assert errs[3] < 1e-4, errs


def test_split_update_large_tiles_match_small_tiles(dm):
    """Launches of at least 256 workgroup tiles use 256x128 tiles (SYRK jobs
    as rectangles with the tiles outside the triangle skipped): the factor
    is bitwise the one of 128x128 tiles (same per-element product order),
    checked here against the same update arithmetic on FP32 MFMA."""
    from dagger_mi355x import mixed
    n, nb = 8192, 512  # step 0: 120 trailing tasks = 960 tiles of 256x128
    A = gen.spd(n, 41)
    f = {}
    for update in (0, 6):
        DA = dm.DArray.from_array(A, dm.Blocks(nb, nb))
        F = mixed.MixedCholesky(DA, options=mixed._OPTS.replace(f32_update=update))
        f[update] = np.triu(F.F.collect().astype(np.float64))
    assert np.linalg.norm(f[6] - f[0]) / np.linalg.norm(f[0]) < 1e-6
    # the strictly lower triangle of the FP32 copy keeps A (nothing written there)
    DA = dm.DArray.from_array(A, dm.Blocks(nb, nb))
    F = mixed.MixedCholesky(DA, options=mixed._OPTS.replace(f32_update=6))
    low = np.tril(F.F.collect(), -1)
    assert np.array_equal(low, np.tril(A.astype(np.float32), -1))


def test_mixed_matches_plain_fp64_solution(dm):
    from dagger_mi355x import mixed
    n, nb = 1024, 256
    A = gen.spd(n, 21)
    DA = dm.DArray.from_array(A, dm.Blocks(nb, nb))
    F = mixed.MixedCholesky(DA)
    b = gen.randn(n * 4, 22).reshape(4, n).T.copy()
    X, info = F.solve(dm.DArray.from_array(b, dm.Blocks(nb, nb)), tol=1e-14)
    DA2 = dm.DArray.from_array(A, dm.Blocks(nb, nb))
    F64 = dm.cholesky_(DA2)
    DB = dm.DArray.from_array(b, dm.Blocks(nb, nb))
    dm.ldiv_(F64, DB)
    assert np.allclose(X.collect(), DB.collect(), rtol=1e-12, atol=1e-13)
    assert info["berr"].max() <= 1e-14


@pytest.mark.parametrize("uplo", ["U", "L"])
def test_fp64_diagonal_potrf_of_fp32_plan(dm, uplo):
    """f64_diag: an FP32 plan factors its diagonal tiles in FP64.  The first
    diagonal tile has no updates, so its FP32 factor is the FP64 Cholesky of
    the FP32 tile rounded to FP32 (within one rounding of it); the whole
    factor still meets the FP32 parity bar against LAPACK in FP64."""
    import ctypes as C
    from dagger_mi355x.datadeps import PlanOptions
    from dagger_mi355x.darray import DeviceInt
    rng = np.random.default_rng(41)
    n, nb = 1024, 256
    G = rng.standard_normal((n, n))
    A32 = (G @ G.T / n + np.eye(n)).astype(np.float32)
    D = dm.DArray.from_array(A32, dm.Blocks(nb, nb))
    info, ginfo = DeviceInt(1), DeviceInt(1)
    o = PlanOptions(num_streams=2, batch=True, lookahead=1, f64_diag=True).c()
    t = D.tiled()
    dm._lib.call("dm_cholesky_tiled", uplo.encode(), C.byref(t), C.c_void_p(info.ptr),
                 C.c_void_p(ginfo.ptr), C.byref(o), None)
    assert int(info.get()[0]) == 0
    F = D.collect().astype(np.float64)
    T0 = A32[:nb, :nb].astype(np.float64)
    L0 = np.linalg.cholesky(T0)
    ref0 = L0.T if uplo == "U" else L0
    got0 = np.triu(F[:nb, :nb]) if uplo == "U" else np.tril(F[:nb, :nb])
    # one FP32 rounding of each entry (relative 2^-24, plus tiny entries)
    assert np.all(np.abs(got0 - ref0) <= (2.0 ** -24 + 1e-12) * np.abs(ref0) + 1e-30)
    ref = np.array(A32, dtype=np.float64, order="F")
    assert O.tile_potrf(uplo, ref) == 0
    assert O.rel_factor_diff(F, ref, uplo) <= 5e-5
