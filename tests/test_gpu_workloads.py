"""The precision, batching and serving forms at their full workload sizes,
checked against closed forms and per-system LAPACK:

* mixed precision + refinement: A = H diag(s) H^T (64 Householder
  reflectors, kappa exact) at N = 16384 (Blocks(512, 512)) for kappa in
  {1e2, 1e4, 1e6} with 1 and 32 right-hand sides, at N = 32768
  (Blocks(1024, 1024)) for kappa = 1e6 with 32, and kappa = 1e8 at
  N = 16384; every column's backward error ||b - A x||_inf /
  (||A||_inf ||x||_inf + ||b||_inf) is computed on the host from the
  generator's algebra (A x = s x - U W^T x), never from the device;
* batched: 4096 systems G G^T / n + I with n log-uniform in [32, 2048] and
  1-8 right-hand sides, FP64 and FP32, systems of n >= 512 through the
  tiled level-synchronous plan and the rest in one persistent launch (and
  FP64 all in the persistent launch); every system's backward error
  <= 10 n eps, and the factor of one system per floor(log2 n) bucket
  against LAPACK potrf;
* serving: a Poisson trace of 64 requests, n log-uniform in [64, 8192],
  1-32 right-hand sides, tolerances 1e-8 and 1e-13, 30% reuse, 200 GB
  budget, requests on both the FP32-factor and the FP64 paths; every
  request meets its tolerance (closed form), and results alone equal the
  results in the mix bitwise on a sample holding both paths.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import dagger_tiled as O
from oracle import gen

pytestmark = pytest.mark.gpu


def _norm_inf(h, chunk=1024):
    """||A||_inf of the Householder matrix from its algebra, by row blocks
    (A[R, :] = diag(s)[R, :] - U[R] W^T) so no n x n array is formed."""
    best = 0.0
    for r0 in range(0, h.n, chunk):
        rows = slice(r0, min(r0 + chunk, h.n))
        blk = -(h.U[rows] @ h.W.T)
        idx = np.arange(rows.start, rows.stop)
        blk[idx - r0, idx] += h.s[rows]
        best = max(best, float(np.abs(blk).sum(1).max()))
    return best


def _closed_form_berr(h, normA, x, b):
    r = b - h.matvec(x)
    x2, b2, r2 = (v.reshape(h.n, -1) for v in (x, b, r))
    return np.abs(r2).max(0) / (normA * np.abs(x2).max(0) + np.abs(b2).max(0))


def _mixed_case(dm, n, nb, kappa, nrhs_list, seed, options=None, force_gmres=False):
    from dagger_mi355x import mixed
    D, _ = mixed.householder_generate(n, dm.Blocks(nb, nb), kappa, seed=seed, nrefl=64)
    h = gen.HouseholderSPD(n, kappa, seed, nrefl=64)
    normA = _norm_inf(h)
    M = mixed.MixedCholesky(D, options=options)
    out = []
    for nrhs in nrhs_list:
        xt = gen.randn(n * nrhs, seed + nrhs).reshape(nrhs, n).T.copy()
        b = h.matvec(xt)
        DB = dm.DArray.from_array(b, dm.Blocks(nb, nb))
        X, info = M.solve(DB, tol=1e-13, max_iter=30, force_gmres=force_gmres)
        x = X.collect()
        berr = _closed_form_berr(h, normA, x, b)
        out.append((nrhs, berr, info, x, xt))
        DB.unsafe_free()
    mixed.release_factor_buffers(M.F)
    M.F.unsafe_free()
    D.unsafe_free()
    return out


@pytest.mark.parametrize("kappa", [1e2, 1e4, 1e6])
def test_mixed_n16384_nb512(dm, kappa):
    """N = 16384, Blocks(512, 512), 1 and 32 right-hand sides: refinement on
    the FP32 factor reaches FP64 backward error 1e-13 in every column."""
    for nrhs, berr, info, x, xt in _mixed_case(dm, 16384, 512, kappa, (1, 32), 70 + int(np.log10(kappa))):
        assert berr.shape == (nrhs,)
        assert berr.max() <= 1e-13, (kappa, nrhs, berr.max(), info["method"], info["iterations"])
        # forward error within the condition number's reach
        ferr = np.abs(x - xt).max() / np.abs(xt).max()
        assert ferr <= 100 * kappa * 1e-13 + 1e-12, (kappa, nrhs, ferr)


def test_mixed_n32768_nb1024_kappa1e6(dm):
    """The headline precision case: N = 32768, Blocks(1024, 1024),
    kappa = 1e6, 32 right-hand sides."""
    [(nrhs, berr, info, x, xt)] = _mixed_case(dm, 32768, 1024, 1e6, (32,), 91)
    assert berr.max() <= 1e-13, (berr.max(), info["method"], info["iterations"])
    assert info["method"] == "ir"


def test_mixed_kappa1e8_reports_method_and_meets_tolerance(dm):
    """kappa = 1e8 at N = 16384 (kappa u32 about 6): the FP32 copy needs the
    scaled diagonal shift to factor, classic refinement on the shifted
    factor converges too slowly or stalls, and GMRES-IR (FP64 GMRES,
    shifted FP32 factor as preconditioner) reaches 1e-13."""
    [(nrhs, berr, info, x, xt)] = _mixed_case(dm, 16384, 512, 1e8, (1,), 88)
    assert berr.max() <= 1e-13, (berr.max(), info)
    assert info["method"] in ("ir", "gmres-ir")
    assert info["shift"] in (0.0, 4.0, 32.0, 256.0)


def test_mixed_gmres_ir_meets_tolerance(dm):
    """GMRES-IR itself (FP64 GMRES, the FP32 factor as preconditioner, all
    right-hand sides in lockstep with device-resident Krylov blocks) at
    kappa = 1e8, N = 16384, 1 and 32 right-hand sides: refinement skipped
    after the first solve; every column meets 1e-13."""
    for nrhs, berr, info, x, xt in _mixed_case(dm, 16384, 512, 1e8, (1, 32), 88,
                                               force_gmres=True):
        assert info["method"] == "gmres-ir", info
        assert berr.shape == (nrhs,)
        assert berr.max() <= 1e-13, (nrhs, berr.max(), info)


def test_mixed_shift_when_the_fp32_copy_is_singular(dm):
    """An SPD A whose FP32 copy is exactly singular: I with the 2 x 2 block
    [1, 1 - 1e-9; 1 - 1e-9, 1] at rows 300, 301 (eigenvalues 2 - 1e-9 and
    1e-9).  1 - 1e-9 rounds to 1 in FP32, so the factorization of the copy
    meets the pivot 1 - 1 = 0 in any arithmetic; the scaled diagonal shift
    (diagonal times 1 + 4 u32) makes the pivot about 4.8e-7, and refinement
    against the unshifted FP64 A (kappa 2e9; classic refinement stalls on
    the 1e-9 direction, GMRES-IR resolves it) reaches 1e-13."""
    from dagger_mi355x import mixed
    n, nb = 1024, 512
    A = np.eye(n)
    A[300, 301] = A[301, 300] = 1.0 - 1e-9
    assert np.float32(A[300, 301]) == np.float32(1.0)
    xt = gen.randn(n, 93).reshape(n, 1)
    b = A @ xt
    D = dm.DArray.from_array(A, dm.Blocks(nb, nb))
    DB = dm.DArray.from_array(b, dm.Blocks(nb, nb))
    M = mixed.MixedCholesky(D)
    assert M.shift == 4.0, M.shift
    X, info = M.solve(DB, tol=1e-13, max_iter=30)
    x = X.collect()
    berr = np.abs(b - A @ x).max() / (np.abs(A).sum(1).max() * np.abs(x).max() + np.abs(b).max())
    assert info["shift"] == 4.0 and info["method"] == "gmres-ir", info
    assert berr <= 1e-13, (berr, info)
    mixed.release_factor_buffers(M.F)
    for d in (M.F, DB, D):
        d.unsafe_free()


def _batched_check(dm, dtype, count=4096, seed=2024, mode="hybrid"):
    from dagger_mi355x import batched
    sizes, nrhs = batched.log_uniform_sizes(count, 32, 2048, seed)
    bs = batched.BatchedSystems(sizes, nrhs, dtype).generate(seed=seed + 1)
    es = bs.dtype.itemsize
    A0 = np.empty(bs.a_elems, dtype=dtype)
    B0 = np.empty(bs.b_elems, dtype=dtype)
    dm._lib.call("dm_memcpy_d2h", A0.ctypes.data_as(C.c_void_p), C.c_void_p(bs.A.ptr),
                 C.c_size_t(bs.a_elems * es), None)
    dm._lib.call("dm_memcpy_d2h", B0.ctypes.data_as(C.c_void_p), C.c_void_p(bs.B.ptr),
                 C.c_size_t(bs.b_elems * es), None)
    panels, _ = batched.autotune(bs, sample=8, reps=1)
    if mode == "hybrid":
        bs.solve(panels=panels)
        bs.close()
    else:
        bs.posv(panels=panels)
    infos = bs.infos()
    assert (infos == 0).all(), np.nonzero(infos)[0][:10]
    X = np.empty(bs.b_elems, dtype=dtype)
    dm._lib.call("dm_memcpy_d2h", X.ctypes.data_as(C.c_void_p), C.c_void_p(bs.B.ptr),
                 C.c_size_t(bs.b_elems * es), None)
    eps = np.finfo(dtype).eps
    worst = 0.0
    for m in range(count):
        n, r = int(sizes[m]), int(nrhs[m])
        b0 = int(bs.b_off[m])
        A = bs.host_matrix(A0, m).astype(np.float64)
        Bm = B0[b0:b0 + n * r].reshape(n, r, order="F").astype(np.float64)
        Xm = X[b0:b0 + n * r].reshape(n, r, order="F").astype(np.float64)
        be = O.backward_error(A, Xm, Bm)
        worst = max(worst, be / (n * eps))
        assert be <= 10 * n * eps, (m, n, r, be)
    # factors against LAPACK: one system per floor(log2 n) bucket
    rtol = 1e-12 if dtype == np.float64 else 5e-5
    seen = set()
    for m in range(count):
        bkt = batched.bucket_of(int(sizes[m]))
        if bkt in seen:
            continue
        seen.add(bkt)
        n = int(sizes[m])
        ref = np.array(bs.host_matrix(A0, m), dtype=np.float64, order="F")
        assert O.tile_potrf("U", ref) == 0
        F, _ = bs.get_system(m)
        assert O.rel_factor_diff(F.astype(np.float64), ref, "U") <= rtol, (m, n)
    assert seen == {batched.bucket_of(int(n)) for n in sizes}
    bs.A.free()
    bs.B.free()
    return worst


@pytest.mark.parametrize("dtype,mode", [(np.float64, "hybrid"), (np.float32, "hybrid"),
                                        (np.float64, "persistent")])
def test_batched_4096_uneven_systems(dm, dtype, mode):
    """hybrid: n >= 512 through the tiled level-synchronous plan, the rest
    in one persistent launch; persistent: every system in that launch."""
    worst = _batched_check(dm, dtype, mode=mode)
    assert worst <= 10


def test_serving_trace_full_range(dm):
    """64 requests, n in [64, 8192], nrhs in [1, 32], both tolerances, 30%
    reuse, 200 GB budget, 1000 req/s: every request meets its tolerance by
    the closed form, requests ran on both paths, and a sample of both paths
    gives bitwise the same x alone as in the mix."""
    from dagger_mi355x import serving
    trace = serving.make_trace(rate=1000.0, count=64, seed=77)
    srv = serving.Server(workers=6, budget_bytes=200e9)
    stats = srv.run(trace)
    srv.close()
    assert stats["completed"] == len(trace)
    mixed_reqs = [r for r in trace if serving.pad_size(r.n) >= srv.mixed_min_n]
    direct_reqs = [r for r in trace if serving.pad_size(r.n) < srv.mixed_min_n]
    assert mixed_reqs and direct_reqs
    assert max(r.n for r in trace) > 4096 and max(r.nrhs for r in trace) > 16
    assert {r.tol for r in trace} == {1e-8, 1e-13}
    norms = {}
    for r in trace:
        key = (r.mid, r.n, r.kappa)
        if key not in norms:
            h = gen.HouseholderSPD(r.n, r.kappa, 1_000_003 + r.mid, nrefl=64)
            norms[key] = (h, _norm_inf(h))
        h, na = norms[key]
        berr = _closed_form_berr(h, na, r.x, r.rhs())
        assert berr.max() <= r.tol, (r.rid, r.n, r.kappa, r.tol, berr.max())
    sample = sorted(mixed_reqs, key=lambda r: r.n)[-2:] + direct_reqs[:2]
    for r in sample:
        alone = serving.Request(r.rid, r.mid, r.n, r.kappa, r.nrhs, r.tol, 0.0, False, r.seed)
        s1 = serving.Server(workers=1, budget_bytes=200e9)
        s1.run([alone])
        s1.close()
        assert np.array_equal(alone.x, r.x), (r.rid, r.n)
