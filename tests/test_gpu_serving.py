"""GPU checks of the serving form: results are the same alone and in the
mix, reach the requested tolerance, and the HBM budget admission works."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _trace(dm, count, rate, seed, n_hi=1500):
    from dagger_mi355x import serving
    return serving.make_trace(rate=rate, count=count, seed=seed, n_lo=64, n_hi=n_hi,
                              k_lo=1e2, k_hi=1e6, reuse=0.3, tols=(1e-8, 1e-13), nrhs_max=8)


@pytest.mark.parametrize("mixed_min_n", [0, 4096])
def test_serving_mix_equals_alone_and_meets_tolerance(dm, mixed_min_n):
    """mixed_min_n 0: every request on the FP32 factor + refinement; 4096:
    every request of this trace (n <= 1500) on the FP64 direct path."""
    from dagger_mi355x import serving
    from oracle import gen
    trace = _trace(dm, 24, rate=400.0, seed=3)
    srv = serving.Server(workers=3, budget_bytes=8e9, mixed_min_n=mixed_min_n)
    stats = srv.run(trace)
    srv.close()
    assert stats["completed"] == len(trace)
    for r in trace:
        assert r.berr <= r.tol, (r.rid, r.berr, r.tol)
    # the same requests alone, one server per request: bitwise identical x
    for r in trace[:6]:
        alone = serving.Request(r.rid, r.mid, r.n, r.kappa, r.nrhs, r.tol, 0.0, False, r.seed)
        s1 = serving.Server(workers=1, budget_bytes=8e9, mixed_min_n=mixed_min_n)
        s1.run([alone])
        s1.close()
        assert np.array_equal(alone.x, r.x), r.rid
    # closed-form check of one solution: A x = b
    r = trace[0]
    h = gen.HouseholderSPD(r.n, r.kappa, 1_000_003 + r.mid, nrefl=64)
    b = r.rhs()
    res = b - h.matvec(r.x)
    assert np.abs(res).max() <= 10 * r.tol * (np.abs(h.dense()).sum(1).max() *
                                            np.abs(r.x).max() + np.abs(b).max())


def test_serving_budget_limits_resident_bytes(dm):
    from dagger_mi355x import serving
    trace = _trace(dm, 12, rate=1000.0, seed=4, n_hi=1024)
    budget = 3 * 1024 * 1024 * 12  # about three 1024-padded factors
    srv = serving.Server(workers=2, budget_bytes=budget)
    stats = srv.run(trace)
    assert srv.used <= budget
    srv.close()
    assert stats["completed"] == len(trace)
    assert all(r.berr <= r.tol for r in trace)


def test_server_runs_several_traces_and_keys_cache_by_matrix(dm):
    """A server serves traces one after another; matrix ids repeat across
    traces with other sizes, so cached factors are keyed by (id, n, kappa)."""
    from dagger_mi355x import serving
    srv = serving.Server(workers=2, budget_bytes=8e9)
    t1 = _trace(dm, 8, rate=500.0, seed=5, n_hi=700)
    t2 = _trace(dm, 8, rate=500.0, seed=6, n_hi=700)
    s1 = srv.run(t1)
    s2 = srv.run(t2)
    srv.close()
    assert s1["completed"] == len(t1) and s2["completed"] == len(t2)
    assert all(r.berr <= r.tol for r in t1 + t2)
