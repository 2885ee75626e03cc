"""GPU checks of the batching form: one persistent launch factoring and
solving many uneven SPD systems, checked per system against LAPACK
(the oracle's dense tile routines) and the solve residual."""
import numpy as np
import pytest

from oracle import dagger_tiled as O

pytestmark = pytest.mark.gpu


def _check(bs, A0s, B0s, rtol):
    infos = bs.infos()
    for m in range(bs.count):
        A0, B0 = A0s[m], B0s[m]
        n = A0.shape[0]
        F, X = bs.get_system(m)
        assert infos[m] == 0, (m, infos[m])
        ref = np.array(A0, dtype=np.float64, order="F")
        info = O.tile_potrf("U", ref)
        assert info == 0
        assert O.rel_factor_diff(F.astype(np.float64), ref, "U") <= rtol, m
        if B0.shape[1]:
            A64 = A0.astype(np.float64)
            be = O.backward_error(A64, X.astype(np.float64), B0.astype(np.float64))
            assert be <= (10 * n * np.finfo(A0.dtype).eps), (m, be)


@pytest.mark.parametrize("dtype,rtol", [(np.float64, 1e-12), (np.float32, 5e-5)])
def test_batched_posv_uneven(dm, dtype, rtol):
    from dagger_mi355x import batched
    sizes, nrhs = batched.log_uniform_sizes(48, 32, 700, seed=5)
    sizes[:4] = [32, 33, 64, 700]     # edges: smallest, ragged panel, exact panel, largest
    bs = batched.BatchedSystems(sizes, nrhs, dtype).generate(seed=100)
    A0s, B0s = zip(*[bs.get_system(m) for m in range(bs.count)])
    bs.posv(panels={})
    _check(bs, A0s, B0s, rtol)


def test_batched_panels_agree_and_autotune(dm):
    from dagger_mi355x import batched
    sizes, nrhs = batched.log_uniform_sizes(24, 32, 400, seed=6)
    bs = batched.BatchedSystems(sizes, nrhs, np.float64).generate(seed=7)
    snap = bs.snapshot()
    outs = []
    for P in batched.PANELS:
        bs.restore(snap)
        bs.posv(panels={b: P for b in range(16)})
        outs.append([bs.get_system(m)[1] for m in range(bs.count)])
    for o in outs[1:]:
        for a, b in zip(o, outs[0]):
            assert np.allclose(a, b, rtol=1e-10, atol=1e-12)
    bs.restore(snap)
    choice, table = batched.autotune(bs, sample=8, reps=1)
    assert set(choice) == {batched.bucket_of(int(n)) for n in sizes}
    assert all(p in batched.PANELS for p in choice.values())


def test_batched_non_pd_system_is_isolated(dm):
    from dagger_mi355x import batched
    sizes = np.array([100, 200, 150])
    nrhs = np.array([2, 2, 2])
    bs = batched.BatchedSystems(sizes, nrhs).generate(seed=9)
    A1, _ = bs.get_system(1)
    A1[120, 120] = -1.0
    bs.put_system(1, A=A1)
    A0s, B0s = zip(*[bs.get_system(m) for m in range(bs.count)])
    bs.posv(panels={})
    infos = bs.infos()
    ref = np.array(A1, order="F")
    assert infos[1] == O.tile_potrf("U", ref) == 121
    assert infos[0] == 0 and infos[2] == 0
    for m in (0, 2):
        F, X = bs.get_system(m)
        assert O.backward_error(A0s[m], X, B0s[m]) <= 10 * sizes[m] * np.finfo(float).eps


@pytest.mark.parametrize("dtype,rtol", [(np.float64, 1e-12), (np.float32, 5e-5)])
@pytest.mark.parametrize("nb", [64, 128, 256])
def test_batched_tiled_hybrid(dm, dtype, rtol, nb):
    """Systems of at least 300 rows through the tiled level-synchronous
    plan (ragged last tiles at every nb), the rest through the persistent
    kernel; each factor against LAPACK and each solve's backward error."""
    from dagger_mi355x import batched
    sizes, nrhs = batched.log_uniform_sizes(40, 32, 1100, seed=31)
    sizes[:4] = [1100, 300, nb + 1, 2 * nb]
    bs = batched.BatchedSystems(sizes, nrhs, dtype).generate(seed=32)
    A0s, B0s = zip(*[bs.get_system(m) for m in range(bs.count)])
    bs.solve(panels={}, tiled_min=300, nb=nb)
    _check(bs, A0s, B0s, rtol)
    bs.close()


def test_batched_tiled_failed_system_keeps_b(dm):
    """A non-positive pivot in a later tile of one tiled system: LAPACK's
    info for it, its B left as it was, every other system solved."""
    import ctypes as C
    from dagger_mi355x import batched
    sizes = np.array([700, 520, 900, 64])
    nrhs = np.array([3, 2, 8, 1])
    bs = batched.BatchedSystems(sizes, nrhs).generate(seed=33)
    A1, _ = bs.get_system(1)
    A1[400, 400] = -2.0
    bs.put_system(1, A=A1)
    A0s, B0s = zip(*[bs.get_system(m) for m in range(bs.count)])
    bs.solve(panels={}, tiled_min=500, nb=128)
    infos = bs.infos()
    ref = np.array(A1, order="F")
    assert infos[1] == O.tile_potrf("U", ref) == 401
    _, X1 = bs.get_system(1)
    assert np.array_equal(X1, B0s[1])
    for m in (0, 2, 3):
        assert infos[m] == 0
        _, X = bs.get_system(m)
        assert O.backward_error(A0s[m], X, B0s[m]) <= 10 * sizes[m] * np.finfo(float).eps
    bs.close()


def test_batched_tiled_per_bucket_tiles_and_autotune(dm):
    """Per-bucket outer tiles (one plan mixing 64, 192 and 384), checked
    per system against LAPACK; the tile autotuner returns one candidate per
    bucket and the plan it picks solves the batch."""
    from dagger_mi355x import batched
    sizes, nrhs = batched.log_uniform_sizes(60, 32, 1500, seed=71)
    bs = batched.BatchedSystems(sizes, nrhs).generate(seed=72)
    A0s, B0s = zip(*[bs.get_system(m) for m in range(bs.count)])
    tiles = {b: (64, 192, 384)[b % 3] for b in range(5, 11)}
    bs.solve(panels={}, tile_nb=tiles)
    _check(bs, A0s, B0s, 1e-12)
    choice, table = batched.autotune_tiled(bs, sample=6, reps=1)
    assert set(choice) == {batched.bucket_of(int(n)) for n in sizes}
    assert all(v in batched.TILE_CANDIDATES for v in choice.values())
    for m in range(bs.count):
        bs.put_system(m, A=A0s[m], B=B0s[m])
    bs.solve(panels={}, tile_nb=choice)
    _check(bs, A0s, B0s, 1e-12)
    bs.close()
