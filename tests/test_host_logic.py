"""Host-side logic of the product package that needs no GPU: task
translation, partitioning, serving traces and the batched size model."""
import numpy as np
import pytest

from oracle import dagger_tiled as O


def test_cumlength_matches_reference_partition():
    from dagger_mi355x.darray import _cumlength
    # src/array/alloc.jl:23-28
    assert _cumlength(10, 3) == [3, 3, 3, 1]
    assert _cumlength(9, 3) == [3, 3, 3]
    assert _cumlength(2, 5) == [2]
    assert _cumlength(32768, 1024) == [1024] * 32
    assert O.cumlength(1000, 128) == _cumlength(1000, 128)


def test_make_task_dimensions_follow_blas_calls():
    import dagger_mi355x as d
    from dagger_mi355x.datadeps import make_task
    T = lambda r, c: d.Tile(1 << 30, r, c, r, np.float64)  # noqa: E731
    t = make_task(d.BLAS.gemm, ("T", "N", -1.0, d.In(T(64, 32)), d.In(T(64, 48)), 1.0,
                                d.InOut(T(32, 48))))
    assert (t.m, t.n, t.k) == (32, 48, 64) and t.transa == b"T"
    t = make_task(d.BLAS.syrk, ("U", "T", -1.0, d.In(T(16, 40)), 1.0, d.InOut(T(40, 40))))
    assert (t.n, t.k) == (40, 16)
    t = make_task(d.BLAS.syrk, ("L", "N", -1.0, d.In(T(40, 16)), 1.0, d.InOut(T(40, 40))))
    assert (t.n, t.k) == (40, 16)
    t = make_task(d.BLAS.trsm, ("R", "L", "T", "N", 1.0, d.In(T(20, 20)), d.InOut(T(7, 20))))
    assert (t.m, t.n) == (7, 20) and t.access[0] == 1 and t.access[2] == 3
    # bare arguments count as In; potrf defaults its tile to InOut
    t = make_task(d.LAPACK.potrf, ("U", T(8, 8)))
    assert t.access[2] == 3 and t.n == 8
    with pytest.raises(TypeError):
        make_task(d.BLAS.gemm, ("N", "N", 1.0, np.zeros((2, 2)), T(2, 2), 0.0, T(2, 2)))


@pytest.mark.parametrize("f,args", [
    # gemm: B's tile is not k x n (k = 64 from op(A) = A^T)
    ("gemm", ("T", "N", -1.0, "A64x32", "B48x48", 1.0, "C32x48")),
    # gemm: C's rows differ from op(A)'s
    ("gemm", ("N", "N", 1.0, "A64x32", "B32x48", 0.0, "C32x48")),
    # gemm: transposed B with the wrong inner size
    ("gemm", ("N", "T", 1.0, "A32x16", "B48x8", 0.0, "C32x48")),
    # syrk: C not op(A)'s order
    ("syrk", ("U", "T", -1.0, "A16x40", 1.0, "C32x32")),
    # trsm: A not square of B's row count (side L) / column count (side R)
    ("trsm", ("L", "U", "T", "N", 1.0, "A20x20", "B7x20")),
    ("trsm", ("R", "L", "T", "N", 1.0, "A7x7", "B7x20")),
    # gemv: x length differs from op(A)'s columns, y from its rows
    ("gemv", ("N", -1.0, "A32x16", "x32x1", 1.0, "y32x1")),
    ("gemv", ("T", -1.0, "A32x16", "x32x1", 1.0, "y32x1")),
    # trsv: x not of A's order
    ("trsv", ("U", "N", "N", "A32x32", "x16x1")),
    ("potrf", ("U", "A8x4")),
])
def test_make_task_rejects_mismatched_tiles(f, args):
    """BLAS.gemm!/syrk!/trsm!/gemv!/trsv! and LAPACK.potrf! throw
    DimensionMismatch on inconsistent operands; a task with such operands
    must never reach the executor, whose kernels trust the shapes."""
    import re as _re
    import dagger_mi355x as d
    from dagger_mi355x.datadeps import make_task

    def conv(a):
        if isinstance(a, str) and _re.fullmatch(r"[A-Za-z]\d+x\d+", a):
            r, c = map(int, a[1:].split("x"))
            return d.In(d.Tile(1 << 30, r, c, r, np.float64))
        return a
    fn = d.LAPACK.potrf if f == "potrf" else getattr(d.BLAS, f)
    with pytest.raises(ValueError, match="DimensionMismatch"):
        make_task(fn, tuple(conv(a) for a in args))


def test_serving_trace_model():
    from dagger_mi355x import serving
    tr = serving.make_trace(rate=500.0, count=4000, seed=1)
    t = np.array([r.t_arrive for r in tr])
    gaps = np.diff(t)
    assert abs(gaps.mean() - 1 / 500) < 0.1 / 500          # Poisson rate
    n = np.array([r.n for r in tr])
    assert n.min() >= 64 and n.max() <= 8192
    ln = np.log(n[[not r.reuse for r in tr]])
    assert abs(ln.mean() - (np.log(64) + np.log(8193)) / 2) < 0.1   # log-uniform
    reuse = np.mean([r.reuse for r in tr])
    assert 0.25 < reuse < 0.35
    assert {r.tol for r in tr} == {1e-8, 1e-13}
    assert all(1 <= r.nrhs <= 32 for r in tr)
    # reuse points at an earlier matrix with the same size and kappa
    first = {}
    for r in tr:
        if r.mid in first:
            assert first[r.mid] == (r.n, r.kappa)
        else:
            assert not r.reuse
            first[r.mid] = (r.n, r.kappa)
    # determinism
    tr2 = serving.make_trace(rate=500.0, count=4000, seed=1)
    assert [(r.n, r.mid, r.nrhs) for r in tr] == [(r.n, r.mid, r.nrhs) for r in tr2]


def test_serving_padding_ladder():
    from dagger_mi355x import serving
    for n in (1, 64, 65, 1000, 1025, 4097, 8192):
        p = serving.pad_size(n)
        assert p >= n and p in serving.LADDER
        nb = serving.tile_for(p)
        assert p % nb == 0
    assert serving.pad_size(8193) == 9216


def test_batched_size_model_and_buckets():
    from dagger_mi355x import batched
    sizes, nrhs = batched.log_uniform_sizes(4096, 32, 2048, seed=2)
    assert sizes.min() >= 32 and sizes.max() <= 2048
    assert abs(np.log(sizes).mean() - (np.log(32) + np.log(2049)) / 2) < 0.05
    assert nrhs.min() >= 1 and nrhs.max() <= 8
    flops = (sizes.astype(float) ** 3 / 3).sum()
    assert 5e11 < flops < 1.5e12   # ~9.4e11 expected (E[n^3] over LogUniform[32,2048])
    assert batched.bucket_of(32) == 5 and batched.bucket_of(2047) == 10
    assert batched.bucket_of(2048) == 11
