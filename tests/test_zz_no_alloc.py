"""The allocation-free TRSM entry point (dm_dtrsm_noalloc) against dm_dtrsm
through the C ABI: same kernels, its workspace in the library's static
per-GPU scratch.  Results must be bitwise equal, on two calls in a row with
the second on a smaller input (the scratch keeps the first call's contents),
for every side/uplo/trans/diag and for solves longer than one recursion
leaf."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class Dev:
    def __init__(self, dm, a):
        self.a = np.asfortranarray(a)
        self.buf = dm.DeviceBuffer(max(self.a.nbytes, 8))
        dm._lib.call("dm_memcpy_h2d", C.c_void_p(self.buf.ptr), self.a.ctypes.data_as(C.c_void_p),
                     C.c_size_t(self.a.nbytes), None)

    @property
    def p(self):
        return C.c_void_p(self.buf.ptr)

    def get(self, dm):
        out = np.empty_like(self.a, order="F")
        dm._lib.call("dm_memcpy_d2h", out.ctypes.data_as(C.c_void_p), C.c_void_p(self.buf.ptr),
                     C.c_size_t(out.nbytes), None)
        return out


def _pair(dm, side, uplo, trans, diag, m, n, alpha, seed):
    """dm_dtrsm and dm_dtrsm_noalloc on copies of one input; both results."""
    rng = np.random.default_rng(seed)
    ka = m if side == "L" else n
    A = 0.1 * rng.standard_normal((ka, ka)) + np.diag(np.full(ka, 2.0))
    B = rng.standard_normal((m, n))
    dA, d1, d2 = Dev(dm, A), Dev(dm, B), Dev(dm, B)
    args = (side.encode(), uplo.encode(), trans.encode(), diag.encode(), m, n, alpha)
    lda, ldb = max(ka, 1), max(m, 1)   # BLAS: leading dimensions are at least 1
    dm._lib.call("dm_dtrsm", *args, dA.p, lda, d1.p, ldb, None)
    dm._lib.call("dm_dtrsm_noalloc", *args, dA.p, lda, d2.p, ldb, None)
    dm._lib.call("dm_stream_sync", None)
    return A, B, d1.get(dm), d2.get(dm)


@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("uplo", ["U", "L"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("diag", ["N", "U"])
def test_noalloc_trsm_equals_trsm_twice(dm, side, uplo, trans, diag):
    seed = sum(map(ord, side + uplo + trans + diag))
    # first call: longer than one recursion leaf (several leaves reuse the
    # scratch); second: a smaller input right after it
    for (m, n) in ((700, 96), (150, 33)) if side == "L" else ((96, 700), (33, 150)):
        A, B, r1, r2 = _pair(dm, side, uplo, trans, diag, m, n, -0.75, seed + m)
        assert np.array_equal(r1, r2), (side, uplo, trans, diag, m, n)
        T = np.triu(A) if uplo == "U" else np.tril(A)
        if diag == "U":
            T = T - np.diag(np.diag(T)) + np.eye(T.shape[0])
        op = T.T if trans == "T" else T
        ref = (np.linalg.solve(op, -0.75 * B) if side == "L"
               else np.linalg.solve(op.T, -0.75 * B.T).T)
        assert np.linalg.norm(r2 - ref) <= 1e-12 * np.linalg.norm(ref)


def test_noalloc_trsm_edge_cases(dm):
    """Empty shapes are no-ops, bad arguments are rejected before any device
    work (LAPACK xerbla numbering), and alpha = 0 zeroes B."""
    info = dm.DeviceBuffer(8)
    A, B, r1, r2 = _pair(dm, "L", "U", "N", "N", 0, 5, 1.0, 3)
    assert r1.size == 0 and r2.size == 0
    A, B, r1, r2 = _pair(dm, "L", "L", "N", "N", 64, 8, 0.0, 4)
    assert np.array_equal(r1, r2) and not r2.any()
    with pytest.raises(dm.DMError) as e:
        dm._lib.call("dm_dtrsm_noalloc", b"X", b"U", b"N", b"N", 4, 4, 1.0, C.c_void_p(info.ptr),
                     4, C.c_void_p(info.ptr), 4, None)
    assert e.value.status == -1
    with pytest.raises(dm.DMError) as e:
        dm._lib.call("dm_dtrsm_noalloc", b"L", b"U", b"N", b"N", 8, 4, 1.0, C.c_void_p(info.ptr),
                     4, C.c_void_p(info.ptr), 8, None)
    assert e.value.status == -9
