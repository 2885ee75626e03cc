"""GPU parity of the device norm2 / chkfinite pass (dm_tiled_norm2) against
the oracle's restatement of src/array/linalg.jl:1-33 and :127-148.

Tolerance: both sides accumulate squares in FP64 in different orders, so
the norms agree to a relative 1e-13 (FP64 and FP32 inputs alike)."""
import numpy as np
import pytest

from oracle import dagger_tiled as O

pytestmark = pytest.mark.gpu

RTOL = 1e-13


@pytest.mark.parametrize("m,n,nb", [(1000, 1000, 128), (300, 300, 64), (257, 129, 32),
                                    (4096, 4096, 1024), (5, 5, 8)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_norm2_matches_oracle(dm, m, n, nb, dtype):
    rng = np.random.default_rng(m + 3 * n + nb)
    A = rng.standard_normal((m, n)).astype(dtype)
    DA = dm.DArray.from_array(A, dm.Blocks(nb, nb))
    A64 = A.astype(np.float64)
    assert abs(dm.norm2(DA) - O.norm2(A64, nb)) <= RTOL * O.norm2(A64, nb)
    if m == n:
        for part in ("upper", "lower"):
            want = O.norm2(A64, nb, part)
            assert abs(dm.norm2(DA, part) - want) <= RTOL * want, part
    assert dm.chkfinite(DA)


def test_norm2_vector_and_empty(dm):
    x = np.random.default_rng(2).standard_normal(1000)
    Dx = dm.DArray.from_array(x, dm.Blocks(96))
    assert abs(dm.norm2(Dx) - np.linalg.norm(x)) <= RTOL * np.linalg.norm(x)
    E = dm.DArray(np.float64, (0, 0), dm.Blocks(4, 4))
    assert dm.norm2(E) == 0.0
    assert dm.chkfinite(E)


@pytest.mark.parametrize("where", [(0, 0), (129, 69), (64, 33)])
@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf])
def test_chkfinite_finds_inf_and_nan(dm, where, bad):
    A = np.ones((130, 70))
    A[where] = bad
    DA = dm.DArray.from_array(A, dm.Blocks(32, 32))
    with pytest.raises(ValueError, match="Inf or NaN"):
        O.chkfinite(A, 32)
    with pytest.raises(ValueError, match="Inf or NaN"):
        dm.chkfinite(DA)


def test_norm2_device_pass_is_deterministic(dm):
    A = np.random.default_rng(9).standard_normal((2048, 2048))
    DA = dm.DArray.from_array(A, dm.Blocks(256, 256))
    vals = {dm.norm2(DA) for _ in range(3)}
    assert len(vals) == 1
