"""CPU restatement of the device input generators (test infrastructure).

TEST INFRASTRUCTURE ONLY (tests/, smoke(), bench.py's cpu_baseline leg).

* philox4x32_10: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel
  random numbers: as easy as 1, 2, 3", SC'11), the integer stream behind
  dm_fill_randn_d: counter = (i lo32, i hi32, 0, 0), key = (seed lo32,
  seed hi32).  The integer outputs must match the device bit for bit.
* randn: two 53-bit uniforms in (0,1) from the four words, Box-Muller
  z = sqrt(-2 ln u) cos(2 pi v).  Equal to the device up to the last-ulp
  differences of log/cos between libm implementations.
* spd(n, seed): A = G G^T / n + I (the BASELINE generator for the plain
  Cholesky configurations).
"""
import numpy as np

_M0, _M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_W0, _W1 = np.uint32(0x9E3779B9), np.uint32(0xBB67AE85)
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10_block(c0, c1, c2, c3, k0, k1):
    """Philox4x32-10 on explicit counter words and key (vectorised)."""
    c0, c1, c2, c3 = (np.asarray(c, dtype=np.uint64) & _MASK for c in (c0, c1, c2, c3))
    k0, k1 = int(k0) & 0xFFFFFFFF, int(k1) & 0xFFFFFFFF
    for _ in range(10):
        p0 = _M0 * c0
        p1 = _M1 * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK
        n0 = hi1 ^ c1 ^ np.uint64(k0)
        n2 = hi0 ^ c3 ^ np.uint64(k1)
        c0, c1, c2, c3 = n0, lo1, n2, lo0
        k0 = (k0 + int(_W0)) & 0xFFFFFFFF
        k1 = (k1 + int(_W1)) & 0xFFFFFFFF
    return tuple(c.astype(np.uint32) for c in (c0, c1, c2, c3))


def philox4x32_10(idx, seed):
    """Four uint32 words for counter (idx lo, idx hi, 0, 0), key = seed."""
    idx = np.asarray(idx, dtype=np.uint64)
    z = np.zeros_like(idx)
    return philox4x32_10_block(idx & _MASK, idx >> np.uint64(32), z, z, seed & 0xFFFFFFFF,
                               (seed >> 32) & 0xFFFFFFFF)


def randn(count, seed, offset=0):
    idx = np.arange(offset, offset + count, dtype=np.uint64)
    w0, w1, w2, w3 = philox4x32_10(idx, seed)
    two53 = 9007199254740992.0
    u = ((w0 >> 5).astype(np.float64) * 67108864.0 + (w1 >> 6).astype(np.float64) + 0.5) / two53
    v = ((w2 >> 5).astype(np.float64) * 67108864.0 + (w3 >> 6).astype(np.float64) + 0.5) / two53
    return np.sqrt(-2.0 * np.log(u)) * np.cos(6.283185307179586 * v)


def gaussian_matrix(n, seed, m=None):
    """n x m (default n x n) column-major G, element (i, j) = randn[i + j*n]."""
    m = n if m is None else m
    return randn(n * m, seed).reshape(m, n).T.copy(order="F")


def spd(n, seed):
    """A = G G^T / n + I."""
    G = gaussian_matrix(n, seed)
    A = G @ G.T / n
    A[np.diag_indices(n)] += 1.0
    return A
