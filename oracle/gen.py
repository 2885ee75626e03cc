"""CPU restatement of the device input generators (test infrastructure).

TEST INFRASTRUCTURE ONLY (tests/, smoke(), bench.py's cpu_baseline leg).

* philox4x32_10: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel
  random numbers: as easy as 1, 2, 3", SC'11), the integer stream behind
  dm_fill_randn_d: counter = (i lo32, i hi32, 0, 0), key = (seed lo32,
  seed hi32).  The integer outputs must match the device bit for bit.
* randn: two 53-bit uniforms in (0,1) from the four words, Box-Muller
  z = sqrt(-2 ln u) cos(2 pi v).  Equal to the device up to the last-ulp
  differences of log/cos between libm implementations.
* spd(n, seed): A = G G^T / n + I (the input generator of the plain
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


class HouseholderSPD:
    """A = H diag(s) H^T with H = H_1 H_2 ... H_r, H_i = I - tau_i v_i v_i^T,
    v_i seeded Gaussian vectors (Philox stream `seed`, vector i at offset
    i*n), s log-spaced from 1 down to 1/kappa, so kappa_2(A) = kappa exactly
    and A^-1 = H diag(1/s) H^T.

    Compact WY (LAPACK dlarft, forward/columnwise): H = I - V T V^T with T
    upper triangular.  Then A = S - Q V^T - V Q^T with
      Y = S V T^T,  Z = T V^T Y (symmetric),  Q = Y - V Z / 2,
    i.e. a rank-2r update of the diagonal: A = S - U W^T, U = [Q V], W = [V Q].
    Products with A and A^-1 cost O(n r) per vector (no n x n matrix)."""

    def __init__(self, n, kappa, seed, nrefl=64):
        self.n, self.kappa, self.r = n, float(kappa), nrefl
        V = randn(n * nrefl, seed).reshape(nrefl, n).T.copy()
        self.V = V
        tau = 2.0 / np.einsum("ij,ij->j", V, V)
        T = np.zeros((nrefl, nrefl))
        for i in range(nrefl):
            T[i, i] = tau[i]
            if i:
                T[:i, i] = -tau[i] * (T[:i, :i] @ (V[:, :i].T @ V[:, i]))
        self.T = T
        self.s = self.kappa ** (-np.arange(n) / max(n - 1, 1))
        Y = self.s[:, None] * (V @ T.T)
        Z = T @ (V.T @ Y)
        Q = Y - 0.5 * (V @ Z)
        self.U = np.ascontiguousarray(np.hstack([Q, V]))
        self.W = np.ascontiguousarray(np.hstack([V, Q]))

    def H(self, x):
        """H x = x - V T V^T x."""
        return x - self.V @ (self.T @ (self.V.T @ x))

    def Ht(self, x):
        return x - self.V @ (self.T.T @ (self.V.T @ x))

    def matvec(self, x):
        """A x = s * x - U (W^T x)."""
        x2 = x.reshape(self.n, -1)
        return (self.s[:, None] * x2 - self.U @ (self.W.T @ x2)).reshape(x.shape)

    def solve(self, b):
        """A^-1 b = H diag(1/s) H^T b."""
        b2 = b.reshape(self.n, -1)
        return self.H(self.Ht(b2) / self.s[:, None]).reshape(b.shape)

    def dense(self):
        return np.diag(self.s) - self.U @ self.W.T


def spd(n, seed):
    """A = G G^T / n + I."""
    G = gaussian_matrix(n, seed)
    A = G @ G.T / n
    A[np.diag_indices(n)] += 1.0
    return A
