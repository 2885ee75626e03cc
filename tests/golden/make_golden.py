"""Writes tests/golden/cholesky_golden.npz: small seeded SPD systems with
their dense LAPACK factor and solve (scipy's OpenBLAS dpotrf/dpotrs), made
in this container.  The reference ships no golden vectors for this path
(its tests compare against LAPACK at run time), so these fixtures are the
same comparison, frozen.  Shapes follow the reference's tests:
  64 x 64, Blocks(16,16)   test/mpi_cholesky.jl:8-23
  128 x 128, Blocks(32,32) test/array/linalg/cholesky.jl:8-12
  128, Blocks(64,128)      test/array/linalg/solve.jl:4-7 (re-blocked)
plus ragged and degenerate sizes.  Inputs use the Philox/Box-Muller
generator of oracle/gen.py: A = G G^T / n + I.

Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np
import scipy.linalg.lapack as lapack

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import gen  # noqa: E402

CASES = [  # name, n, nb, uplo, seed, nrhs
    ("n64_b16_U", 64, 16, "U", 42, 16),
    ("n64_b16_L", 64, 16, "L", 42, 16),
    ("n128_b32_U", 128, 32, "U", 7, 8),
    ("n128_b32_L", 128, 32, "L", 7, 1),
    ("n100_b32_U", 100, 32, "U", 11, 5),
    ("n1_b1_U", 1, 1, "U", 3, 1),
    ("n7_b3_L", 7, 3, "L", 5, 2),
]


def main():
    out = {}
    for name, n, nb, uplo, seed, nrhs in CASES:
        A = gen.spd(n, seed)
        F, info = lapack.dpotrf(A, lower=1 if uplo == "L" else 0, clean=0)
        assert info == 0
        B = gen.randn(n * nrhs, seed + 1000).reshape(nrhs, n).T.copy()
        X, info2 = lapack.dpotrs(F, B, lower=1 if uplo == "L" else 0)
        assert info2 == 0
        out[f"{name}/A"] = A
        out[f"{name}/F"] = F
        out[f"{name}/B"] = B
        out[f"{name}/X"] = X
        out[f"{name}/meta"] = np.array([n, nb, 1 if uplo == "L" else 0, seed, nrhs])
    # non positive definite: the reference's own construction (A[1,1] = -100)
    A = gen.spd(128, 9)
    A[0, 0] = -100.0
    out["npd_first/A"] = A
    # failure inside the third diagonal tile (global row 70, Blocks(32,32))
    A = gen.spd(128, 10)
    A[70, 70] = -5.0
    out["npd_tile3/A"] = A
    out["npd_tile3/dense_info"] = np.array([lapack.dpotrf(A, lower=0, clean=0)[1]])
    np.savez_compressed(os.path.join(HERE, "cholesky_golden.npz"), **out)
    print("wrote", len(out), "arrays")


if __name__ == "__main__":
    main()
