"""MI355X (gfx950) backend for Dagger.jl's tiled Cholesky / ldiv! path.

Python host mirror of the Julia-side API over the C ABI in
include/dagger_mi355x.h (libdagger_mi355x.so, hand-written HIP kernels):

    from dagger_mi355x import DArray, Blocks, cholesky_, ldiv_
    DA = DArray.from_array(A, Blocks(1024, 1024))   # tiles resident in HBM
    F = cholesky_(DA)                              # LinearAlgebra.cholesky!
    ldiv_(F, DB)                                   # LinearAlgebra.ldiv!

The library must be built (python dagger.jl_amd/build.py); nothing here
falls back to a CPU implementation.
"""
from . import _lib
from ._lib import DMError, load as load_library
from .darray import Blocks, DArray, DeviceBuffer, DeviceInt, Tile, zeros, randn, spd_generate
from .datadeps import (In, Out, InOut, BLAS, LAPACK, potrf_checked, spawn, spawn_datadeps,
                       PlanOptions, SEQUENTIAL, Dag)
from .linalg import (Cholesky, PosDefException, cholesky, cholesky_, ldiv_, ldiv, ldiv_triangular_, mul_, mul,
                     trsm_, trsv_,
                     ishermitian, norm2, chkfinite, DEFAULT_OPTIONS)
from .lu import LU, RowMaximum, NoPivot, SingularException, lu, lu_

__all__ = [
    "Blocks", "DArray", "DeviceBuffer", "DeviceInt", "Tile", "zeros", "randn", "spd_generate",
    "In", "Out", "InOut", "BLAS", "LAPACK", "potrf_checked", "spawn", "spawn_datadeps",
    "PlanOptions", "SEQUENTIAL", "Dag", "Cholesky", "PosDefException", "cholesky",
    "cholesky_", "ldiv_", "ldiv", "ldiv_triangular_", "mul_", "mul", "trsm_", "trsv_", "ishermitian", "norm2", "chkfinite",
    "DEFAULT_OPTIONS", "DMError", "LU", "RowMaximum", "NoPivot", "SingularException", "lu", "lu_",
    "load_library",
    "init",
]


def init(device=0):
    """Bind the library to a device (dm_init)."""
    _lib.call("dm_init", int(device))
