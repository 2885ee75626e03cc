"""Resident DArrays: Blocks partitions whose tiles live in HBM.

Mirrors Dagger's `Blocks`/`DArray` (src/array/darray.jl:107-152) and the
partition rule `_cumlength` (src/array/alloc.jl:23-28): an m x n array cut
into mb x nb tiles, edge tiles smaller.  Every tile (a Dagger "chunk") is a
dense column-major array whose leading dimension is its row count, exactly
like the per-chunk `Array`s Dagger builds (src/array/darray.jl:568-581).
Here all tiles of one DArray share one device allocation, packed in
column-major tile order:

    offset(i, j) = j*nb*m + i*mb*cols(j)      (elements)

so a tile task receives a resident device pointer and nothing is copied per
datadeps region (the host<->device moves of ext/ROCExt.jl:184-330 are gone).
"""
import ctypes as C
import weakref

import numpy as np

from . import _lib
from ._lib import call

_DTYPES = {np.dtype(np.float64): _lib.DM_F64, np.dtype(np.float32): _lib.DM_F32,
           np.dtype(np.complex128): _lib.DM_C64, np.dtype(np.complex64): _lib.DM_C32}


class Blocks:
    """Block partitioning, `Blocks(mb, nb)` / `Blocks(mb)` for vectors."""

    def __init__(self, *blocksize):
        if not blocksize or any(int(b) <= 0 for b in blocksize):
            raise ValueError("block sizes must be positive")
        self.blocksize = tuple(int(b) for b in blocksize)

    def __eq__(self, other):
        return isinstance(other, Blocks) and self.blocksize == other.blocksize

    def __hash__(self):
        return hash(self.blocksize)

    def __repr__(self):
        return f"Blocks{self.blocksize}"


def _cumlength(length, step):
    """Tile extents along one axis (src/array/alloc.jl:23-28)."""
    nice, extra = divmod(length, step)
    return [step] * nice + ([extra] if extra else [])


class DeviceBuffer:
    """An owned device allocation (dm_malloc / dm_free)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        call("dm_malloc", C.byref(p), C.c_size_t(self.nbytes))
        self.ptr = p.value or 0
        self._fin = weakref.finalize(self, DeviceBuffer._release, self.ptr)

    @staticmethod
    def _release(ptr):
        if ptr:
            lib = _lib.load()
            lib.dm_forget_matrix(C.c_void_p(ptr))
            lib.dm_free(C.c_void_p(ptr))

    def free(self):
        """unsafe_free! (src/memory-spaces.jl:660-671)."""
        self._fin()
        self.ptr = 0


class Scratch:
    """Named device buffers reused across calls on one stream.  A caller
    that stages many transfers or fills through one Scratch allocates and
    frees nothing once warm; per-call temporaries would not do for
    concurrent streams, because hipFree waits for the whole device.  A
    buffer grows to the next power of two (at least 1 MiB) when a larger one
    is asked for; the one it replaces is kept until free(), so growing never
    waits for the device either."""

    def __init__(self):
        self._bufs = {}
        self._retired = []

    def get(self, name, nbytes):
        b = self._bufs.get(name)
        if b is None or b.nbytes < nbytes:
            if b is not None:
                self._retired.append(b)  # queued work on the stream may still use it
            b = DeviceBuffer(1 << max(20, (max(int(nbytes), 1) - 1).bit_length()))
            self._bufs[name] = b
        return b

    def nbytes(self):
        return sum(b.nbytes for b in self._bufs.values()) + sum(b.nbytes for b in self._retired)

    def free(self):
        """Release every buffer (the caller has synchronized the stream)."""
        for b in list(self._bufs.values()) + self._retired:
            b.free()
        self._bufs.clear()
        self._retired = []


class _Temp:
    """A temporary device buffer: from `scratch` when given (kept), else
    allocated here and freed on exit."""

    def __init__(self, scratch, name, nbytes):
        self.scratch, self.name, self.nbytes = scratch, name, nbytes

    def __enter__(self):
        if self.scratch is not None:
            self.buf = self.scratch.get(self.name, self.nbytes)
        else:
            self.buf = DeviceBuffer(self.nbytes)
        return self.buf

    def __exit__(self, *exc):
        if self.scratch is None:
            self.buf.free()
        return False


class DeviceInt:
    """A device int32 array (LAPACK info / flags)."""

    def __init__(self, count=1, value=0):
        self.count = count
        self.buf = DeviceBuffer(4 * count)
        self.fill(value)

    @property
    def ptr(self):
        return self.buf.ptr

    def fill(self, value, stream=None):
        call("dm_fill_i32", C.c_void_p(self.ptr), int(value), C.c_size_t(self.count), stream)

    def free(self):
        self.buf.free()

    def get(self, stream=None):
        out = np.zeros(self.count, dtype=np.int32)
        call("dm_memcpy_d2h", out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr),
             C.c_size_t(4 * self.count), stream)
        return out


class Tile:
    """One chunk of a DArray: rows x cols, column-major, ld = rows."""

    __slots__ = ("ptr", "rows", "cols", "ld", "dtype", "owner", "index")

    def __init__(self, ptr, rows, cols, ld, dtype, owner=None, index=None):
        self.ptr, self.rows, self.cols, self.ld = ptr, rows, cols, ld
        self.dtype = np.dtype(dtype)
        self.owner = owner
        self.index = index

    @property
    def shape(self):
        return (self.rows, self.cols)

    def __repr__(self):
        return f"Tile{self.index}({self.rows}x{self.cols}, {self.dtype})"


class DArray:
    """A 1-D or 2-D array partitioned by `Blocks`, resident in HBM."""

    def __init__(self, dtype, shape, blocks, _buffer=None):
        dtype = np.dtype(dtype)
        if dtype not in _DTYPES:
            raise TypeError(f"unsupported element type {dtype}")
        if isinstance(shape, int):
            shape = (shape,)
        if len(shape) not in (1, 2):
            raise ValueError("DArray supports vectors and matrices")
        if len(blocks.blocksize) != len(shape):
            raise ValueError("Blocks dimensionality does not match the array")
        self.dtype = dtype
        self.shape = tuple(int(s) for s in shape)
        self.partitioning = blocks
        self.m = self.shape[0]
        self.n = self.shape[1] if len(self.shape) == 2 else 1
        self.mb = blocks.blocksize[0]
        self.nb = blocks.blocksize[1] if len(self.shape) == 2 else 1
        self.row_ext = _cumlength(self.m, self.mb)
        self.col_ext = _cumlength(self.n, self.nb)
        nbytes = self.m * self.n * dtype.itemsize
        self.buffer = _buffer if _buffer is not None else DeviceBuffer(max(nbytes, 8))
        self._chunks = None

    # ------------------------------------------------------------ layout
    @property
    def ndim(self):
        return len(self.shape)

    @property
    def mt(self):
        return len(self.row_ext)

    @property
    def nt(self):
        return len(self.col_ext)

    def tile_offset(self, i, j):
        """Element offset of tile (i, j) in the packed allocation."""
        return j * self.nb * self.m + i * self.mb * self.col_ext[j]

    def tile(self, i, j=0):
        rows, cols = self.row_ext[i], self.col_ext[j]
        ptr = self.buffer.ptr + self.tile_offset(i, j) * self.dtype.itemsize
        return Tile(ptr, rows, cols, rows, self.dtype, owner=self, index=(i, j))

    @property
    def chunks(self):
        """Tiles as a grid (`A.chunks` in Dagger): chunks[i, j] or chunks[i]."""
        if self._chunks is None:
            if self.ndim == 1:
                g = np.empty(self.mt, dtype=object)
                for i in range(self.mt):
                    g[i] = self.tile(i, 0)
            else:
                g = np.empty((self.mt, self.nt), dtype=object)
                for i in range(self.mt):
                    for j in range(self.nt):
                        g[i, j] = self.tile(i, j)
            self._chunks = g
        return self._chunks

    def tiled(self):
        """The dm_tiled descriptor of this array (vectors as m x 1)."""
        return _lib.dm_tiled(C.c_void_p(self.buffer.ptr), self.m, self.n, self.mb, self.nb,
                             _DTYPES[self.dtype])

    # ------------------------------------------------------------ data
    @classmethod
    def from_array(cls, A, blocks, stream=None):
        """`DArray(A, Blocks(...))` / `distribute` (src/array/darray.jl:605-621)."""
        A = np.asarray(A)
        if A.dtype not in _DTYPES:
            A = A.astype(np.float64)
        D = cls(A.dtype, A.shape, blocks)
        D.copy_from_host(A, stream)
        return D

    def copy_from_host(self, A, stream=None, scratch=None):
        """`copyto!(DA, A)` (src/array/copy.jl:146-151): one H2D + pack kernel.
        `scratch` (a Scratch) provides the staging buffer."""
        A = np.asfortranarray(np.asarray(A, dtype=self.dtype).reshape(self.m, self.n, order="F")
                              if self.ndim == 1 else np.asarray(A, dtype=self.dtype))
        if A.shape != (self.m, self.n):
            raise ValueError("shape mismatch")
        if self.m * self.n == 0:
            return self
        with _Temp(scratch, "stage", A.nbytes) as tmp:
            call("dm_memcpy_h2d", C.c_void_p(tmp.ptr), A.ctypes.data_as(C.c_void_p),
                 C.c_size_t(A.nbytes), stream)
            t = self.tiled()
            call("dm_tiled_pack", C.byref(t), C.c_void_p(tmp.ptr), self.m, 0.0, stream)
            call("dm_stream_sync", stream)
        return self

    def collect(self, stream=None, scratch=None):
        """`collect(DA)` (src/array/darray.jl:236-271): unpack kernel + D2H.
        `scratch` (a Scratch) provides the staging buffer."""
        out = np.empty((self.m, self.n), dtype=self.dtype, order="F")
        if self.m * self.n:
            with _Temp(scratch, "stage", out.nbytes) as tmp:
                t = self.tiled()
                call("dm_tiled_unpack", C.c_void_p(tmp.ptr), self.m, C.byref(t), stream)
                call("dm_memcpy_d2h", out.ctypes.data_as(C.c_void_p), C.c_void_p(tmp.ptr),
                     C.c_size_t(out.nbytes), stream)
        return out.reshape(self.m) if self.ndim == 1 else out

    def copy(self, stream=None):
        D = DArray(self.dtype, self.shape, self.partitioning)
        call("dm_memcpy_d2d", C.c_void_p(D.buffer.ptr), C.c_void_p(self.buffer.ptr),
             C.c_size_t(self.m * self.n * self.dtype.itemsize), stream)
        return D

    def reblock(self, blocks, stream=None):
        """A copy partitioned by `blocks` (the buffer of copy_buffered,
        src/array/copy.jl:11-37): unpack to a full matrix, pack again."""
        D = DArray(self.dtype, self.shape, blocks)
        if self.m * self.n:
            tmp = DeviceBuffer(self.m * self.n * self.dtype.itemsize)
            t = self.tiled()
            call("dm_tiled_unpack", C.c_void_p(tmp.ptr), self.m, C.byref(t), stream)
            t2 = D.tiled()
            call("dm_tiled_pack", C.byref(t2), C.c_void_p(tmp.ptr), self.m, 0.0, stream)
            call("dm_stream_sync", stream)
            tmp.free()
        return D

    def copy_into(self, other, stream=None):
        """copyto!(other, self) between DArrays of equal shape (any blocks)."""
        if other.shape != self.shape or other.dtype != self.dtype:
            raise ValueError("shape/dtype mismatch")
        if other.partitioning == self.partitioning:
            call("dm_memcpy_d2d", C.c_void_p(other.buffer.ptr), C.c_void_p(self.buffer.ptr),
                 C.c_size_t(self.m * self.n * self.dtype.itemsize), stream)
            return other
        tmp = DeviceBuffer(self.m * self.n * self.dtype.itemsize)
        t = self.tiled()
        call("dm_tiled_unpack", C.c_void_p(tmp.ptr), self.m, C.byref(t), stream)
        t2 = other.tiled()
        call("dm_tiled_pack", C.byref(t2), C.c_void_p(tmp.ptr), self.m, 0.0, stream)
        call("dm_stream_sync", stream)
        tmp.free()
        return other

    def unsafe_free(self):
        self.buffer.free()

    def __repr__(self):
        return f"DArray{{{self.dtype}}}({self.shape}, {self.partitioning})"


def zeros(blocks, dtype, *shape):
    """`zeros(Blocks(...), T, dims...)` allocated in HBM (src/array/alloc.jl:61-121)."""
    D = DArray(dtype, shape if len(shape) > 1 else shape[0], blocks)
    nbytes = D.m * D.n * D.dtype.itemsize
    if nbytes:
        z = np.zeros(D.m * D.n, dtype=D.dtype)
        call("dm_memcpy_h2d", C.c_void_p(D.buffer.ptr), z.ctypes.data_as(C.c_void_p),
             C.c_size_t(nbytes), None)
        call("dm_stream_sync", None)
    return D


def randn(blocks, m, n, seed, offset=0, stream=None):
    """`randn(Blocks(...), m, n)` generated in HBM: element (i, j) is the
    Philox/Box-Muller variate with index offset + i + j*m (dm_fill_randn_d)."""
    D = DArray(np.float64, (m, n) if n is not None else m, blocks)
    count = D.m * D.n
    if count:
        F = DeviceBuffer(count * 8)
        call("dm_fill_randn_d", C.c_void_p(F.ptr), C.c_size_t(count), C.c_uint64(seed),
             C.c_uint64(offset), stream)
        t = D.tiled()
        call("dm_tiled_pack", C.byref(t), C.c_void_p(F.ptr), D.m, 0.0, stream)
        call("dm_stream_sync", stream)
        F.free()
    return D


def spd_generate(n, blocks, seed, dtype=np.float64, stream=None):
    """A = G G^T / n + I with G an n x n matrix of Philox-Box-Muller N(0,1)
    variates (dm_fill_randn_d), built entirely on the device: G in a full
    column-major buffer, one FP64 GEMM, then the tile pack adds I."""
    if np.dtype(dtype) != np.float64:
        raise TypeError("the device generator produces FP64")
    G = DeviceBuffer(n * n * 8)
    call("dm_fill_randn_d", C.c_void_p(G.ptr), C.c_size_t(n * n), C.c_uint64(seed),
         C.c_uint64(0), stream)
    F = DeviceBuffer(n * n * 8)
    call("dm_dgemm", b"N", b"T", n, n, n, 1.0 / n, C.c_void_p(G.ptr), n, C.c_void_p(G.ptr), n,
         0.0, C.c_void_p(F.ptr), n, stream)
    G.free()
    D = DArray(np.float64, (n, n), blocks)
    t = D.tiled()
    call("dm_tiled_pack", C.byref(t), C.c_void_p(F.ptr), n, 1.0, stream)
    call("dm_stream_sync", stream)
    F.free()
    return D
