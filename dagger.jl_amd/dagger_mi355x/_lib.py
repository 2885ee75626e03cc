"""ctypes binding of libdagger_mi355x.so (include/dagger_mi355x.h).

This is the Python form of the `ccall` layer a Julia extension would use.
There is no fallback: if the shared library is missing or does not load,
every entry point raises.
"""
import ctypes as C
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libdagger_mi355x.so")

DM_F64, DM_F32, DM_C64, DM_C32 = 0, 1, 2, 3
DM_OP_POTRF, DM_OP_TRSM, DM_OP_SYRK, DM_OP_GEMM, DM_OP_TRSV, DM_OP_GEMV = 1, 2, 3, 4, 5, 6
DM_OP_GETRF, DM_OP_LASWP = 7, 8
DM_IN, DM_OUT, DM_INOUT = 1, 2, 3
DM_HIP_ERROR_BASE = 1000

c_char = C.c_char
vp = C.c_void_p
i32 = C.c_int
i64 = C.c_int64
f64 = C.c_double
f32 = C.c_float
sz = C.c_size_t


class dm_task(C.Structure):
    _fields_ = [
        ("op", i32), ("dtype", i32),
        ("side", c_char), ("uplo", c_char), ("transa", c_char), ("transb", c_char),
        ("diag", c_char), ("pad_", c_char * 3),
        ("m", i32), ("n", i32), ("k", i32),
        ("alpha", f64), ("beta", f64),
        ("a", vp), ("lda", i32),
        ("b", vp), ("ldb", i32),
        ("c", vp), ("ldc", i32),
        ("info", vp),
        ("access", i32 * 3),
    ]


class dm_plan_opts(C.Structure):
    _fields_ = [("num_streams", i32), ("batch", i32), ("lookahead", i32), ("use_graph", i32),
                ("profile", i32), ("f32_update", i32), ("f64_diag", i32)]


class dm_group_profile(C.Structure):
    _fields_ = [("op", i32), ("ntasks", i32), ("stream", i32), ("level", i32), ("flops", f64),
                ("ms", f32), ("variant", i32)]


class dm_dag_stats(C.Structure):
    _fields_ = [("num_tasks", i32), ("num_edges", i32), ("num_groups", i32),
                ("num_levels", i32), ("num_launches", i32), ("num_streams", i32),
                ("graph", i32)]


class dm_tiled(C.Structure):
    _fields_ = [("base", vp), ("m", i64), ("n", i64), ("mb", i32), ("nb", i32), ("dtype", i32)]


# name -> (restype, argtypes)
_PROTOS = {
    "dm_init": (i32, [i32]),
    "dm_device_count": (i32, [C.POINTER(i32)]),
    "dm_last_error_string": (C.c_char_p, []),
    "dm_version": (i32, []),
    "dm_malloc": (i32, [C.POINTER(vp), sz]),
    "dm_free": (i32, [vp]),
    "dm_memcpy_h2d": (i32, [vp, vp, sz, vp]),
    "dm_memcpy_d2h": (i32, [vp, vp, sz, vp]),
    "dm_memcpy_d2d": (i32, [vp, vp, sz, vp]),
    "dm_memcpy2d_h2d": (i32, [vp, sz, vp, sz, sz, sz, vp]),
    "dm_memcpy2d_d2h": (i32, [vp, sz, vp, sz, sz, sz, vp]),
    "dm_memcpy2d_d2d": (i32, [vp, sz, vp, sz, sz, sz, vp]),
    "dm_fill_i32": (i32, [vp, i32, sz, vp]),
    "dm_stream_create": (i32, [C.POINTER(vp)]),
    "dm_stream_destroy": (i32, [vp]),
    "dm_stream_sync": (i32, [vp]),
    "dm_device_sync": (i32, []),
    "dm_event_create": (i32, [C.POINTER(vp)]),
    "dm_event_destroy": (i32, [vp]),
    "dm_event_record": (i32, [vp, vp]),
    "dm_stream_wait_event": (i32, [vp, vp]),
    "dm_event_sync": (i32, [vp]),
    "dm_event_elapsed_ms": (i32, [C.POINTER(f32), vp, vp]),
    "dm_graph_begin": (i32, [vp]),
    "dm_graph_end": (i32, [vp, C.POINTER(vp)]),
    "dm_graph_launch": (i32, [vp, vp]),
    "dm_graph_destroy": (i32, [vp]),
    "dm_dpotrf": (i32, [c_char, i32, vp, i32, vp, vp]),
    "dm_spotrf": (i32, [c_char, i32, vp, i32, vp, vp]),
    "dm_dtrsm": (i32, [c_char, c_char, c_char, c_char, i32, i32, f64, vp, i32, vp, i32, vp]),
    "dm_strsm": (i32, [c_char, c_char, c_char, c_char, i32, i32, f32, vp, i32, vp, i32, vp]),
    "dm_dtrsm_noalloc": (i32, [c_char, c_char, c_char, c_char, i32, i32, f64, vp, i32, vp, i32,
                               vp]),
    "dm_dsyrk": (i32, [c_char, c_char, i32, i32, f64, vp, i32, f64, vp, i32, vp]),
    "dm_ssyrk": (i32, [c_char, c_char, i32, i32, f32, vp, i32, f32, vp, i32, vp]),
    "dm_dgemm": (i32, [c_char, c_char, i32, i32, i32, f64, vp, i32, vp, i32, f64, vp, i32, vp]),
    "dm_zpotrf": (i32, [c_char, i32, vp, i32, vp, vp]),
    "dm_cpotrf": (i32, [c_char, i32, vp, i32, vp, vp]),
    "dm_ztrsm": (i32, [c_char, c_char, c_char, c_char, i32, i32, vp, vp, i32, vp, i32, vp]),
    "dm_ctrsm": (i32, [c_char, c_char, c_char, c_char, i32, i32, vp, vp, i32, vp, i32, vp]),
    "dm_zherk": (i32, [c_char, c_char, i32, i32, f64, vp, i32, f64, vp, i32, vp]),
    "dm_cherk": (i32, [c_char, c_char, i32, i32, f32, vp, i32, f32, vp, i32, vp]),
    "dm_zgemm": (i32, [c_char, c_char, i32, i32, i32, vp, vp, i32, vp, i32, vp, vp, i32, vp]),
    "dm_cgemm": (i32, [c_char, c_char, i32, i32, i32, vp, vp, i32, vp, i32, vp, vp, i32, vp]),
    "dm_sgemm": (i32, [c_char, c_char, i32, i32, i32, f32, vp, i32, vp, i32, f32, vp, i32, vp]),
    "dm_sgemm_bf16x": (i32, [i32, c_char, c_char, i32, i32, i32, f32, vp, i32, vp, i32, f32, vp,
                             i32, vp]),
    "dm_dtrsv": (i32, [c_char, c_char, c_char, i32, vp, i32, vp, i32, vp]),
    "dm_dgemv": (i32, [c_char, i32, i32, f64, vp, i32, vp, i32, f64, vp, i32, vp]),
    "dm_dag_create": (i32, [C.POINTER(vp)]),
    "dm_dag_destroy": (i32, [vp]),
    "dm_dag_add_task": (i32, [vp, C.POINTER(dm_task)]),
    "dm_dag_edges": (i32, [vp, C.POINTER(i32), C.POINTER(i32)]),
    "dm_dag_plan": (i32, [vp, C.POINTER(dm_plan_opts)]),
    "dm_dag_run": (i32, [vp, vp]),
    "dm_dag_get_stats": (i32, [vp, C.POINTER(dm_dag_stats)]),
    "dm_dag_profile": (i32, [vp, C.POINTER(dm_group_profile), C.POINTER(i32)]),
    "dm_dag_build_cholesky": (i32, [C.POINTER(vp), c_char, C.POINTER(dm_tiled), vp]),
    "dm_getrf_tiled": (i32, [i32, C.POINTER(dm_tiled), vp, vp, C.POINTER(dm_plan_opts), vp]),
    "dm_tiled_laswp": (i32, [C.POINTER(dm_tiled), vp, i32, vp]),
    "dm_dag_build_potrs": (i32, [C.POINTER(vp), c_char, C.POINTER(dm_tiled), C.POINTER(dm_tiled)]),
    "dm_cholesky_tiled": (i32, [c_char, C.POINTER(dm_tiled), vp, vp, C.POINTER(dm_plan_opts), vp]),
    "dm_potrs_tiled": (i32, [c_char, C.POINTER(dm_tiled), C.POINTER(dm_tiled),
                             C.POINTER(dm_plan_opts), vp]),
    "dm_trsm_tiled": (i32, [c_char, c_char, c_char, c_char, f64, C.POINTER(dm_tiled),
                            C.POINTER(dm_tiled), C.POINTER(dm_plan_opts), vp]),
    "dm_ishermitian_tiled": (i32, [C.POINTER(dm_tiled), vp, vp, vp]),
    "dm_tiled_norm2": (i32, [C.POINTER(dm_tiled), i32, vp, vp, vp, vp]),
    "dm_batched_posv": (i32, [i32, i32, vp, vp, vp, i32, vp, vp, i32, i32, vp]),
    "dm_batched_tiled_create": (i32, [C.POINTER(vp), i32, i32, vp, vp, vp, i32, vp, i32, vp]),
    "dm_batched_tiled_run": (i32, [vp, vp, i32, vp]),
    "dm_batched_tiled_destroy": (i32, [vp]),
    "dm_gemm_tiled": (i32, [c_char, c_char, f64, C.POINTER(dm_tiled), C.POINTER(dm_tiled), f64,
                            C.POINTER(dm_tiled), C.POINTER(dm_plan_opts), vp]),
    "dm_tiled_convert": (i32, [C.POINTER(dm_tiled), C.POINTER(dm_tiled), vp]),
    "dm_tiled_norm_inf": (i32, [C.POINTER(dm_tiled), vp, vp, vp]),
    "dm_tiled_identity": (i32, [C.POINTER(dm_tiled), vp]),
    "dm_tiled_col_norm_inf": (i32, [C.POINTER(dm_tiled), vp, vp]),
    "dm_tiled_col_dot": (i32, [C.POINTER(dm_tiled), C.POINTER(dm_tiled), vp, vp]),
    "dm_tiled_col_axpby": (i32, [vp, C.c_double, C.POINTER(dm_tiled), C.c_double,
                                 C.POINTER(dm_tiled), vp]),
    "dm_tiled_axpy_s2d": (i32, [C.POINTER(dm_tiled), C.POINTER(dm_tiled), vp]),
    "dm_tiled_scale_diag": (i32, [C.POINTER(dm_tiled), f64, vp]),
    "dm_add_diag_d": (i32, [vp, i64, i64, vp, vp]),
    "dm_tiled_pack": (i32, [C.POINTER(dm_tiled), vp, i64, f64, vp]),
    "dm_tiled_unpack": (i32, [vp, i64, C.POINTER(dm_tiled), vp]),
    "dm_plan_cache_clear": (i32, []),
    "dm_forget_matrix": (i32, [vp]),
    "dm_fill_randn_d": (i32, [vp, sz, C.c_uint64, C.c_uint64, vp]),
}


class DMError(RuntimeError):
    """Raised when a library call returns a non-zero status."""

    def __init__(self, name, status, msg):
        super().__init__(f"{name} returned {status}: {msg}")
        self.status = status


_lib = None
_lock = threading.Lock()


def load(path=None):
    """Load the shared library (once).  Raises if it is missing."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        p = path or LIB_PATH
        if not os.path.exists(p):
            raise FileNotFoundError(
                f"{p} not found: build it with `python dagger.jl_amd/build.py`")
        lib = C.CDLL(p)
        for name, (res, args) in _PROTOS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def exported_symbols():
    return list(_PROTOS)


def call(name, *args):
    """Call a dm_* entry point; raise DMError on non-zero status."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.dm_last_error_string()
        raise DMError(name, rc, msg.decode() if msg else "")
    return rc


def ch(c):
    return c.encode() if isinstance(c, str) else c
