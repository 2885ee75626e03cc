"""The Julia glue in INTEGRATION.md binds the C ABI correctly: every
`ccall((:dm_..., LIB), Ret, (ArgTypes...), args...)` names an entry point
declared in include/dagger_mi355x.h, passes as many argument types as the
declaration has parameters, and each Julia type matches the C type (scalars
exactly, pointers by level and pointee).  Julia is not installed here, so
this is the check the glue gets short of running it."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dagger_mi355x.h")
DOC = os.path.join(ROOT, "INTEGRATION.md")

# pointer typedefs of the header (opaque handles)
HANDLES = {"dm_stream_t", "dm_event_t", "dm_graph_t", "dm_dag_t", "dm_batch_plan_t"}
SCALARS = {"Cint": "int", "Cchar": "char", "Cdouble": "double", "Cfloat": "float",
           "Csize_t": "size_t", "UInt64": "uint64_t", "Int64": "int64_t"}
POINTEES = {"Cint": "int", "Float64": "double", "F64": "double", "Float32": "float",
            "F32": "float", "Cdouble": "double", "Cfloat": "float", "DMTiled": "dm_tiled",
            "ComplexF64": "double", "ComplexF32": "float",  # {re, im} scalars by reference
            "DMPlanOpts": "dm_plan_opts", "DMTask": "dm_task"}


def _split_top(s, sep=","):
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "({[":
            depth += 1
        elif ch in ")}]":
            depth -= 1
        if ch == sep and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


def header_decls():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    decls = {}
    for m in re.finditer(r"\b(int|const char\*)\s+(dm_\w+)\s*\(([^)]*)\)\s*;", txt):
        ret, name, params = m.group(1), m.group(2), m.group(3).strip()
        ps = [] if params in ("", "void") else [p.strip() for p in params.split(",")]
        types = []
        for p in ps:
            p = re.sub(r"\s+", " ", p)
            t = re.sub(r"\s*\b\w+$", "", p) if re.search(r"[\w*]\s+\w+$", p) else p
            types.append(t.replace("const ", "").replace(" *", "*").strip())
        decls[name] = (ret, types)
    return decls


def glue_ccalls():
    txt = open(DOC).read()
    blocks = re.findall(r"```julia\n(.*?)```", txt, flags=re.S)
    calls = []
    for code in blocks:
        for m in re.finditer(r"ccall\(\(\s*(?::(\w+)|\$\(QuoteNode\(fn\)\))\s*,\s*LIB\s*\)\s*,", code):
            name = m.group(1)
            rest = code[m.end():]
            # return type, then the parenthesised argument-type tuple
            ret, _, rest2 = rest.partition(",")
            rest2 = rest2.lstrip()
            assert rest2.startswith("("), rest2[:40]
            depth, i = 0, 0
            for i, ch in enumerate(rest2):
                depth += ch == "("
                depth -= ch == ")"
                if depth == 0:
                    break
            tup = rest2[1:i]
            args = [a for a in _split_top(tup) if a]
            calls.append((name, ret.strip(), args, code[max(0, m.start() - 300):m.start()]))
    return calls


def _check_type(jt, ct, where):
    jt = jt.strip()
    if jt in SCALARS:
        assert SCALARS[jt] == ct, (where, jt, ct)
        return
    m = re.fullmatch(r"(Ptr|Ref)\{(.*)\}", jt)
    assert m, (where, jt)
    inner = m.group(2).strip()
    c_level = ct.count("*") + (1 if ct.rstrip("*") in HANDLES else 0)
    base = ct.rstrip("*")
    inner_m = re.fullmatch(r"Ptr\{(.*)\}", inner)
    if inner_m:                      # pointer to pointer
        assert c_level == 2, (where, jt, ct)
        return
    assert c_level == 1, (where, jt, ct)
    if inner == "Cvoid" or inner.startswith("$"):
        return
    assert inner in POINTEES, (where, jt)
    assert POINTEES[inner] == base, (where, jt, ct)


def test_glue_binds_declared_entry_points_with_matching_types():
    decls = header_decls()
    calls = glue_ccalls()
    assert len(calls) >= 25
    seen = set()
    for name, ret, args, ctx in calls:
        if name is None:             # the @eval potrf! loops: every name is checked here
            for fn in ("dm_dpotrf", "dm_spotrf", "dm_zpotrf", "dm_cpotrf"):
                assert fn in decls
                assert len(args) == len(decls[fn][1])
            continue
        assert name in decls, name
        seen.add(name)
        cret, ctypes_ = decls[name]
        assert len(args) == len(ctypes_), (name, args, ctypes_)
        if ret == "Cstring":
            assert cret == "const char*", name
        else:
            assert ret == "Cint" and cret == "int", (name, ret)
        for jt, ct in zip(args, ctypes_):
            _check_type(jt, ct, name)
    # the generics the ROCExt checklist asks for are all bound
    for must in ("dm_memcpy_h2d", "dm_memcpy_d2h", "dm_memcpy_d2d", "dm_memcpy2d_h2d",
                 "dm_memcpy2d_d2h", "dm_malloc", "dm_free", "dm_fill_i32", "dm_fill_randn_d",
                 "dm_dgemm", "dm_dsyrk", "dm_dtrsm", "dm_dtrsv", "dm_dgemv", "dm_sgemm",
                 "dm_zgemm", "dm_cgemm", "dm_zherk", "dm_cherk", "dm_ztrsm", "dm_ctrsm",
                 "dm_stream_sync", "dm_device_sync", "dm_init", "dm_stream_create",
                 "dm_cholesky_tiled"):
        assert must in seen, must


def test_glue_defines_every_plugin_generic():
    """The ROCExt checklist (ext/ROCExt.jl:169-447, src/gpu.jl:88-262):
    movement, allocation, scope, context/sync and GPURef methods exist, and
    the modules the glue uses are imported."""
    code = "\n".join(re.findall(r"```julia\n(.*?)```", open(DOC).read(), flags=re.S))
    for pat in (r"using Dagger, LinearAlgebra, Distributed",
                r"using LinearAlgebra: BlasInt",
                r"\nmove!\(::Dagger\.CPURAMMemorySpace, ::MI355XSpace",
                r"\nmove!\(::MI355XSpace, ::Dagger\.CPURAMMemorySpace",
                r"\nmove!\(::MI355XSpace, ::MI355XSpace",
                r"function move\(from::Dagger\.CPUProc, to::MI355XProc, x::Array",
                r"function move\(from::MI355XProc, to::Dagger\.CPUProc, x::MIArray",
                r"function move\(from::MI355XProc, to::MI355XProc, x::MIArray",
                r"move\(::Dagger\.CPUProc, ::MI355XProc, x::Function\)",
                r"GPURef\(move\(from, to, x\[\]\)",
                r"function move!\(dep_mod",
                r"allocate_array_func\(::MI355XProc, ::typeof\(zeros\)\)",
                r"allocate_array_func\(::MI355XProc, ::typeof\(randn\)\)",
                r"allocate_array_func\(::MI355XProc, ::typeof\(rand\)\)",
                r"allocate_array_func\(::MI355XProc, ::typeof\(ones\)\)",
                r"allocate_array_func\(::MI355XProc, ::Dagger\.AllocateUndef",
                r"Dagger\.to_scope\(::Val\{:mi355x_gpu\}",
                r"with_context!\(p::MI355XProc\)",
                r"gpu_synchronize\(p::MI355XProc\)",
                r"function execute!\(p::MI355XProc",
                r"function unsafe_free!\(A::MIArray\)",
                r"move\(::Dagger\.CPUProc, ::MI355XProc, ::typeof\(LAPACK\.potrf!\)\)"):
        assert re.search(pat, code), pat
    # potrf! reuses the stream's info word: no allocation per call
    potrf = code[code.index("function mi_potrf!"):code.index("move(::Dagger.CPUProc")]
    assert "dm_malloc" not in potrf and "dm_free" not in potrf


@pytest.mark.parametrize("jt,ct,ok", [("Cint", "int", True), ("Cint", "double", False),
                                      ("Ptr{Float64}", "double*", True),
                                      ("Ptr{Float64}", "float*", False),
                                      ("Ptr{Ptr{Cvoid}}", "void**", True),
                                      ("Ptr{Ptr{Cvoid}}", "dm_stream_t*", True),
                                      ("Ptr{Cvoid}", "dm_stream_t", True),
                                      ("Ptr{Cvoid}", "void**", False),
                                      ("Ref{DMTiled}", "dm_tiled*", True)])
def test_type_checker_itself(jt, ct, ok):
    if ok:
        _check_type(jt, ct, "t")
    else:
        with pytest.raises(AssertionError):
            _check_type(jt, ct, "t")
