"""The Julia binding (flou.jl_amd/julia/FlouHIP.jl) against the C ABI it calls.

There is no Julia in this image, so the binding is checked as text, against the two other
statements of the same interface:
  * its `struct Desc` has the fields of `flou_hip_desc` (include/flouhip.h) and of the
    Python mirror (flouhip._capi.Desc), in the same order with the same C types, so the
    isbits struct Julia passes by reference has the C layout;
  * every `ccall` names a function the header declares, with as many arguments as the
    prototype and a Julia type that passes each C parameter type;
  * its status codes and enum values equal the header's.
"""
import ctypes
import os
import re

import pytest

from flouhip import _capi as C

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JL = os.path.join(ROOT, "flou.jl_amd", "julia", "FlouHIP.jl")
HEADER = os.path.join(ROOT, "include", "flouhip.h")


def _strip_c_comments(text):
    return re.sub(r"/\*.*?\*/", "", text, flags=re.S)


def _strip_jl_comments(text):
    return "\n".join(line.split("#", 1)[0] if not line.lstrip().startswith("#") else "" for line in text.splitlines())


JULIA_TYPES = {"Int32": "i32", "Int64": "i64", "Float64": "f64", "NTuple{3,Float64}": "f64x3",
               "Ptr{Float64}": "p_f64", "Ptr{Int64}": "p_i64", "Ptr{Int8}": "p_i8", "Ptr{UInt8}": "p_u8",
               "Ptr{Int32}": "p_i32"}
C_TYPES = {"int32_t": "i32", "int64_t": "i64", "double": "f64"}


def julia_desc():
    text = _strip_jl_comments(open(JL).read())
    body = re.search(r"^struct Desc\n(.*?)^end", text, flags=re.S | re.M).group(1)
    fields = []
    for line in body.splitlines():
        line = line.strip()
        if not line:
            continue
        name, typ = line.split("::")
        fields.append((name.strip(), JULIA_TYPES[typ.strip()]))
    return fields


def header_desc():
    text = _strip_c_comments(open(HEADER).read())
    body = re.search(r"typedef struct flou_hip_desc \{(.*?)\} flou_hip_desc;", text, flags=re.S).group(1)
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        m = re.fullmatch(r"(const\s+)?(\w+)\s*(\*)?\s*(\w+)(\[(\d+)\])?", decl)
        assert m, decl
        base = C_TYPES.get(m.group(2)) or {"int8_t": "i8", "uint8_t": "u8"}[m.group(2)]
        if m.group(3):
            typ = "p_" + base
        elif m.group(6):
            typ = f"{base}x{m.group(6)}"
        else:
            typ = base
        fields.append((m.group(4), typ))
    return fields


def ctypes_desc():
    out = []
    for name, t in C.Desc._fields_:
        if t is ctypes.c_int32:
            typ = "i32"
        elif t is ctypes.c_int64:
            typ = "i64"
        elif t is ctypes.c_double:
            typ = "f64"
        elif isinstance(t, type) and issubclass(t, ctypes.Array):
            assert t._type_ is ctypes.c_double
            typ = f"f64x{t._length_}"
        else:
            typ = "p_" + {ctypes.c_double: "f64", ctypes.c_int64: "i64", ctypes.c_int8: "i8",
                          ctypes.c_uint8: "u8", ctypes.c_int32: "i32"}[t._type_]
        out.append((name, typ))
    return out


def test_julia_desc_matches_header_and_python_mirror():
    jl, h, py = julia_desc(), header_desc(), ctypes_desc()
    assert len(h) == 36
    assert jl == h, [(a, b) for a, b in zip(jl, h) if a != b] or (len(jl), len(h))
    assert py == h
    # Julia lays out an isbits struct as C does (each field at the next multiple of its
    # alignment, the size a multiple of the largest): the offsets of the Julia field list
    # are those of the ctypes structure, which is the C layout
    size_align = {"i32": (4, 4), "i64": (8, 8), "f64": (8, 8), "f64x3": (24, 8)}
    off, offsets, amax = 0, [], 1
    for _, typ in jl:
        sz, al = size_align.get(typ, (8, 8))  # pointers: 8 bytes
        off = (off + al - 1) // al * al
        offsets.append(off)
        off += sz
        amax = max(amax, al)
    size = (off + amax - 1) // amax * amax
    assert offsets == [getattr(C.Desc, name).offset for name, _ in C.Desc._fields_]
    assert size == ctypes.sizeof(C.Desc)


def header_prototypes():
    text = _strip_c_comments(open(HEADER).read())
    protos = {}
    for m in re.finditer(r"\b[\w\s\*]+?\b(flou_hip_\w+)\s*\(([^)]*)\)\s*;", text):
        params = [p.strip() for p in m.group(2).split(",") if p.strip() and p.strip() != "void"]
        protos[m.group(1)] = params
    return protos


def _c_param_kind(p):
    p = re.sub(r"\bconst\b", "", p).strip()
    if "*" in p:
        return "ptr"
    base = p.split()[0]
    return {"double": "f64", "int32_t": "i32", "int64_t": "i64", "size_t": "size"}[base]


def _jl_type_kind(t):
    t = t.strip()
    if t.startswith("Ptr{") or t.startswith("Ref{") or t in ("Cstring",):
        return "ptr"
    return {"Float64": "f64", "Int32": "i32", "Int64": "i64", "Csize_t": "size", "Cint": "i32"}[t]


def _split_types(tup):
    parts, depth, cur = [], 0, ""
    for ch in tup:
        if ch == "{":
            depth += 1
        elif ch == "}":
            depth -= 1
        if ch == "," and depth == 0:
            parts.append(cur)
            cur = ""
        else:
            cur += ch
    if cur.strip():
        parts.append(cur)
    return [p.strip() for p in parts if p.strip()]


def julia_ccalls():
    text = _strip_jl_comments(open(JL).read())
    calls = re.findall(r"ccall\(\(:(flou_hip_\w+),\s*libflouhip\),\s*(\w+),\s*\(([^()]*)\)", text, flags=re.S)
    return [(name, ret, _split_types(tup)) for name, ret, tup in calls]


def test_julia_ccalls_match_the_prototypes():
    protos = header_prototypes()
    calls = julia_ccalls()
    assert len(calls) >= 20
    used = set()
    for name, ret, types in calls:
        assert name in protos, name
        params = protos[name]
        assert len(types) == len(params), (name, types, params)
        for t, p in zip(types, params):
            assert _jl_type_kind(t) == _c_param_kind(p), (name, t, p)
        used.add(name)
    # the binding covers the rhs, the stages, the reductions, the limiter, the monitors,
    # the output projection and both adjoint entry points
    for name in ("flou_hip_create", "flou_hip_rhs", "flou_hip_timeintegrate", "flou_hip_max_dt",
                 "flou_hip_integrate", "flou_hip_set_limiter", "flou_hip_monitor", "flou_hip_equispaced",
                 "flou_hip_set_bc", "flou_hip_rhs_vjp", "flou_hip_rk_step_record", "flou_hip_rk_step_vjp"):
        assert name in used, name


def _header_enum_values():
    text = _strip_c_comments(open(HEADER).read())
    return {k: int(v) for k, v in re.findall(r"\bFLOU_HIP_(\w+)\s*=\s*(-?\d+)", text)}


def test_julia_constants_match_the_header():
    hv = _header_enum_values()
    text = _strip_jl_comments(open(JL).read())
    seen = 0
    for names, values in re.findall(r"^const ([A-Z0-9_, ]+) = ([^\n]+)$", text, flags=re.M):
        names = [n.strip() for n in names.split(",")]
        vals = [int(re.sub(r"Int32\((-?\d+)\)", r"\1", v.strip())) for v in values.split(",")]
        if len(names) != len(vals):
            continue
        for n, v in zip(names, vals):
            key = {"SCHEME_SSPRK33": "SSPRK33", "SCHEME_CK2N54": "CK2N54", "SCHEME_ORK256": "ORK256"}.get(n, n)
            assert key in hv, n
            assert hv[key] == v, (n, v, hv[key])
            seen += 1
    assert seen >= 30


@pytest.mark.parametrize("op,code", [("StrongDivOperator", "STRONG"), ("SplitDivOperator", "SPLIT"),
                                     ("HybridDivOperator", "HYBRID")])
def test_julia_maps_every_operator(op, code):
    text = open(JL).read()
    assert re.search(rf"operator_codes\(op::{op}\)", text) and code in text
