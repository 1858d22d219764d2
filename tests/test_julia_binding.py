"""Mechanical checks of the Julia binding (julia/HIPExtDummy.jlh, julia/ext/*.jl).
No Julia toolchain exists here, so the binding is checked as text:

* every `ccall((:ph_*, libpartia_hip), R, (T...), args...)` names a function
  that include/partia_hip.h declares, with as many argument types and values
  as the C declaration has parameters, each Julia type compatible with the C
  type (and the return type with the C return type);
* the Julia mirrors of ph_catalog and ph_request have the C structs' fields,
  in order, with compatible types;
* every `name.field` access on a parameter annotated with one of the
  reference's types names a field of that struct
  (tests/golden/reference_structs.json, read from the reference's source);
* the names imported from Partia exist in the reference, and each driver
  method the binding adds has the positional parameter types of a reference
  method of the same function, apart from the backend."""
import json
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JULIA = [os.path.join(ROOT, "julia", "HIPExtDummy.jlh"), os.path.join(ROOT, "julia", "ext", "HIPExt.jl"),
         os.path.join(ROOT, "julia", "ext", "HIPChainRulesExt.jl")]
HEADER = os.path.join(ROOT, "include", "partia_hip.h")
FIXTURE = os.path.join(ROOT, "tests", "golden", "reference_structs.json")
OPAQUE = {"ph_context", "ph_lbvh", "ph_batch", "void"}


def strip_c_comments(s):
    return re.sub(r"/\*.*?\*/", "", s, flags=re.S)


def strip_julia_comments(s):
    s = re.sub(r'"""(.*?)"""', lambda m: '"' * 3 + " " * 0 + '"' * 3, s, flags=re.S)  # docstrings
    return re.sub(r"#[^\n]*", "", s)


def split_top(s):
    out, depth, cur, quote = [], 0, "", False
    for ch in s:
        if ch == '"':
            quote = not quote
        if not quote:
            if ch in "({[":
                depth += 1
            elif ch in ")}]":
                depth -= 1
        if ch == "," and depth == 0 and not quote:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


def matching(s, i):
    """Index of the bracket closing the one at s[i]."""
    depth = 0
    for j in range(i, len(s)):
        if s[j] in "({[":
            depth += 1
        elif s[j] in ")}]":
            depth -= 1
            if depth == 0:
                return j
    raise ValueError("unbalanced brackets")


def c_param_type(p):
    p = " ".join(p.split())
    m = re.match(r"(.*?)\s*([A-Za-z_][A-Za-z0-9_]*)$", p)
    t = m.group(1) if m and m.group(1).strip() else p
    return re.sub(r"\s*\*", "*", t).strip()


def header_functions():
    src = strip_c_comments(open(HEADER).read())
    out = {}
    for m in re.finditer(r"\n\s*((?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**)\s*(ph_[a-z0-9_]+)\s*\(([^;]*?)\)\s*;", src):
        ret = re.sub(r"\s*\*", "*", " ".join(m.group(1).split()))
        params = [c_param_type(p) for p in split_top(m.group(3))]
        out[m.group(2)] = (ret, params)
    return out


def header_struct(name):
    src = strip_c_comments(open(HEADER).read())
    m = re.search(r"typedef\s+struct\s+%s\s*\{(.*?)\}\s*%s;" % (name, name), src, flags=re.S)
    fields = []
    for decl in m.group(1).split(";"):
        decl = " ".join(decl.split())
        if not decl:
            continue
        arr = re.match(r"((?:const\s+)?[A-Za-z_]\w*)\s*(.+)$", decl)
        base, names = arr.group(1), arr.group(2)
        for nm in names.split(","):
            nm = nm.strip()
            stars = len(nm) - len(nm.lstrip("*"))
            nm = nm.lstrip("*").strip()
            dims = re.findall(r"\[(\d+)\]", nm)
            nm = re.sub(r"\[.*", "", nm)
            fields.append((nm, base + "*" * stars, int(dims[0]) if dims else 0))
    return fields


def c_compatible(jt, ct, ret=False):
    """Is Julia ccall type jt a correct binding of C type ct?"""
    jt = jt.replace(" ", "")
    ct = re.sub(r"\bconst\b", "", ct).replace(" ", "")
    scalar = {"Cint": {"int", "int32_t"}, "Int32": {"int32_t", "int"}, "Int64": {"int64_t"},
              "Float64": {"double"}, "UInt8": {"uint8_t"}, "Cstring": {"char*"}}
    if jt in scalar:
        return ct in scalar[jt]
    if ret:
        return False
    m = re.fullmatch(r"(Ptr|Ref)\{(.*)\}", jt)
    if not m:
        return False
    kind, inner = m.group(1), m.group(2)
    if not ct.endswith("*"):
        return False
    pointee = ct[:-1]
    if kind == "Ref" and inner == "Ptr{Cvoid}":  # an output handle: T**
        return pointee.endswith("*") and pointee[:-1] in OPAQUE
    if kind == "Ref":
        return {"PhCatalog": "ph_catalog", "PhRequest": "ph_request"}.get(inner) == pointee
    if inner == "Cvoid":  # any data or handle pointer
        return not pointee.endswith("*")
    if pointee == "void":  # typed Julia pointer to a C void*
        return not inner.startswith("Ptr{")
    if inner == "Ptr{Float64}":
        return pointee == "double*"
    ptr_scalar = {"Float64": "double", "Float32": "float", "Int32": "int32_t", "Int64": "int64_t",
                  "UInt64": "uint64_t", "UInt8": "uint8_t", "PhRequest": "ph_request"}
    return ptr_scalar.get(inner) == pointee


def julia_sources():
    return {p: strip_julia_comments(open(p).read()) for p in JULIA}


def ccalls():
    out = []
    for path, src in julia_sources().items():
        for m in re.finditer(r"\bccall\(", src):
            j = matching(src, m.end() - 1)
            args = split_top(src[m.end():j])
            fn = re.match(r"\(\s*:(ph_[a-z0-9_]+)\s*,\s*libpartia_hip\s*\)", args[0])
            assert fn, "%s: ccall target %r is not (:ph_*, libpartia_hip)" % (path, args[0])
            types = args[2].strip()
            assert types.startswith("(") and types.endswith(")"), (path, types)
            tl = split_top(types[1:-1])
            out.append((os.path.basename(path), fn.group(1), args[1].strip(), tl, args[3:]))
    return out


def test_every_ccall_matches_the_header():
    decl = header_functions()
    calls = ccalls()
    assert len(calls) >= 12
    for path, fn, ret, types, values in calls:
        assert fn in decl, "%s: %s is not declared in include/partia_hip.h" % (path, fn)
        cret, cparams = decl[fn]
        assert c_compatible(ret, cret, ret=True), (path, fn, ret, cret)
        assert len(types) == len(cparams), "%s: %s has %d argument types, the header %d" % (
            path, fn, len(types), len(cparams))
        assert len(values) == len(cparams), "%s: %s passes %d values, the header has %d parameters" % (
            path, fn, len(values), len(cparams))
        for k, (jt, ct) in enumerate(zip(types, cparams)):
            assert c_compatible(jt, ct), "%s: %s argument %d: Julia %s does not bind C %s" % (path, fn, k + 1, jt, ct)


def test_the_binding_covers_the_drivers_entry_points():
    used = {fn for _, fn, _, _, _ in ccalls()}
    for fn in ("ph_ctx_create", "ph_last_error", "ph_lbvh_build", "ph_lbvh_get", "ph_lbvh_destroy",
               "ph_point_samples", "ph_line_samples", "ph_structured_grid_coords", "ph_line_samples_adjoint_device",
               "ph_batch_create", "ph_batch_run", "ph_batch_destroy"):
        assert fn in used, fn


def julia_struct(src, name):
    m = re.search(r"\bstruct\s+%s\b(.*?)\n\s*(end|function)\b" % name, src, flags=re.S)
    assert m, name
    return [(a, t.strip()) for a, t in re.findall(r"\n\s*([A-Za-z_]\w*)\s*::\s*([^\n]+)", m.group(1))]


def test_catalog_mirror_matches_the_c_struct():
    src = julia_sources()[JULIA[1]]
    jf = julia_struct(src, "PhCatalog")
    cf = header_struct("ph_catalog")
    assert [a for a, _ in jf] == [nm for nm, _, _ in cf]
    for (_, jt), (nm, ct, dim) in zip(jf, cf):
        assert dim == 0
        assert c_compatible(jt, ct) or (jt in ("Int32",) and ct == "int32_t"), (nm, jt, ct)


def test_request_mirror_matches_the_c_struct():
    jf = julia_struct(julia_sources()[JULIA[1]], "PhRequest")
    cf = header_struct("ph_request")
    assert [a for a, _ in jf] == [nm for nm, _, _ in cf]
    for (_, jt), (nm, ct, dim) in zip(jf, cf):
        if dim:
            tup = re.fullmatch(r"NTuple\{(\d+),\s*(\w+)\}", jt)
            assert tup and int(tup.group(1)) == dim and c_compatible(tup.group(2), ct), (nm, jt, ct)
        else:
            assert c_compatible(jt, ct), (nm, jt, ct)


def annotated_functions(src):
    """(name, [(param, type head)], body) of every top-level function."""
    out = []
    for m in re.finditer(r"(?m)^function\s+([\w.]+)\s*\(", src):
        close = matching(src, m.end() - 1)
        params = []
        for a in split_top(src[m.end():close].split(";")[0]):
            a = a.split("=")[0]
            pm = re.match(r"\s*([A-Za-z_]\w*)?\s*::\s*([A-Za-z_][\w.]*)", a)
            if pm:
                params.append((pm.group(1), pm.group(2).split(".")[-1]))
        end = re.search(r"(?m)^end\b", src[close:])
        out.append((m.group(1), params, src[close:close + end.start()]))
    return out


def test_field_accesses_name_reference_fields():
    ref = json.load(open(FIXTURE))
    structs = {k: set(v["fields"]) for k, v in ref["structs"].items()}
    srcs = julia_sources()
    # the binding's own structs
    structs["HIPComputeBackend"] = {a for a, _ in julia_struct(srcs[JULIA[0]], "HIPComputeBackend")}
    for nm in ("DeviceLBVH", "CatalogArrays", "PhCatalog", "DeviceBatch", "PhRequest"):
        structs[nm] = {a for a, _ in julia_struct(srcs[JULIA[1]], nm)}
    checked = 0
    for path, src in srcs.items():
        for fname, params, body in annotated_functions(src):
            for pname, ptype in params:
                if not pname or ptype not in structs:
                    continue
                for field in re.findall(r"\b%s\.([A-Za-zρ_][\wρ]*)" % re.escape(pname), body):
                    assert field in structs[ptype], "%s: %s(%s :: %s) reads .%s, not a field of %s (%s)" % (
                        os.path.basename(path), fname, pname, ptype, field, ptype,
                        ref["structs"].get(ptype, {}).get("source", "the binding"))
                    checked += 1
    assert checked >= 20


def test_imported_names_and_driver_signatures_match_the_reference():
    ref = json.load(open(FIXTURE))
    srcs = julia_sources()
    ext = srcs[JULIA[1]]
    imp = re.search(r"using Partia:\s*(.*?)\n\n", ext, flags=re.S).group(1)
    own = {"HIPComputeBackend"}
    for nm in (x.strip() for x in imp.split(",")):
        assert nm in own or nm in ref["names"] or nm in ref["structs"], nm
    extended = 0
    for fname, params, _ in annotated_functions(ext):
        if not fname.startswith("Partia.") or fname.split(".")[1] not in ref["functions"]:
            continue
        mine = [t for _, t in params]
        assert mine[0] == "HIPComputeBackend", fname
        methods = [m["param_types"] for m in ref["functions"][fname.split(".")[1]]]
        generic = {"AS", "B", "Any"}

        def same(a, b):
            return len(a) == len(b) and all(x == y or y in generic or x in generic for x, y in zip(a[1:], b[1:]))
        assert any(same(mine, r) for r in methods), "%s%s matches no reference method %s" % (fname, mine, methods)
        extended += 1
    assert extended >= 6


def test_status_codes_map_to_the_reference_exceptions():
    ext = julia_sources()[JULIA[1]]
    hdr = strip_c_comments(open(HEADER).read())
    codes = dict(re.findall(r"#define\s+(PH_ERR_[A-Z_]+)\s+\((-\d+)\)", hdr))
    for nm, val in re.findall(r"const\s+(PH_ERR_[A-Z_]+)\s*=\s*Cint\((-\d+)\)", ext):
        assert codes[nm] == val, nm
    body = re.search(r"function check\(.*?\n end|function check\(.*?\nend", ext, flags=re.S).group(0)
    assert "DimensionMismatch" in body and "ArgumentError" in body and "AssertionError" in body


def block_balance(src):
    """(openers, ends) of Julia block keywords outside brackets, strings and comments."""
    src = re.sub(r'"(\\.|[^"\\])*"', '""', src)
    openers = {"module", "baremodule", "function", "macro", "struct", "begin", "for", "while", "if", "let", "do",
               "try", "quote"}
    depth, opens, ends = 0, 0, 0
    for tok in re.findall(r"[A-Za-z_][\w!]*|[()\[\]{}]", src):
        if tok in "([{":
            depth += 1
        elif tok in ")]}":
            depth -= 1
        elif depth == 0 and tok in openers:
            opens += 1
        elif depth == 0 and tok == "end":
            ends += 1
    return opens, ends


@pytest.mark.parametrize("path", JULIA)
def test_julia_sources_have_balanced_blocks(path):
    """Every block opener has its `end` outside brackets (a parse check the absent
    toolchain cannot do); a planted missing `end` fails it."""
    src = strip_julia_comments(open(path).read())
    opens, ends = block_balance(src)
    assert opens == ends and ends > 0, (opens, ends)
    cut = src[:src.rindex("end")] + src[src.rindex("end") + 3:]
    assert block_balance(cut)[0] != block_balance(cut)[1]


@pytest.mark.parametrize("mistake", ["coor", "arity", "type", "catalog_int64"])
def test_planted_mistakes_are_caught(tmp_path, mistake):
    """The checks fail on a copy of the binding with a planted mistake: the
    PointSamples field `coor` read from an InterpolationInput (whose field is
    `coord`), a ccall missing an argument, a 32-bit sample count for an int64_t
    parameter, and Int64 catalog slots where ph_catalog holds int32_t."""
    src = open(JULIA[1]).read()
    if mistake == "coor":
        bad = src.replace("input.coord[1]", "input.coor[1]", 1)
        check = test_field_accesses_name_reference_fields
    elif mistake == "arity":
        bad = src.replace("(Ptr{Cvoid}, Int32, Ptr{UInt64}), t.handle, PH_GET_ORDER,\n                order)",
                          "(Ptr{Cvoid}, Int32), t.handle, PH_GET_ORDER)", 1)
        check = test_every_ccall_matches_the_header
    elif mistake == "type":
        bad = src.replace("(Ptr{Cvoid}, Ptr{Cvoid}, Ref{PhCatalog}, Int32, Int32, Int64, Ptr{Float64}, Ptr{Float64},\n"
                          "                     Ptr{Float64}, Ptr{Float64}),",
                          "(Ptr{Cvoid}, Ptr{Cvoid}, Ref{PhCatalog}, Int32, Int32, Int32, Ptr{Float64}, Ptr{Float64},\n"
                          "                     Ptr{Float64}, Ptr{Float64}),", 1)
        check = test_every_ccall_matches_the_header
    else:
        bad = src.replace("    scalar_slots :: Ptr{Int32}\n", "    scalar_slots :: Ptr{Int64}\n", 1)
        check = test_catalog_mirror_matches_the_c_struct
    assert bad != src
    p = tmp_path / "HIPExt.jl"
    p.write_text(bad)
    saved = JULIA[1]
    JULIA[1] = str(p)
    try:
        with pytest.raises(AssertionError):
            check()
    finally:
        JULIA[1] = saved
