"""The Julia binding (tensornetworks.jl_amd/julia/TensorNetworksAMD) against the C-ABI it
calls, checked without a Julia toolchain: every ccall names a symbol that include/tnhip.h
declares and libtnhip.so exports, with the declared number of arguments and matching C
types (return type included), and the Julia mirrors of the parameter structs have the C
structs' fields in order. A wrong argument count or type in a ccall is undefined behaviour
at run time, so this is where such a mistake is caught."""
import ctypes
import re

import pytest

from conftest import PKG, ROOT

JL = PKG / "julia" / "TensorNetworksAMD"
SRC = [JL / "src" / "TensorNetworksAMD.jl", JL / "ext" / "TensorNetworksAMDChainRulesCoreExt.jl"]


def _strip_c(text):
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.sub(r"//[^\n]*", "", text)


def c_type_class(t):
    """Normalise a C parameter type to the Julia type class a ccall must use."""
    t = " ".join(t.replace("*", " * ").split())
    stars = t.count("*")
    base = [w for w in t.replace("*", " ").split() if w != "const"]
    base = base[:-1] if len(base) > 1 else base  # drop the parameter name
    b = base[0] if base else ""
    if stars == 0:
        return {"int": "Cint", "double": "Cdouble", "int64_t": "Int64"}[b]
    if stars == 2:
        return "Ptr{Ptr{Cvoid}}"
    if b in ("tnh_ctx", "tnh_buf", "tnh_dmrg", "void"):
        return "Ptr{Cvoid}"
    if b == "char":
        return "Cstring" if "const" in t else "Ptr{UInt8}"
    return {"int": "Ptr{Cint}", "double": "Ptr{Cdouble}", "int64_t": "Ptr{Int64}",
            "tnh_dmrg_params": "Ptr{DmrgParams}", "tnh_qjmc_problem": "Ptr{QjmcProblem}"}[b]


def header_decls():
    text = _strip_c((ROOT / "include" / "tnhip.h").read_text())
    decls = {}
    for m in re.finditer(r"([A-Za-z_][\w \*]*?)\b(tnh_[a-z0-9_]+)\s*\(([^;{]*?)\)\s*;", text, flags=re.S):
        ret, name, params = m.group(1).strip(), m.group(2), m.group(3).strip()
        ps = [] if params in ("", "void") else [p.strip() for p in params.split(",")]
        rc = {"int": "Cint", "const char *": "Cstring", "void *": "Ptr{Cvoid}"}[" ".join(ret.replace("*", " * ").split())]
        decls[name] = (rc, [c_type_class(p) for p in ps])
    return decls


JL_ALIASES = {"Float64": "Cdouble", "Ptr{Float64}": "Ptr{Cdouble}", "Int32": "Cint", "Ptr{Int32}": "Ptr{Cint}"}


def _split_top(s):
    """Split a comma-separated list at depth 0 of (), {}, []."""
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "({[":
            depth += 1
        elif ch in ")}]":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


def _balanced(text, start):
    """Index just past the parenthesis group opening at text[start] == '('."""
    depth = 0
    for i in range(start, len(text)):
        if text[i] == "(":
            depth += 1
        elif text[i] == ")":
            depth -= 1
            if depth == 0:
                return i + 1
    raise ValueError("unbalanced parentheses")


def julia_ccalls():
    calls = []
    for path in SRC:
        text = re.sub(r"#[^\n]*", "", path.read_text())
        for m in re.finditer(r"ccall\(\s*\(\s*:(tnh_[a-z0-9_]+)\s*,\s*libtnhip\s*\)\s*,", text):
            i = m.end()
            # return type, then the argument-type tuple, then the arguments
            j = text.index(",", i)
            ret = text[i:j].strip()
            k = text.index("(", j)
            end = _balanced(text, k)
            types = _split_top(text[k + 1:end - 1])
            close = _balanced(text, m.start() + len("ccall"))
            args = _split_top(text[end:close - 1].lstrip(", \n"))
            calls.append((path.name, m.group(1), ret, [JL_ALIASES.get(t, t) for t in types], args))
    return calls


def test_binding_files_exist():
    for p in SRC + [JL / "Project.toml"]:
        assert p.exists(), p
    proj = (JL / "Project.toml").read_text()
    assert "ChainRulesCore" in proj and "[extensions]" in proj


def test_every_ccall_matches_the_header():
    decls = header_decls()
    calls = julia_ccalls()
    assert len(calls) >= 25
    for fname, name, ret, types, args in calls:
        assert name in decls, f"{fname}: {name} is not declared in include/tnhip.h"
        cret, ctypes_ = decls[name]
        assert JL_ALIASES.get(ret, ret) == cret, (name, ret, cret)
        assert len(types) == len(ctypes_), (name, types, ctypes_)
        assert types == ctypes_, (name, types, ctypes_)
        assert len(args) == len(types), (name, "argument values vs types", args, types)


def test_ccall_symbols_are_exported():
    lib = ctypes.CDLL(str(PKG / "tnhip" / "libtnhip.so"))
    missing = sorted({n for _, n, _, _, _ in julia_ccalls() if not hasattr(lib, n)})
    assert not missing, missing


def test_binding_covers_the_reference_entry_points():
    """The drop-in surface: dmrg, tebd, applygates!, applygate!, replacesites!, svd and
    movecenter! get GPU methods; the rrule is code, not a comment."""
    text = (JL / "src" / "TensorNetworksAMD.jl").read_text()
    for fn in ["dmrg", "tebd", "applygates!", "applygate!", "replacesites!", "svd", "movecenter!"]:
        assert re.search(r"^function " + re.escape(fn) + r"\(g::GPU,", text, flags=re.M), fn
    for fn in ["product", "calculate", "buildleft!", "buildright!"]:
        assert re.search(r"^(function )?" + re.escape(fn) + r"\(p::GPUProjMPS", text, flags=re.M), fn
    ext = SRC[1].read_text()
    assert re.search(r"^function ChainRulesCore\.rrule\(::typeof\(coupling_energy\)", ext, flags=re.M)
    # the DMRG state is destroyed on every path out of the sweep loop
    body = text[text.index("function dmrg(g::GPU"):]
    body = body[:body.index("\nend\n")]
    assert "finally" in body and body.index("tnh_dmrg_destroy") > body.index("finally")


def _c_struct_fields(name):
    text = _strip_c((ROOT / "include" / "tnhip.h").read_text())
    m = re.search(r"typedef struct " + name + r"\s*\{(.*?)\}\s*" + name + ";", text, flags=re.S)
    fields = []
    for decl in m.group(1).split(";"):
        decl = " ".join(decl.split())
        if not decl:
            continue
        head, *rest = decl.split(",")
        t = head.rsplit(" ", 1)[0].replace("*", "").strip()
        star = "*" in head
        names = [head.rsplit(" ", 1)[1]] + [r.strip() for r in rest]
        for n in names:
            fields.append((n.replace("*", "").strip(), c_type_class(t + (" *" if star or "*" in n else "") + " x")))
    return fields


def _julia_struct_fields(name):
    text = re.sub(r"#[^\n]*", "", (JL / "src" / "TensorNetworksAMD.jl").read_text())
    m = re.search(r"struct " + name + r"\n(.*?)\nend", text, flags=re.S)
    fields = []
    for part in re.split(r"[;\n]", m.group(1)):
        part = part.strip()
        if part:
            n, t = part.split("::")
            fields.append((n.strip(), JL_ALIASES.get(t.strip(), t.strip())))
    return fields


@pytest.mark.parametrize("c_name,jl_name", [("tnh_dmrg_params", "DmrgParams"),
                                            ("tnh_qjmc_problem", "QjmcProblem")])
def test_parameter_structs_match(c_name, jl_name):
    c = _c_struct_fields(c_name)
    j = _julia_struct_fields(jl_name)
    assert [n for n, _ in c] == [n for n, _ in j]
    assert [t for _, t in c] == [t for _, t in j], (c, j)
