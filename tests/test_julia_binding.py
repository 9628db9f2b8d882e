"""The Julia binding's ccalls against the C ABI (no Julia needed).

Every `ccall((:fmm4rb_x, libfmm4rb), Ret, (Types...), args...)` in
fmm4rbgpu.jl_amd/julia/FMM4RBHIP.jl, and in the Julia code blocks of the integration
guide at the repository root, must name a function that include/fmm4rb.h declares, with the Julia types of the
header's return and parameter types, and pass as many arguments as it declares.
Also checks that the binding registers its update_particles_field! method by the
reference package's UUID (any load order), not by a name lookup in Main."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JL = os.path.join(ROOT, "fmm4rbgpu.jl_amd", "julia", "FMM4RBHIP.jl")
HDR = os.path.join(ROOT, "include", "fmm4rb.h")
INTEG = os.path.join(ROOT, "INTEGRATION.md")

# C parameter/return types of the ABI -> the Julia types a ccall must declare
C2JL = {
    "int": {"Cint"},
    "int64_t": {"Int64"},
    "double": {"Float64"},
    "const char *": {"Cstring", "Ptr{UInt8}"},
    "double *": {"Ptr{Float64}"},
    "const double *": {"Ptr{Float64}"},
    "int64_t *": {"Ptr{Int64}"},
    "const int64_t *": {"Ptr{Int64}"},
    "fmm4rb_ctx *": {"Ptr{Cvoid}"},
    "fmm4rb_ctx **": {"Ptr{Ptr{Cvoid}}"},
    "fmm4rb_stats *": {"Ptr{Cvoid}"},  # the binding passes C_NULL
    "void *": {"Ptr{Cvoid}"},
}


def _norm_ctype(t):
    t = re.sub(r"\s+", " ", t.strip())
    t = re.sub(r"\s*\*", " *", t).replace("* *", "**")
    return t


def header_protos():
    txt = open(HDR).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    protos = {}
    for m in re.finditer(r"([A-Za-z_][\w \t\*]*?)\b(fmm4rb_\w+)\s*\(([^)]*)\)\s*;", txt):
        ret, name, params = _norm_ctype(m.group(1)), m.group(2), m.group(3).strip()
        types = []
        if params and params != "void":
            for p in params.split(","):
                p = p.strip()
                pm = re.match(r"(.*?)(\w+)$", p)
                types.append(_norm_ctype(pm.group(1)))
        protos[name] = (ret, types)
    return protos


def _split_top(s):
    """Split at commas outside (), {}, []."""
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


def _balanced(s, i):
    """s[i] == '(' -> the text inside the matching parentheses and the index after."""
    assert s[i] == "("
    depth = 0
    for j in range(i, len(s)):
        if s[j] == "(":
            depth += 1
        elif s[j] == ")":
            depth -= 1
            if depth == 0:
                return s[i + 1:j], j + 1
    raise ValueError("unbalanced parentheses")


def _strip_jl_comments(src):
    return "\n".join(re.sub(r"#(?![^\"]*\"[^\"]*$).*$", "", line) for line in src.splitlines())


def ccalls(src):
    """[(symbol, return type, [arg types], number of arguments or None)]."""
    src = _strip_jl_comments(src)
    out = []
    for m in re.finditer(r"\bccall\s*\(", src):
        inner, _ = _balanced(src, m.end() - 1)
        parts = _split_top(inner)
        sym = re.match(r"\(\s*:(\w+)\s*,", parts[0]).group(1)
        ret = parts[1]
        tup = parts[2].strip()
        assert tup.startswith("(") and tup.endswith(")"), tup
        types = [t for t in _split_top(tup[1:-1]) if t]
        args = parts[3:]
        nargs = len(args)
        if len(args) == 1 and args[0].endswith("..."):
            # a splatted tuple: the nearest preceding `name = (...)` assignment
            name = args[0][:-3].strip()
            defs = [d for d in re.finditer(r"\b%s\s*=\s*\(" % re.escape(name), src)
                    if d.start() < m.start()]
            assert defs, f"no definition of {name} before ccall of {sym}"
            tup_txt, _ = _balanced(src, defs[-1].end() - 1)
            nargs = len(_split_top(tup_txt))
        out.append((sym, ret, types, nargs))
    return out


def _julia_blocks(md):
    return "\n".join(re.findall(r"```julia\n(.*?)```", md, flags=re.S))


def test_header_parses():
    p = header_protos()
    assert p["fmm4rb_field"][0] == "int" and len(p["fmm4rb_field"][1]) == 13
    assert p["fmm4rb_last_error"] == ("const char *", [])
    assert p["fmm4rb_create"][1] == ["fmm4rb_ctx **", "int"]


@pytest.mark.parametrize("where", ["binding", "integration"])
def test_ccalls_match_header(where):
    src = open(JL).read() if where == "binding" else _julia_blocks(open(INTEG).read())
    protos = header_protos()
    calls = ccalls(src)
    assert calls, "no ccall found"
    for sym, ret, types, nargs in calls:
        assert sym in protos, f"{sym} is not declared in include/fmm4rb.h"
        cret, ctypes_ = protos[sym]
        assert ret in C2JL[cret], f"{sym}: return {ret} vs C {cret}"
        assert len(types) == len(ctypes_), f"{sym}: {len(types)} types vs {len(ctypes_)} in C"
        for k, (jt, ct) in enumerate(zip(types, ctypes_)):
            assert jt in C2JL[ct], f"{sym} argument {k}: {jt} vs C {ct}"
        assert nargs == len(ctypes_), f"{sym}: {nargs} arguments passed, {len(ctypes_)} declared"


def test_binding_covers_the_drop_in_entry_points():
    syms = {c[0] for c in ccalls(open(JL).read())}
    for s in ("fmm4rb_create", "fmm4rb_destroy", "fmm4rb_last_error", "fmm4rb_field",
              "fmm4rb_field_refit", "fmm4rb_field_vjp", "fmm4rb_field_pullback"):
        assert s in syms, s


def test_binding_registration_is_load_order_independent():
    src = _strip_jl_comments(open(JL).read())
    # the reference package and ChainRulesCore by UUID (FMM4RBGPU's Project.toml)
    assert '"6188f75b-cc19-42af-84de-782f8f05992b"' in src
    assert '"d360d2e6-b24c-11e9-a2a3-2a2ae2dbcce4"' in src
    assert "Base.package_callbacks" in src and "Base.loaded_modules" in src
    assert "isdefined(Main" not in src
    # without FMM4RBGPU the module's own update_particles_field! raises
    assert re.search(r"function update_particles_field!\(particles, alg::FMMHIP; lambda\)", src)
    assert "error(" in src[src.index("function update_particles_field!(particles, alg::FMMHIP"):]
