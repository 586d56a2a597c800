"""The Julia binding (oetqf.jl_amd/julia/OetqfAMD.jl) cannot run here (no
Julia), so its ccall sites are checked against include/oetqf_amd.h as text:
every ccall names a declared entry point, its return type matches, its
argument-type tuple has the header's parameter kinds in order (pointer,
double, int64_t, int), and it passes as many values as it declares types."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JL = os.path.join(ROOT, "oetqf.jl_amd", "julia", "OetqfAMD.jl")
HDR = os.path.join(ROOT, "include", "oetqf_amd.h")


def _split_top(s):
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "([{":
            depth += 1
        elif ch in ")]}":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


def _ccalls(src):
    i = 0
    while True:
        i = src.find("ccall(", i)
        if i < 0:
            return
        j = k = i + len("ccall(")
        depth = 1
        while depth:
            depth += {"(": 1, ")": -1}.get(src[k], 0)
            k += 1
        yield src[j:k - 1]
        i = k


def _c_kind(p):
    p = p.strip()
    if "*" in p or "[" in p:
        return "ptr"
    for kind in ("double", "int64_t", "int"):
        if re.match(r"(const\s+)?%s\b" % kind, p):
            return kind
    return "unknown: " + p


def _jl_kind(t):
    t = t.strip()
    if t.startswith(("Ptr{", "Ref{")) or t == "Cstring":
        return "ptr"
    return {"Cdouble": "double", "Float64": "double", "Int64": "int64_t", "Cint": "int",
            "Int32": "int"}.get(t, "unknown: " + t)


def _header():
    src = open(HDR).read()
    decls = re.findall(r"^(int|void|const char \*)\s*(oq_[a-z0-9_]+)\s*\(([^;]*?)\)\s*;", src, re.M | re.S)
    out = {}
    for ret, name, params in decls:
        params = " ".join(params.split())
        out[name] = (ret, [] if params in ("", "void") else [_c_kind(p) for p in params.split(",")])
    return out


def _problems(jl_src, hdr):
    """(problems, names seen, ccall count) of the binding source against the header."""
    rets = {"Cint": "int", "Cstring": "const char *", "Cvoid": "void"}
    bad, seen, n = [], set(), 0
    for body in _ccalls(jl_src):
        args = _split_top(body)
        names = re.findall(r":(oq_[a-z0-9_]+)", args[0])
        tup = args[2]
        if not names or not (tup.startswith("(") and tup.endswith(")")):
            bad.append(("unparsed", body[:60]))
            continue
        kinds = [_jl_kind(t) for t in _split_top(tup[1:-1])]
        if len(args) - 3 != len(kinds):
            bad.append((names, "passes %d values for %d types" % (len(args) - 3, len(kinds))))
        for name in names:
            if name not in hdr:
                bad.append((name, "not declared"))
                continue
            ret, want = hdr[name]
            if rets.get(args[1]) != ret:
                bad.append((name, "returns %s, header %s" % (args[1], ret)))
            if kinds != want:
                bad.append((name, kinds, want))
            seen.add(name)
        n += 1
    return bad, seen, n


def test_julia_ccalls_match_the_header():
    bad, seen, n = _problems(open(JL).read(), _header())
    assert not bad, bad
    assert n >= 30
    # the binding covers the reference's entry points: Green's functions, RHS, steppers, ensembles
    for name in ("oq_gf_okada_ff", "oq_gf_okada_mf", "oq_gf_hex8_fm", "oq_gf_hex8_mm", "oq_rhs_create",
                 "oq_rhs_create_from_mesh", "oq_rhs_eval", "oq_stepper_create", "oq_stepper_create_vcabm5",
                 "oq_stepper_run", "oq_ensemble_create", "oq_ensemble_run"):
        assert name in seen, name


def test_planted_binding_mistakes_are_caught():
    """A Float64 where the header has int64_t, and a dropped value, in the
    binding's oq_stepper_run call, must both be reported."""
    src, hdr = open(JL).read(), _header()
    good = "(Ptr{Cvoid}, Int64, Ref{Cint}), s[], nstep, rc)"
    assert good in src
    bad, _, _ = _problems(src.replace(good, "(Ptr{Cvoid}, Float64, Ref{Cint}), s[], nstep, rc)"), hdr)
    assert any(b[0] == "oq_stepper_run" for b in bad)
    bad, _, _ = _problems(src.replace(good, "(Ptr{Cvoid}, Int64, Ref{Cint}), s[], nstep)"), hdr)
    assert any("values" in str(b) for b in bad)
