"""The Julia binding in INTEGRATION.md against include/cbm_hip.h (CPU only;
there is no Julia here): the C structs mirrored field by field (name, order,
C type), every ModelDesc/SimDesc/LossDesc construction, every ccall (declared
name and argument count) and the reference's dispatch seams (the platform is
community.abm.platform; Community is not parametric)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CTYPE = {"int32_t": "Int32", "int64_t": "Int64", "uint64_t": "UInt64", "double": "Float64"}


def header():
    with open(os.path.join(ROOT, "include", "cbm_hip.h")) as f:
        return re.sub(r"/\*.*?\*/", "", f.read(), flags=re.S)


def julia():
    with open(os.path.join(ROOT, "INTEGRATION.md")) as f:
        text = f.read()
    return "\n".join(re.findall(r"```julia\n(.*?)```", text, flags=re.S))


def c_struct(name):
    body = re.search(r"typedef struct %s \{(.*?)\}" % name, header(), flags=re.S).group(1)
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        m = re.match(r"(?:const\s+)?(\w+)\s*(\*?)\s*(\w+)(?:\[(\d+)\])?$", decl)
        assert m, decl
        ctype, ptr, fname, n = m.groups()
        jt = CTYPE[ctype]
        if ptr:
            jt = f"Ptr{{{jt}}}"
        elif n:
            jt = f"NTuple{{{n},{jt}}}"
        fields.append((fname, jt))
    return fields


def jl_struct(name):
    m = re.search(r"(?:Base\.@kwdef\s+)?struct %s\n(.*?)\nend" % name, julia(), flags=re.S)
    assert m, name
    fields = []
    for line in m.group(1).splitlines():
        for part in line.split("#")[0].split(";"):
            part = part.strip()
            if part:
                fm = re.match(r"(\w+)::([\w{},]+)", part)
                assert fm, part
                fields.append(fm.groups())
    return fields


@pytest.mark.parametrize("c,jl", [("cbm_model_desc", "ModelDesc"), ("cbm_sim_desc", "SimDesc"),
                                  ("cbm_loss_desc", "LossDesc")])
def test_julia_structs_mirror_the_header(c, jl):
    assert jl_struct(jl) == c_struct(c)


def calls(src, fn):
    """Argument lists of every call fn(...) (brackets balanced)."""
    out = []
    for m in re.finditer(r"(?<![\w.])%s\(" % fn, src):
        depth, i = 1, m.end()
        while depth:
            depth += {"(": 1, ")": -1}.get(src[i], 0)
            i += 1
        out.append(src[m.end():i - 1])
    return out


def split_args(args):
    parts, depth, cur = [], 0, ""
    for ch in args:
        if ch in "([{":
            depth += 1
        elif ch in ")]}":
            depth -= 1
        if ch == "," and depth == 0:
            parts.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        parts.append(cur.strip())
    return parts


@pytest.mark.parametrize("jl,required", [("ModelDesc", ()), ("SimDesc", ("n", "n_max", "box_lo", "box_hi")),
                                         ("LossDesc", None)])
def test_every_construction_matches_the_struct(jl, required):
    names = [f for f, _ in jl_struct(jl)]
    src = julia()
    body = re.sub(r"struct %s\n.*?\nend" % jl, "", src, flags=re.S)
    found = calls(body, jl)
    assert found, jl
    for args in found:
        parts = split_args(args.replace(";", ","))
        kw = [p.split("=")[0].strip() for p in parts if re.match(r"\w+\s*=", p)]
        if kw:
            assert len(kw) == len(parts), args
            assert set(kw) <= set(names), set(kw) - set(names)
            assert set(required or ()) <= set(kw), args
        else:
            assert len(parts) == len(names), (jl, len(parts), len(names), args)


def c_arity():
    out = {}
    for m in re.finditer(r"\b(cbm_\w+)\s*\(([^)]*)\)\s*;", header()):
        params = m.group(2).strip()
        out[m.group(1)] = 0 if params in ("", "void") else len(params.split(","))
    return out


def test_every_ccall_names_a_declared_entry_point_with_its_arity():
    arity = c_arity()
    src = julia()
    seen = set()
    for args in calls(src, "ccall"):
        parts = split_args(args)
        name = re.match(r"\(:(\w+),", parts[0]).group(1)
        assert name in arity, name
        types = parts[2]
        assert types.startswith("(") and types.endswith(")"), types
        n = len(split_args(types[1:-1]))
        assert n == arity[name], (name, n, arity[name])
        assert len(parts) - 3 == n, (name, parts)
        seen.add(name)
    assert {"cbm_create", "cbm_step", "cbm_upload", "cbm_download", "cbm_step_save"} <= seen


def test_dispatch_uses_the_reference_seams():
    src = julia()
    assert "Community{" not in src
    assert "com.platform" not in src and "abm.neighborsAlg" not in src
    for seam in ("loadToPlatform!", "bringFromPlatform!", "step!(community)", "evolve!", "update!",
                 "computeNeighbors!"):
        assert seam in src, seam
    assert len(re.findall(r"typeof\((?:com|community)\.abm\.platform\) <: HIP", src)) == 6
