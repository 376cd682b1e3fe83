"""CPU test: the Rust declarations a maintainer would paste into the sys
crate agree with the C header and with the
Python ctypes mirror: every #[repr(C)] struct has the header's fields in the
header's order with matching types, every header function is declared with
the same argument count and types, and the ctypes _fields_ match too.  A
drifted struct would make one side read the other's padding."""
import ctypes as C
import os
import re

import pytest

from conftest import ROOT, load_package

RUST_C = {"i32": "int32", "i64": "int64", "f64": "double", "usize": "size", "c_int": "int32",
          "c_double": "double", "c_char": "char"}
C_TYPES = {"int32_t": "int32", "int": "int32", "int64_t": "int64", "double": "double", "size_t": "size",
           "char": "char", "void": "void"}


def header():
    return open(os.path.join(ROOT, "include", "hfg.h")).read()


def rust_block():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    m = re.search(r"```rust\n// hf-rs-gpu-sys/src/lib.rs\n(.*?)```", text, re.S)
    assert m, "sys block missing"
    return m.group(1)


def norm_c(t):
    """'const double*' -> ('double', 1); 'hfg_ctx**' -> ('hfg_ctx', 2)."""
    t = t.replace("const", " ").strip()
    stars = t.count("*")
    base = t.replace("*", " ").split()[0]
    return (C_TYPES.get(base, base), stars)


def norm_rust(t):
    t = t.strip()
    stars = 0
    while True:
        m = re.match(r"\*(const|mut)\s+(.*)", t)
        if not m:
            break
        stars += 1
        t = m.group(2).strip()
    return (RUST_C.get(t, t), stars)


def c_structs():
    out = {}
    for body, name in re.findall(r"typedef struct \{(.*?)\}\s*(\w+);", header(), re.S):
        body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
        fields = []
        for decl in body.split(";"):
            decl = decl.strip()
            if not decl:
                continue
            m = re.match(r"((?:const\s+)?\w+\s*\**)\s*(.*)", decl)
            base = m.group(1)
            for nm in m.group(2).split(","):
                nm = nm.strip()
                stars = base.count("*") + nm.count("*")
                fields.append((nm.replace("*", "").strip(), (norm_c(base.replace("*", ""))[0], stars)))
        out[name] = fields
    return out


def rust_structs():
    out = {}
    for name, body in re.findall(r"pub struct (\w+)\s*\{(.*?)\}", rust_block(), re.S):
        out[name] = [(f, norm_rust(t)) for f, t in re.findall(r"pub (\w+):\s*([^,]+?),?\s*(?:\n|$)", body)]
    return out


def c_functions():
    text = re.sub(r"/\*.*?\*/", "", header(), flags=re.S)
    out = {}
    for ret, name, args in re.findall(r"([\w\s\*]+?)\b(hfg_\w+)\s*\(([^)]*)\);", text):
        args = args.strip()
        if args in ("", "void"):
            out[name] = []
            continue
        sig = []
        for a in args.split(","):
            a = a.strip()
            m = re.match(r"(.*?)(\w+)$", a)
            sig.append(norm_c(m.group(1)))
        out[name] = sig
    return out


def rust_functions():
    out = {}
    for name, args in re.findall(r"pub fn (hfg_\w+)\((.*?)\)", rust_block(), re.S):
        args = args.strip()
        out[name] = [] if not args else [norm_rust(a.split(":", 1)[1]) for a in args.split(",") if a.strip()]
    return out


def test_every_c_struct_matches_its_rust_declaration():
    cs, rs = c_structs(), rust_structs()
    assert {"hfg_basis_view", "hfg_scf_options", "hfg_scf_result", "hfg_dfhf_options",
            "hfg_dfhf_result"} <= set(cs)
    for name, fields in cs.items():
        assert name in rs, "struct %s is not declared in the sys block" % name
        assert [f for f, _ in rs[name]] == [f for f, _ in fields], name
        for (f, ct), (_, rt) in zip(fields, rs[name]):
            assert ct == rt, "%s.%s: C %s vs Rust %s" % (name, f, ct, rt)


def test_every_c_function_is_declared_in_the_sys_block():
    cf, rf = c_functions(), rust_functions()
    assert len(cf) >= 35
    assert set(cf) == set(rf), (sorted(set(cf) - set(rf)), sorted(set(rf) - set(cf)))
    for name, sig in cf.items():
        assert len(sig) == len(rf[name]), name
        for k, (a, b) in enumerate(zip(sig, rf[name])):
            # opaque handle types keep their name; everything else maps 1:1
            assert a == b, "%s argument %d: C %s vs Rust %s" % (name, k, a, b)


CT = {C.c_int32: ("int32", 0), C.c_int64: ("int64", 0), C.c_double: ("double", 0)}


def test_ctypes_mirrors_match_the_header():
    hf = load_package()
    cs = c_structs()
    pairs = {"hfg_basis_view": hf.BasisView, "hfg_scf_options": hf.ScfOptions, "hfg_scf_result": hf.ScfResult,
             "hfg_dfhf_options": hf.DfhfOptions, "hfg_dfhf_result": hf.DfhfResult,
             "hfg_server_options": hf.ServerOptions, "hfg_step_result": hf.StepResult,
             "hfg_request_result": hf.RequestResult, "hfg_server_stats_t": hf.ServerStats,
             "hfg_cd_info": hf.CdInfo}
    assert set(pairs) == set(cs), sorted(set(cs) ^ set(pairs))
    for cname, py in pairs.items():
        fields = cs[cname]
        assert len(py._fields_) == len(fields), cname
        for (f, t), (pf, pt) in zip(fields, py._fields_):
            if t[1]:
                assert pt is C.c_void_p, (cname, f)
            else:
                assert CT[pt] == t, (cname, f, pt, t)


def test_parser_catches_a_dropped_field():
    """The check above is live: a Rust block missing incremental_fock fails it."""
    rs = rust_structs()
    fields = [f for f, _ in rs["hfg_scf_options"]][:-1]
    assert fields != [f for f, _ in c_structs()["hfg_scf_options"]]
