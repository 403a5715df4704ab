"""The Rust -sys crate under rust/ is not compiled here (no cargo in this
image): its declarations are checked against include/spec_spars_hip.h as
text -- every function it declares exists in the header with the same
parameter count, and its repr(C) structs list the header's fields in the
same order."""
import os
import re

from conftest import ROOT


def _c_functions():
    txt = open(os.path.join(ROOT, "include", "spec_spars_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    out = {}
    for m in re.finditer(r"\b(?:int|void|const char\s*\*)\s*(ss_[a-z0-9_]+)\s*\(([^)]*)\)\s*;", txt):
        args = [a for a in m.group(2).split(",") if a.strip() and a.strip() != "void"]
        out[m.group(1)] = len(args)
    return out, txt


def test_sys_crate_functions_match_the_header():
    rs = open(os.path.join(ROOT, "rust", "spec_spars_hip_sys", "src", "lib.rs")).read()
    funcs, _ = _c_functions()
    decl = re.findall(r"pub fn (ss_[a-z0-9_]+)\(([^)]*)\)", rs, flags=re.S)
    assert len(decl) >= 20
    for name, args in decl:
        assert name in funcs, name
        n = len([a for a in args.split(",") if a.strip()])
        assert n == funcs[name], (name, n, funcs[name])


def test_sys_crate_structs_match_the_header():
    rs = open(os.path.join(ROOT, "rust", "spec_spars_hip_sys", "src", "lib.rs")).read()
    _, txt = _c_functions()
    for st in ("ss_solve_opts", "ss_solve_report", "ss_sparsify_opts", "ss_sparsify_report"):
        c_body = re.search(r"typedef struct %s \{(.*?)\}\s*%s;" % (st, st), txt, flags=re.S).group(1)
        c_fields = re.findall(r"\b([a-z_0-9]+)\s*;", c_body)
        r_body = re.search(r"pub struct %s \{(.*?)\n\}" % st, rs, flags=re.S).group(1)
        r_fields = re.findall(r"pub ([a-z_0-9]+):", r_body)
        assert c_fields == r_fields, (st, c_fields, r_fields)


def test_shim_calls_only_declared_functions():
    shim = open(os.path.join(ROOT, "rust", "reference_crate_src", "hip.rs")).read()
    rs = open(os.path.join(ROOT, "rust", "spec_spars_hip_sys", "src", "lib.rs")).read()
    for name in re.findall(r"sys::(ss_[a-z0-9_]+)\(", shim):
        assert "pub fn %s(" % name in rs, name
    assert "pub fn run_solve_lap(shared_jl_cols: FlattenedVec" in shim
