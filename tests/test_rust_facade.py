"""The Rust facade crate (rust/) against the C header it binds: no Rust toolchain is installed
here, so the crate is checked as text -- every extern "C" declaration in rust/src/ffi.rs names a
function of include/voronoids.h with the same parameter and return types under the C -> Rust
type map, vo_stats has the header's fields in the header's order, and the crate's sources have
balanced delimiters.  The reference's public API names (src/delaunay_tree.rs, src/geometry.rs,
src/scheduler.rs) are present with the reference's signatures."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "voronoids.h")
CRATE = os.path.join(ROOT, "rust")

C_TO_RUST = {
    "int": "c_int", "int64_t": "i64", "uint64_t": "u64", "double": "f64", "void": "()",
    "const double*": "*const f64", "double*": "*mut f64",
    "const int32_t*": "*const i32", "int32_t*": "*mut i32", "int32_t**": "*mut *mut i32",
    "int64_t*": "*mut i64", "const int64_t*": "*const i64",
    "vo_tri*": "*mut vo_tri", "const vo_tri*": "*const vo_tri", "vo_tri**": "*mut *mut vo_tri",
    "const char*": "*const c_char", "void*": "*mut c_void", "vo_stats*": "*mut vo_stats",
}


def _norm_c(t: str) -> str:
    t = re.sub(r"\s+", " ", t.strip())
    t = re.sub(r"\s*\*", "*", t)
    return t


def header_functions() -> dict:
    text = re.sub(r"/\*.*?\*/", "", open(HDR).read(), flags=re.S)
    out = {}
    for m in re.finditer(r"^\s*((?:const\s+)?[a-z_0-9]+\s*\**)\s*(vo_[a-z_0-9]+)\s*\(([^)]*)\)\s*;", text, re.M):
        ret, name, params = m.groups()
        ps = [] if params.strip() in ("", "void") else [p.strip() for p in params.split(",")]
        types = []
        for p in ps:
            p = re.sub(r"\s+", " ", p)
            mm = re.match(r"(.*?)([A-Za-z_][A-Za-z_0-9]*)$", p)
            types.append(_norm_c(mm.group(1)))
        out[name] = (_norm_c(ret), types)
    return out


def rust_externs() -> dict:
    text = open(os.path.join(CRATE, "src", "ffi.rs")).read()
    block = text[text.index('extern "C" {'):]
    out = {}
    for m in re.finditer(r"pub fn (vo_[a-z_0-9]+)\s*\((.*?)\)\s*(?:->\s*([^;]+?))?\s*;", block, re.S):
        name, params, ret = m.groups()
        ps = [p.strip() for p in params.split(",") if p.strip()]
        types = [re.sub(r"\s+", " ", p.split(":", 1)[1].strip()) for p in ps]
        out[name] = ((ret or "()").strip(), types)
    return out


def test_every_extern_matches_the_header():
    h, r = header_functions(), rust_externs()
    assert len(r) >= 24
    for name, (ret, types) in r.items():
        assert name in h, f"{name} is not declared in include/voronoids.h"
        hret, htypes = h[name]
        assert C_TO_RUST[hret] == ret, (name, hret, ret)
        assert [C_TO_RUST[t] for t in htypes] == types, (name, htypes, types)


def test_vo_stats_layout_matches_the_header():
    text = re.sub(r"/\*.*?\*/", "", open(HDR).read(), flags=re.S)
    body = re.search(r"typedef struct \{(.*?)\} vo_stats;", text, re.S).group(1)
    want = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        ctype, names = decl.split(None, 1)
        for n in names.split(","):
            n = n.strip()
            m = re.match(r"(\w+)\[(\d+)\]", n)
            rust_t = {"int64_t": "i64", "double": "f64"}[ctype]
            want.append((m.group(1), f"[{rust_t}; {m.group(2)}]") if m else (n, rust_t))
    rs = open(os.path.join(CRATE, "src", "ffi.rs")).read()
    sbody = re.search(r"pub struct vo_stats \{(.*?)\n\}", rs, re.S).group(1)
    got = re.findall(r"pub (\w+): ([^,]+),", sbody)
    assert got == want


def test_sources_have_balanced_delimiters():
    for dirpath, _, files in os.walk(CRATE):
        for f in files:
            if not f.endswith(".rs"):
                continue
            text = open(os.path.join(dirpath, f)).read()
            # strip comments, strings and char literals before counting
            text = re.sub(r"//[^\n]*", "", text)
            text = re.sub(r'"(?:\\.|[^"\\])*"', '""', text)
            text = re.sub(r"'(?:\\.|[^'\\])'", "''", text)
            stack = []
            pairs = {")": "(", "]": "[", "}": "{"}
            for i, ch in enumerate(text):
                if ch in "([{":
                    stack.append(ch)
                elif ch in ")]}":
                    assert stack and stack[-1] == pairs[ch], (f, i)
                    stack.pop()
            assert not stack, f


def test_reference_api_names_and_signatures_are_present():
    src = {m: open(os.path.join(CRATE, "src", m + ".rs")).read() for m in ("delaunay_tree", "geometry", "scheduler", "lib")}
    want = {
        "delaunay_tree": [
            r"pub struct Simplex<const N: usize, const M: usize>", r"pub struct Vertex<const N: usize>",
            r"pub struct DelaunayTree<const N: usize, const M: usize>", r"pub max_simplex_id: usize",
            r"pub fn new\(vertices: Vec<\[f64; 3\]>\) -> Self", r"pub fn new\(vertices: Vec<\[f64; 2\]>\) -> Self",
            r"pub fn locate\(&self, vertex: \[f64; N\]\) -> Vec<usize>",
            r"pub fn get_new_simplices\(\s*&self,\s*killed_site_id: usize,\s*vertex: \[f64; N\],\s*vertex_id: usize,\s*\)",
            r"pub fn insert_point\(&mut self, update: &TreeUpdate<N, M>\)",
            r"pub fn insert_points_parallel\(&mut self, updates: &Vec<TreeUpdate<N, M>>\)",
            r"pub fn add_points_to_tree\(&mut self, vertices: Vec<\[f64; N\]>\)",
            r"pub fn check_delaunay\(&self\) -> bool",
            r"pub struct TreeUpdate<const N: usize, const M: usize>",
            r"pub fn new\(id: usize, vertex: \[f64; N\], tree: &DelaunayTree<N, M>\) -> Self",
        ],
        "geometry": [
            r"pub fn circumsphere<const N: usize, const M: usize>\(vertices: \[\[f64; N\]; M\]\) -> \(\[f64; N\], f64\)",
            r"pub fn in_sphere<const N: usize>\(vertex: \[f64; N\], center: \[f64; N\], radius: f64\) -> bool",
            r"pub fn bounding_sphere<const N: usize>\(vertices: Vec<\[f64; N\]>\) -> \(\[f64; N\], f64\)",
        ],
        "scheduler": [
            r"pub fn make_queue<const N: usize, const M: usize>\(\s*vertices: Vec<\[f64; N\]>,\s*tree: &DelaunayTree<N, M>,\s*\) -> Vec<\(usize, \[f64; N\], Vec<usize>\)>",
            r"pub fn find_placement<const N: usize>\(queue: &Vec<\(usize, \[f64; N\], Vec<usize>\)>\) -> Vec<usize>",
        ],
        "lib": [r"pub mod delaunay_tree;", r"pub mod geometry;", r"pub mod scheduler;"],
    }
    for mod, pats in want.items():
        for p in pats:
            assert re.search(p, src[mod]), (mod, p)
    assert re.search(r'\[lib\]\s*\n(?:#.*\n)*name = "voronoids"', open(os.path.join(CRATE, "Cargo.toml")).read())
