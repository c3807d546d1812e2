"""The Rust side of the boundary (rust/rz-sys, rust/rz-render) matches
include/rz.h: the same entry points with the same parameter counts, the same
constants, and the rz_batch_item field order. No Rust toolchain exists here,
so this is a source-level check of the crates."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _read(*p):
    with open(os.path.join(ROOT, *p)) as f:
        return f.read()


def _c_functions():
    text = re.sub(r"/\*.*?\*/", "", _read("include", "rz.h"), flags=re.S)
    out = {}
    for m in re.finditer(r"\b(rz_[a-z_0-9]+)\s*\(([^;]*?)\)\s*;", text, flags=re.S):
        args = [a.strip() for a in m.group(2).split(",") if a.strip() and a.strip() != "void"]
        out[m.group(1)] = len(args)
    return out


def _rust_functions():
    text = _read("rust", "rz-sys", "src", "lib.rs")
    out = {}
    for m in re.finditer(r"pub fn (rz_[a-z_0-9]+)\s*\((.*?)\)\s*(->[^;]*)?;", text, flags=re.S):
        args = [a for a in m.group(2).split(",") if a.strip()]
        out[m.group(1)] = len(args)
    return out


def test_rz_sys_declares_every_entry_point():
    c, r = _c_functions(), _rust_functions()
    assert sorted(c) == sorted(r)
    assert c == r  # same parameter counts


def test_rz_sys_constants_match_header():
    h = re.sub(r"/\*.*?\*/", "", _read("include", "rz.h"), flags=re.S)
    rs = _read("rust", "rz-sys", "src", "lib.rs")
    consts = dict(re.findall(r"\b(RZ_[A-Z_0-9]+)\s*=\s*(-?\d+)", h))
    consts.update({k: v.rstrip("u") for k, v in re.findall(r"#define\s+(RZ_[A-Z_0-9]+)\s+(\d+)u?", h)})
    rust = dict(re.findall(r"pub const (RZ_[A-Z_0-9]+): \w+ = (-?\d+);", rs))
    assert consts, "no constants parsed"
    for k, v in consts.items():
        assert rust.get(k) == v, (k, v, rust.get(k))


def test_batch_item_fields_in_c_order():
    h = _read("include", "rz.h")
    body = re.search(r"typedef struct rz_batch_item \{(.*?)\} rz_batch_item;", h, flags=re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    c_fields = re.findall(r"\*?\s*(\w+)(?:\[\d+\])?\s*[,;]", body)
    rs = _read("rust", "rz-sys", "src", "lib.rs")
    rbody = re.search(r"pub struct rz_batch_item \{(.*?)\}", rs, flags=re.S).group(1)
    r_fields = re.findall(r"pub (\w+):", rbody)
    assert c_fields == r_fields, (c_fields, r_fields)


def test_rz_render_calls_only_declared_functions():
    used = set(re.findall(r"sys::(rz_[a-z_0-9]+)\(", _read("rust", "rz-render", "src", "lib.rs")))
    assert used and used <= set(_c_functions())
    # the reference's draw API is all there
    text = _read("rust", "rz-render", "src", "lib.rs")
    for item in ("pub fn new(width: usize, height: usize) -> Renderer", "pub fn uniforms(&mut self) -> &mut Uniforms",
                 "pub fn render(&mut self, mesh: &Mesh, vertex_shader: VertexShader, fragment_shader: FragmentShader)",
                 "pub fn framebuffer(&mut self) -> &[u32]", "pub fn bind_texture(&mut self, index: usize, tex: Texture)",
                 "pub fn write_block(&mut self) -> &mut UniformBlock"):
        assert item in text, item
