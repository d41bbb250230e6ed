"""The Rust bindings (geon_amd/rust/geon-hip-sys) agree with the C header.

There is no Rust toolchain in this image, so the crate is checked as text:
every entry point of include/geon_hip.h is declared in its extern block with
the same parameter count, and every structure has the same fields in the
same order with the Rust types of the C types. The shim crate
(geon_amd/rust/geon-hip) implements the reference's Metric trait with every
method of src/metric/mod.rs:19-109.
"""
import os
import re

from tests.conftest import ROOT

HEADER = os.path.join(ROOT, "include", "geon_hip.h")
SYS = os.path.join(ROOT, "geon_amd", "rust", "geon-hip-sys", "src", "lib.rs")
SHIM = os.path.join(ROOT, "geon_amd", "rust", "geon-hip", "src", "lib.rs")

C2RUST = {"int32_t": "i32", "uint32_t": "u32", "uint64_t": "u64", "double": "f64",
          "float": "f32", "int": "c_int", "size_t": "usize"}


def _strip_c(text):
    return re.sub(r"/\*.*?\*/", "", text, flags=re.S)


def _c_functions():
    text = _strip_c(open(HEADER, encoding="utf-8").read())
    out = {}
    for m in re.finditer(r"\b(geon_[a-z0-9_]+)\s*\(([^)]*)\)\s*;", text):
        params = [p for p in m.group(2).split(",") if p.strip() and p.strip() != "void"]
        out[m.group(1)] = len(params)
    return out


def _rust_functions():
    text = open(SYS, encoding="utf-8").read()
    block = text[text.index('extern "C" {'):]
    out = {}
    for m in re.finditer(r"pub fn (geon_[a-z0-9_]+)\s*\(([^)]*)\)", block, flags=re.S):
        params = [p for p in m.group(2).split(",") if p.strip()]
        out[m.group(1)] = len(params)
    return out


def _c_structs():
    text = _strip_c(open(HEADER, encoding="utf-8").read())
    out = {}
    for m in re.finditer(r"typedef struct (\w+) \{(.*?)\} (\w+);", text, flags=re.S):
        fields = []
        for decl in m.group(2).split(";"):
            decl = decl.strip()
            if not decl:
                continue
            f = re.match(r"(\w+)\s+(\w+)(\[\d+\])?$", decl)
            assert f, decl
            ctype, name, dim = f.groups()
            fields.append((name, C2RUST.get(ctype, ctype), dim))
        out[m.group(3)] = fields
    return out


def _rust_structs():
    text = open(SYS, encoding="utf-8").read()
    out = {}
    for m in re.finditer(r"pub struct (geon_\w+_t) \{(.*?)\n\}", text, flags=re.S):
        fields = []
        for line in m.group(2).split("\n"):
            f = re.match(r"\s*pub (\w+): (\[(\w+); (\d+)\]|\w+),", line)
            if f:
                name, ty, elem, n = f.groups()
                fields.append((name, elem or ty, f"[{n}]" if n else None))
        if fields:
            out[m.group(1)] = fields
    return out


def test_every_entry_point_is_bound():
    c, r = _c_functions(), _rust_functions()
    assert len(c) >= 45
    assert set(c) == set(r), sorted(set(c) ^ set(r))
    for name, n in c.items():
        assert r[name] == n, (name, n, r[name])


def test_structures_have_the_c_layout():
    c, r = _c_structs(), _rust_structs()
    for name in ("geon_metric_t", "geon_camera_t", "geon_render_cfg_t", "geon_render_stats_t",
                 "geon_sched_cfg_t", "geon_serve_cfg_t"):
        assert c[name] == r[name], (name, c[name], r[name])
    # the serving result embeds the statistics
    assert [f[0] for f in c["geon_serve_result_t"]] == [f[0] for f in r["geon_serve_result_t"]]


def test_constants_match():
    text = _strip_c(open(HEADER, encoding="utf-8").read())
    rust = open(SYS, encoding="utf-8").read()
    consts = dict(re.findall(r"#define (GEON_[A-Z0-9_]+)\s+(-?\d+)\b", text))
    assert len(consts) >= 25
    for k, v in consts.items():
        assert re.search(rf"pub const {k}: i32 = {v};", rust), k


def test_shim_implements_the_metric_trait():
    text = open(SHIM, encoding="utf-8").read()
    impl = text[text.index("impl Metric for HipMetric"):]
    impl = impl[:impl.index("\n}\n")]
    # the trait's methods (src/metric/mod.rs:21-108)
    for method in ("g", "dg", "g_inv", "christoffel", "fd_step", "terminate", "name",
                   "asymptotic_direction"):
        assert re.search(rf"fn {method}\(&self", impl), method
    # the GPU functions keep the reference's names and argument order
    assert re.search(r"pub fn render\(metric: &HipMetric, camera: &Camera, cfg: &RenderConfig\) "
                     r"-> \(Vec<f32>, RenderStats\)", text)
    assert re.search(r"pub fn trace\(metric: &HipMetric, y0: State, step: Scalar, max_steps: usize\)"
                     r" -> TraceResult", text)
    for f in ("riemann", "ricci", "ricci_scalar", "einstein", "stress_energy", "nec_contraction",
              "wec_contraction"):
        assert re.search(rf"pub fn {f}\(metric: &HipMetric, x: &Vec4", text), f
