"""WGSL front end on the CPU: the reference's shader-compiler known-answer
tests (tests/wgsl_kats.py, from naga-cranelift/src/tests.rs) run through the
translator, with the emitted functions compiled by the host C++ compiler
against the same prelude the device build uses; the constructs the
reference leaves as todo!() are rejected with errors that say so; and the
pipelines of the shipped shaders compile for gfx950 with hipRTC (no device
needed).

The GPU runs of the same tests are in test_gpu_wgsl.py.
"""
from __future__ import annotations

import json
import shutil
import struct
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT / "wgpu-cpu_amd"))
sys.path.insert(0, str(ROOT / "tests"))

import wgc  # noqa: E402
from wgsl_kats import kats  # noqa: E402

CSRC = ROOT / "wgpu-cpu_amd" / "csrc"
SHADERS = ROOT / "wgpu-cpu_amd" / "shaders"


def _lib_or_skip():
    try:
        wgc.load_library()
    except Exception as e:  # pragma: no cover - only without a build
        pytest.skip(f"libwgc.so not built: {e}")


def words_to(out_type, words):
    if out_type == "i32":
        return struct.unpack("<i", struct.pack("<I", words[0]))[0]
    if out_type == "u32":
        return words[0]
    if out_type == "f32":
        return struct.unpack("<f", struct.pack("<I", words[0]))[0]
    if out_type == "vec4i":
        return list(struct.unpack("<4i", struct.pack("<4I", *words[:4])))
    if out_type == "vec4f":
        return list(struct.unpack("<4f", struct.pack("<4I", *words[:4])))
    raise ValueError(out_type)


def host_program(cases) -> str:
    """One C++ program running every case's entry point with HostEnv; prints
    one line per invocation: case, index, abort code, 4 location-0 words, 4
    position words (hex)."""
    parts = ['#include "wgsl_host_env.h"', "#include <cstdio>"]
    runs = []
    for i, k in enumerate(cases):
        info = json.loads(wgc.wgsl_translate(k.source, 1))
        (ep,) = info["entries"]
        parts.append(f"namespace k{i} {{\n{wgc.wgsl_translate(k.source, 0)}\n}}")
        bufs = ""
        for s, b in enumerate(k.buffers):
            arr = ", ".join(str(x) for x in b)
            bufs += (f"    static const uint8_t b{i}_{s}[] = {{{arr}}};\n"
                     f"    env.bufs[{s}] = b{i}_{s}; env.sizes[{s}] = {len(b)};\n")
        runs.append(f"""  for (uint32_t n = 0; n < {k.vertex_count}u; ++n) {{
    HostEnv env;
    env.vi = n;
{bufs}    uint32_t rc = k{i}::{ep['symbol']}(env);
    uint32_t p[4];
    memcpy(p, env.pos, sizeof p);
    printf("{i} %u %u %08x %08x %08x %08x %08x %08x %08x %08x\\n", n, rc, env.loc[0][0], env.loc[0][1],
           env.loc[0][2], env.loc[0][3], p[0], p[1], p[2], p[3]);
  }}""")
    parts.append("int main()\n{\n" + "\n".join(runs) + "\n  return 0;\n}\n")
    return "\n".join(parts)


@pytest.fixture(scope="module")
def host_results(tmp_path_factory):
    _lib_or_skip()
    cxx = shutil.which("g++") or shutil.which("c++")
    if not cxx:
        pytest.skip("no host C++ compiler")
    cases = kats()
    d = tmp_path_factory.mktemp("wgsl_host")
    src = d / "kats.cpp"
    src.write_text(host_program(cases))
    exe = d / "kats"
    subprocess.run([cxx, "-std=c++17", "-O1", "-ffp-contract=off", "-fno-fast-math", "-I", str(CSRC), "-I",
                    str(ROOT / "tests"), str(src), "-o", str(exe)], check=True, capture_output=True, text=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    res = {}
    for line in out.splitlines():
        f = line.split()
        i, n, rc = int(f[0]), int(f[1]), int(f[2])
        words = [int(x, 16) for x in f[3:7]]
        pos = list(struct.unpack("<4f", struct.pack("<4I", *[int(x, 16) for x in f[7:11]])))
        res[(i, n)] = (rc, words, pos)
    return cases, res


def check_case(k, per_invocation):
    """per_invocation[n] = (abort code, location-0 words, position)."""
    if k.expect_abort:
        assert per_invocation[0][0] == k.expect_abort, k.name
        return
    assert per_invocation[0][0] == 0, (k.name, per_invocation[0][0])
    if k.positions is not None:
        got = [per_invocation[n][2] for n in range(k.vertex_count)]
        assert got == [list(p) for p in k.positions], k.name
    if k.out_type != "none":
        got = words_to(k.out_type, per_invocation[0][1])
        assert got == k.expected, (k.name, got, k.expected)


def test_host_kats(host_results):
    cases, res = host_results
    assert len(cases) >= 60
    for i, k in enumerate(cases):
        check_case(k, [res[(i, n)] for n in range(k.vertex_count)])


def test_reflection_of_shipped_shaders():
    _lib_or_skip()
    info = json.loads(wgc.wgsl_translate((SHADERS / "mvp_texture.wgsl").read_text(), 1))
    kinds = sorted((r["group"], r["binding"], r["kind"]) for r in info["resources"])
    assert kinds == [(0, 0, "uniform"), (1, 0, "texture"), (1, 1, "sampler")]
    stages = sorted(e["stage"] for e in info["entries"])
    assert stages == ["fragment", "vertex"]
    vs = [e for e in info["entries"] if e["stage"] == "vertex"][0]
    assert [(v["location"], v["ncomp"]) for v in vs["inputs"]] == [(0, 4), (1, 2)]
    assert vs["position_out"]


@pytest.mark.parametrize("src,needle", [
    # math.rs:21-30, swizzle / splat are not the issue here
    ("fn f(x: f32) -> f32 { return sqrt(x); }", "math.rs"),
    ("fn f(x: vec4f) -> bool { return all(x == x); }", "vector comparisons"),
    ("fn f(a: mat4x4f, b: mat4x4f) -> mat4x4f { return a * b; }", "binary.rs"),
    ("fn f(a: vec4f, b: mat4x4f) -> vec4f { return a * b; }", "binary.rs"),
    ("fn f(x: f32) -> f32 { return dpdx(x); }", "derivative.rs"),
    ("@group(0) @binding(0) var t: texture_2d<f32>;\nfn f() -> vec4f { return textureLoad(t, vec2i(0, 0), 0); }",
     "image.rs"),
])
def test_reference_todo_constructs_are_rejected(src, needle):
    _lib_or_skip()
    with pytest.raises(wgc.WgcError) as e:
        wgc.wgsl_translate(src)
    assert needle in str(e.value)
    assert e.value.code == wgc.ERR_UNSUPPORTED


@pytest.mark.parametrize("src,needle", [
    ("fn f() -> i32 { return 1 / 0; }", "division by zero"),
    ("fn f() -> i32 { let x: i32 = 3000000000; return x; }", "does not fit in i32"),
    ("fn f() -> i32 { return y; }", "unknown identifier"),
    ("fn f() { var a = 1; var a = 2; }", "redeclaration"),
    ("fn f() -> u32 { return 1.5; }", "cannot convert"),
    ("fn f() { f(); }", "recursion"),
    ("fn f() -> i32 { let a = array<i32, 2>(1, 2); return a[2]; }", "out of bounds"),
])
def test_invalid_wgsl_is_rejected_with_a_position(src, needle):
    _lib_or_skip()
    with pytest.raises(wgc.WgcError) as e:
        wgc.wgsl_translate(src)
    assert needle in str(e.value)
    assert str(e.value).split(":")[-3].strip().split()[-1].isdigit()   # line:col: message


def test_abstract_constants_fold_in_binary64():
    """naga evaluates AbstractFloat in f64 and rounds once to f32:
    16777217.0 - 1.0 is 2^24 (in f32 arithmetic it would be 2^24 - 1)."""
    _lib_or_skip()
    code = wgc.wgsl_translate("const c: f32 = 16777217.0 - 1.0;\nfn f() -> f32 { return c; }")
    assert "0x1p+24f" in code
    code = wgc.wgsl_translate("fn f() -> f32 { return 16777217.0f - 1.0f; }")
    assert "0x1.fffffep+23f" in code


@pytest.mark.parametrize("shader", ["mvp_color", "mvp_texture", "colored_triangle", "clip_color"])
def test_shipped_shader_pipelines_compile_for_gfx950(shader):
    """The pipeline kernels of each shipped shader build with hipRTC."""
    _lib_or_skip()
    src = (SHADERS / f"{shader}.wgsl").read_text()
    wgc.wgsl_check_pipeline(src)


def test_stage_interface_mismatch_is_rejected():
    _lib_or_skip()
    src = """
    struct V { @builtin(position) p: vec4f, @location(0) @interpolate(linear) c: vec2f, }
    @vertex fn vs() -> V { return V(vec4f(), vec2f()); }
    @fragment fn fs(@location(0) @interpolate(linear) c: vec3f) -> @location(0) vec4f { return vec4f(c, 1.0); }
    """
    with pytest.raises(wgc.WgcError) as e:
        wgc.wgsl_check_pipeline(src)
    assert "does not match" in str(e.value)


def test_default_perspective_interpolation_needs_extended_semantics():
    """fragment.rs:343-345: perspective interpolation is todo!() in the reference."""
    _lib_or_skip()
    src = """
    struct V { @builtin(position) p: vec4f, @location(0) c: vec4f, }
    @vertex fn vs() -> V { return V(vec4f(), vec4f()); }
    @fragment fn fs(@location(0) c: vec4f) -> @location(0) vec4f { return c; }
    """
    with pytest.raises(wgc.WgcError) as e:
        wgc.wgsl_check_pipeline(src)
    assert "todo!()" in str(e.value)


# ---------------------------------------------------------------- the reference's own shaders

REFERENCE = Path("/root/reference")
# Interfaces read off the reference's shader files: per file, the entry
# points' user outputs (location, components), whether the vertex stage
# reads vertex_index, and the module's resources (group, binding, kind).
REFERENCE_SHADERS = {
    "wgpu-cpu-tests/src/tests/colored_triangle.wgsl": ([(0, 4)], True, []),
    "wgpu-cpu/examples/hello_shader.wgsl": ([(0, 4)], True, []),
    "wgpu-cpu/examples/hello_mesh.wgsl": ([(0, 4)], True, [(0, 0, "uniform")]),
    "wgpu-cpu/examples/hello_texture.wgsl": ([(0, 2)], True, [(0, 0, "uniform"), (1, 0, "texture"),
                                                             (1, 1, "sampler")]),
}


@pytest.mark.parametrize("rel", sorted(REFERENCE_SHADERS))
def test_reference_shader_files_translate(rel):
    """The reference's own WGSL files (read in place; skipped where the
    reference checkout is absent) pass the front end unchanged, with the
    interface their text declares, and their pipelines compile for gfx950."""
    path = REFERENCE / rel
    if not path.exists():
        pytest.skip("reference checkout not present")
    _lib_or_skip()
    src = path.read_text()
    iface = json.loads(wgc.wgsl_translate(src, 1))
    outputs, vertex_index, resources = REFERENCE_SHADERS[rel]
    entries = {e["stage"]: e for e in iface["entries"]}
    assert set(entries) == {"vertex", "fragment"}
    vs, fs = entries["vertex"], entries["fragment"]
    assert vs["name"] == "vs_main" and fs["name"] == "fs_main"
    assert [(o["location"], o["ncomp"]) for o in vs["outputs"]] == outputs
    # the fragment stage reads what the vertex stage writes, and writes one vec4
    assert [(i["location"], i["ncomp"]) for i in fs["inputs"]] == outputs
    assert [(o["location"], o["ncomp"]) for o in fs["outputs"]] == [(0, 4)]
    assert vs["position_out"] and vs["vertex_index"] == vertex_index
    assert not fs["may_discard"] and not fs["frag_depth"]
    assert [(r["group"], r["binding"], r["kind"]) for r in iface["resources"]] == resources
    wgc.wgsl_check_pipeline(src)
