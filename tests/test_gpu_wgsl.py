"""WGSL programs on the GPU, through the C ABI.

1. The reference's shader-compiler known-answer tests (tests/wgsl_kats.py,
   from naga-cranelift/src/tests.rs) run as device entry points.
2. The WGSL modules of the built-in programs (wgpu-cpu_amd/shaders) render
   the parity scenes bit-identically to the CPU oracle's render of the
   built-in program: the translator and the hipRTC-compiled kernels take the
   place of the hand-written shader functions without changing a bit.
3. WGSL-only features against closed forms: discard (equals back-face
   culling when it discards back faces), storage and uniform buffers
   (equal to the same positions given as vertex attributes), aborts.
"""
import copy
import struct
import sys

import numpy as np
import pytest

import wgc
from oracle import bridge
from wgc import runner, scenes

from test_gpu_parity import announce, assert_parity, clip_color_draw, random_clip_triangles
from test_wgsl import check_case
from wgsl_kats import kats

pytestmark = pytest.mark.gpu


# ---------------------------------------------------------------- KATs

@pytest.mark.parametrize("k", kats(), ids=lambda k: k.name)
def test_gpu_kat(gpu_device, k):
    announce(f"kat {k.name}")
    module = gpu_device.create_shader_module_wgsl(k.source)
    groups, keep = [], []
    for b in k.buffers:
        buf = gpu_device.create_buffer_init(b)
        keep.append(buf)
        groups = [gpu_device.create_bind_group([(0, buf)])]
    res = gpu_device.run_entry_point(module, None, 0, k.vertex_count, groups)
    per = []
    for r in res:
        pos = list(np.frombuffer(r[1:5].astype(np.uint32).tobytes(), np.float32).astype(float))
        per.append((int(r[0]), [int(x) for x in r[5:9]], pos))
    check_case(k, per)


# ---------------------------------------------------------------- built-in programs as WGSL

def _c1():
    return scenes.config1_plain_small()


def _colored(ff, cull):
    def make():
        s = scenes.colored_triangle(ff, cull)
        s.name = f"colored_triangle ff={ff} cull={cull}"
        return s
    return make


def _clipping():
    vb = random_clip_triangles(4000, 91, 320, 240, w_range=(0.3, 3.0), spread=2.0, size_px=(5.0, 400.0))
    vb[:, :, 2] = (vb[:, :, 2] / vb[:, :, 3] * 2.0 - 0.5) * vb[:, :, 3]
    return scenes.PassSpec(320, 240, has_depth=True, clear_depth=1.0,
                           draws=[clip_color_draw(vb, depth=(wgc.LESS, True))], name="clipping_all_planes")


def _depth_ties():
    w, h = 160, 120
    vb = random_clip_triangles(3000, 9, w, h, z_levels=[0.25, 0.5, 0.75], size_px=(3.0, 30.0))
    init = np.where(scenes.splitmix_f32(77, w * h) < 0.5, 0.5, 0.625).astype(np.float32)
    return scenes.PassSpec(w, h, clear_color=(0.2, 0.2, 0.2, 1.0), has_depth=True, clear_depth=None,
                           depth_init=init.tobytes(), draws=[clip_color_draw(vb, depth=(wgc.LESS_EQUAL, True))],
                           name="depth ties")


@pytest.mark.parametrize("make", [
    _c1, _colored(wgc.FRONT_CW, wgc.CULL_BACK), _colored(wgc.FRONT_CCW, wgc.CULL_NONE),
    lambda: scenes.config2_textured(extended=False),
    lambda: scenes.config3_overdraw(n=20_000, width=640, height=610),
    _clipping, _depth_ties,
], ids=["c1", "colored_cw_back", "colored_ccw_none", "c2_1080p", "c3_small", "clipping", "depth_ties"])
def test_wgsl_programs_match_the_oracle(gpu_device, make):
    assert_parity(scenes.as_wgsl(make()), gpu_device)


@pytest.mark.parametrize("make", [
    lambda: scenes.config2_textured(extended=True, width=640, height=360),
    scenes.perspective_quad, scenes.bilinear_strip,
    lambda: scenes.config3_overdraw(n=20_000, width=640, height=610, blend=True),
], ids=["c2_extended", "perspective_quad", "bilinear_strip", "c3_blend"])
def test_wgsl_programs_extended(gpu_device_ext, make):
    """@interpolate(perspective), bilinear sampling, sRGB and blending."""
    assert_parity(scenes.as_wgsl(make()), gpu_device_ext)


def test_wgsl_batched_passes(gpu_device):
    """The batched form (one launch group for many passes), launched directly
    and replayed as a hipGraph."""
    announce("wgsl batched")
    specs = []
    for i, (w, h, tris, tex) in enumerate(scenes.batched_sizes(0x5EED_0004, 12)):
        w, h, tris = max(16, w // 4), max(16, h // 4), min(tris, 5000.0)
        specs.append(scenes.as_wgsl(scenes.batched_pass(0x5EED_0004, i, w, h, tris, tex)))
    graphs = gpu_device.get_option(wgc.OPTION_GRAPHS)
    try:
        for use_graphs in (0, 1):
            gpu_device.set_option(wgc.OPTION_GRAPHS, use_graphs)
            pp = runner.PreparedPasses(gpu_device, specs)
            for _ in range(3):      # with graphs: the second captures the graph, the third replays it
                pp.reset_attachments()
                pp.submit()
                results = pp.results()
                for s, r in zip(specs, results):
                    ref = bridge.render(s)
                    assert np.array_equal(r.color, ref.color), s.name
                    assert np.array_equal(r.depth.view(np.uint32), ref.depth.view(np.uint32)), s.name
    finally:
        gpu_device.set_option(wgc.OPTION_GRAPHS, graphs)


# ---------------------------------------------------------------- WGSL-only features

CLIP_VS = """
struct VsIn {
    @location(0) position: vec4f,
    @location(1) color: vec4f,
}
struct VsOut {
    @builtin(position) position: vec4f,
    @location(0) @interpolate(linear) color: vec4f,
}
@vertex
fn vs_main(v: VsIn) -> VsOut {
    return VsOut(v.position, v.color);
}
"""


def _overdraw_both_windings(n=6000, seed=41, w=256, h=192):
    vb = random_clip_triangles(n, seed, w, h, size_px=(4.0, 50.0))
    return vb, w, h


def test_discard_of_back_faces_equals_back_face_culling(gpu_device):
    """A fragment shader that discards when !front_facing leaves exactly what
    culling back faces leaves (colour and depth); the pipeline runs the
    ordered path because the shader can discard."""
    announce("discard == cull back")
    vb, w, h = _overdraw_both_windings()
    ref_spec = scenes.PassSpec(w, h, has_depth=True, clear_depth=1.0,
                               draws=[clip_color_draw(vb, depth=(wgc.LESS, True), cull_mode=wgc.CULL_BACK)])
    ref = bridge.render(ref_spec)
    spec = copy.deepcopy(ref_spec)
    spec.draws[0].cull_mode = wgc.CULL_NONE
    spec.draws[0].wgsl = CLIP_VS + """
@fragment
fn fs_main(@builtin(front_facing) front: bool, f: VsOut) -> @location(0) vec4f {
    if !front {
        discard;
    }
    return f.color;
}
"""
    got = runner.render(spec, gpu_device)
    assert np.array_equal(got.color, ref.color)
    assert np.array_equal(got.depth.view(np.uint32), ref.depth.view(np.uint32))


def test_storage_buffer_vertices_equal_vertex_attributes(gpu_device):
    """Vertices read from a storage buffer by vertex_index (a struct load with
    a dynamic index into a runtime-sized array) give the render of the same
    data as vertex attributes."""
    announce("storage vertices")
    vb, w, h = _overdraw_both_windings(seed=43)
    ref_spec = scenes.PassSpec(w, h, has_depth=True, clear_depth=1.0,
                               draws=[clip_color_draw(vb, depth=(wgc.LESS, True))])
    ref = bridge.render(ref_spec)
    spec = copy.deepcopy(ref_spec)
    d = spec.draws[0]
    d.uniform = vb.tobytes()          # bound at @group(0) @binding(0)
    d.vertex_data = b""
    d.attributes = []
    d.vertex_stride = 0
    d.wgsl = """
struct Vert { position: vec4f, color: vec4f, }
@group(0) @binding(0) var<storage, read> verts: array<Vert>;
struct VsOut {
    @builtin(position) position: vec4f,
    @location(0) @interpolate(linear) color: vec4f,
}
@vertex
fn vs_main(@builtin(vertex_index) i: u32) -> VsOut {
    let v = verts[i];
    if arrayLength(&verts) != """ + str(vb.shape[0] * 3) + """u {
        return VsOut(vec4f(), vec4f());
    }
    return VsOut(v.position, v.color);
}
@fragment
fn fs_main(f: VsOut) -> @location(0) vec4f {
    return f.color;
}
"""
    got = runner.render(spec, gpu_device)
    assert np.array_equal(got.color, ref.color)
    assert np.array_equal(got.depth.view(np.uint32), ref.depth.view(np.uint32))


def test_uniform_struct_layout(gpu_device):
    """A uniform struct with vec3 / f32 / vec2 members (offsets 0, 12, 16 in
    the WGSL layout) transforms the positions; the same transform in numpy
    binary32 arithmetic (separately rounded, no fused multiply-add) as
    vertex attributes gives the same render."""
    announce("uniform layout")
    vb, w, h = _overdraw_both_windings(seed=47)
    scale = np.array([0.75, 1.25, 1.0], np.float32)
    bias = np.float32(0.125)
    offset = np.array([0.1, -0.2], np.float32)
    moved = vb.copy()
    moved[:, :, 0] = (vb[:, :, 0] * scale[0]).astype(np.float32) + offset[0] * vb[:, :, 3]
    moved[:, :, 1] = (vb[:, :, 1] * scale[1]).astype(np.float32) + offset[1] * vb[:, :, 3]
    moved[:, :, 2] = vb[:, :, 2] * scale[2] + bias
    ref_spec = scenes.PassSpec(w, h, has_depth=True, clear_depth=1.0,
                               draws=[clip_color_draw(moved.astype(np.float32), depth=(wgc.LESS, True))])
    ref = bridge.render(ref_spec)
    spec = scenes.PassSpec(w, h, has_depth=True, clear_depth=1.0,
                           draws=[clip_color_draw(vb, depth=(wgc.LESS, True))])
    d = spec.draws[0]
    u = struct.pack("<3ff2f2f", *scale, bias, *offset, 0.0, 0.0)
    assert len(u) == 32
    d.uniform = u
    d.wgsl = """
struct U { scale: vec3f, bias: f32, offset: vec2f, }
@group(0) @binding(0) var<uniform> u: U;
struct VsIn { @location(0) position: vec4f, @location(1) color: vec4f, }
struct VsOut { @builtin(position) position: vec4f, @location(0) @interpolate(linear) color: vec4f, }
@vertex
fn vs_main(v: VsIn) -> VsOut {
    let p = v.position;
    let x = p.x * u.scale.x + u.offset.x * p.w;
    let y = p.y * u.scale.y + u.offset.y * p.w;
    let z = p.z * u.scale.z + u.bias;
    return VsOut(vec4f(x, y, z, p.w), v.color);
}
@fragment
fn fs_main(f: VsOut) -> @location(0) vec4f { return f.color; }
"""
    got = runner.render(spec, gpu_device)
    assert np.array_equal(got.color, ref.color)
    assert np.array_equal(got.depth.view(np.uint32), ref.depth.view(np.uint32))


def test_shader_abort_is_an_execution_error(gpu_device):
    """An integer division by zero aborts the invocation (binary.rs:451-476);
    the reference panics ("fragment shader failed"), here the submission
    reports an execution error and the device stays usable."""
    announce("shader abort")
    vb, w, h = _overdraw_both_windings(n=50, seed=53)
    spec = scenes.PassSpec(w, h, draws=[clip_color_draw(vb)])
    spec.draws[0].wgsl = CLIP_VS + """
@fragment
fn fs_main(f: VsOut) -> @location(0) vec4f {
    let zero = i32(f.color.x * 0.0);
    let q = 7 / zero;
    return vec4f(f32(q), 0.0, 0.0, 1.0);
}
"""
    with pytest.raises(wgc.WgcError) as e:
        runner.render(spec, gpu_device)
    assert e.value.code == wgc.ERR_EXECUTION
    assert_parity(scenes.config1_plain_small(), gpu_device)
