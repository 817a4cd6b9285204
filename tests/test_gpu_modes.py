"""Reference paths outside the benchmark workloads, each with the built-in
program and its WGSL form, against the oracle: sampler address modes,
flat interpolation, BGRA targets and vertex buffers stepped per instance."""
import sys

import numpy as np
import pytest

import wgc
from oracle import bridge
from wgc import runner, scenes

pytestmark = pytest.mark.gpu

FORMS = ["builtin", "wgsl"]


def announce(name):
    print(f"\n[gpu case] {name}", flush=True)
    sys.stdout.flush()


def check(spec, dev_ref, dev_ext, form):
    if form == "wgsl":
        spec = scenes.as_wgsl(spec)
    announce(spec.name)
    dev = dev_ext if spec.semantics == wgc.SEMANTICS_EXTENDED else dev_ref
    ref = bridge.render(spec, frag_counts=True)
    got = runner.render(spec, dev, fragment_counts=True)
    assert got.fragments == ref.fragments
    assert np.array_equal(got.frag_count, ref.frag_count), "coverage"
    if ref.depth is not None:
        assert np.array_equal(got.depth.view(np.uint32), ref.depth.view(np.uint32)), "depth"
    if not np.array_equal(got.color, ref.color):
        diff = np.argwhere((got.color != ref.color).any(-1))
        y, x = diff[0]
        raise AssertionError(f"colour differs at {len(diff)} px, first ({x},{y}): gpu {got.color[y, x]} "
                             f"oracle {ref.color[y, x]}")
    return ref


ADDR = [(wgc.CLAMP_TO_EDGE, wgc.CLAMP_TO_EDGE), (wgc.MIRROR_REPEAT, wgc.MIRROR_REPEAT),
        (wgc.CLAMP_TO_EDGE, wgc.MIRROR_REPEAT), (wgc.MIRROR_REPEAT, wgc.REPEAT)]


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("mode", ["reference", "extended_nearest", "extended_bilinear"])
@pytest.mark.parametrize("au,av", ADDR)
def test_sampler_address_modes(gpu_device, gpu_device_ext, au, av, mode, form):
    """uv from -1.3 to 2.7 over an 8x8 texture: ClampToEdge clamps,
    MirrorRepeat reflects every other period (binding.rs:150-164)."""
    spec = scenes.addressing_quad(au, av, extended=mode != "reference", bilinear=mode == "extended_bilinear")
    ref = check(spec, gpu_device, gpu_device_ext, form)
    # the image depends on the mode: ClampToEdge repeats the edge texel
    # along the columns left of u = 0
    if au == wgc.CLAMP_TO_EDGE and mode == "reference":
        left = ref.color[:, :10]
        assert (left == left[:, :1]).all()


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("scene", ["sphere", "overdraw"])
def test_flat_interpolation(gpu_device, gpu_device_ext, scene, form):
    """@interpolate(flat): every fragment takes the first vertex's value
    (fragment.rs:287-297, 395-405)."""
    if scene == "sphere":
        spec = scenes.config1_plain_small()
    else:
        spec = scenes.config3_overdraw(n=3000, width=320, height=300)
    spec.draws[0].interp = wgc.INTERP_FLAT
    spec.name = f"flat {scene}"
    ref = check(spec, gpu_device, gpu_device_ext, form)
    # a flat triangle is one colour: at most one distinct colour per
    # triangle, and the clear colour
    tris = spec.draws[0].count // 3
    assert len(np.unique(ref.color.reshape(-1, 4), axis=0)) <= tris + 1


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("fmt", [wgc.BGRA8_UNORM, wgc.BGRA8_UNORM_SRGB])
@pytest.mark.parametrize("blend", [False, True])
def test_bgra_targets(gpu_device, gpu_device_ext, fmt, blend, form):
    """Bgra8Unorm(Srgb) targets store blue first (texture.rs:402-407); the
    blend reads the destination back in the same order."""
    if blend:
        spec = scenes.config3_overdraw(n=4000, width=320, height=300, blend=True)
    else:
        spec = scenes.config1_plain_small()
    spec.color_format = fmt
    spec.name = f"bgra fmt={fmt} blend={blend}"
    ref = check(spec, gpu_device, gpu_device_ext, form)
    rgba = scenes.PassSpec(**{**spec.__dict__, "color_format": wgc.RGBA8_UNORM})
    ref_rgba = bridge.render(rgba)
    assert np.array_equal(ref.color[..., [2, 1, 0, 3]], ref_rgba.color)


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("first", [0, 3])
@pytest.mark.parametrize("mode", ["less", "less_equal", "blend"])
def test_instance_step_buffer(gpu_device, gpu_device_ext, first, mode, form):
    """Colours from a buffer stepped per instance (vertex.rs:191-195),
    instance_count 4, first_instance 0 and 3."""
    depth = {"less": (wgc.LESS, True), "less_equal": (wgc.LESS_EQUAL, True), "blend": None}[mode]
    spec = scenes.instanced_triangles(first_instance=first, depth=depth, blend=mode == "blend")
    ref = check(spec, gpu_device, gpu_device_ext, form)
    if mode != "blend":
        # Less keeps the first instance's colour, LessEqual the last one's
        cols = np.frombuffer(spec.draws[0].instance_buffer[0], np.float32).reshape(-1, 4)
        k = first if mode == "less" else first + 3
        want = np.array([min(255, int(c * 255)) for c in cols[k]], np.uint8)
        covered = ref.depth < 1.0
        assert covered.any()
        assert (ref.color[covered][:, :3] == want[:3]).all()
