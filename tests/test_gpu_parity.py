"""Parity of the HIP path with the CPU oracle, through the C ABI.

Coverage (per-pixel fragment counts), depth (bits) and colour (bytes) must be
identical to the oracle's render of the same seeded scene; the fragment total
must match.  The workloads run at their full sizes except where noted.
"""
import sys

import numpy as np
import pytest

import wgc
from oracle import bridge
from wgc import runner, scenes

pytestmark = pytest.mark.gpu


def announce(name):
    print(f"\n[gpu case] {name}", flush=True)
    sys.stdout.flush()


def device_for(spec, dev_ref, dev_ext):
    return dev_ext if spec.semantics == wgc.SEMANTICS_EXTENDED else dev_ref


def assert_parity(spec, dev):
    announce(spec.name or "scene")
    ref = bridge.render(spec, frag_counts=True)
    got = runner.render(spec, dev, fragment_counts=True)
    assert got.fragments == ref.fragments, (got.fragments, ref.fragments)
    if not np.array_equal(got.frag_count, ref.frag_count):
        diff = np.argwhere(got.frag_count != ref.frag_count)
        y, x = diff[0]
        raise AssertionError(f"coverage differs at {len(diff)} px, first ({x},{y}): "
                             f"gpu {got.frag_count[y, x]} oracle {ref.frag_count[y, x]}")
    if ref.depth is not None:
        gd, rd = got.depth.view(np.uint32), ref.depth.view(np.uint32)
        if not np.array_equal(gd, rd):
            diff = np.argwhere(gd != rd)
            y, x = diff[0]
            raise AssertionError(f"depth differs at {len(diff)} px, first ({x},{y}): "
                                 f"gpu {got.depth[y, x]!r} oracle {ref.depth[y, x]!r}")
    if ref.color is not None:
        if not np.array_equal(got.color, ref.color):
            diff = np.argwhere((got.color != ref.color).any(-1))
            y, x = diff[0]
            maxerr = int(np.abs(got.color.astype(int) - ref.color.astype(int)).max())
            raise AssertionError(f"colour differs at {len(diff)} px (max {maxerr} LSB), first ({x},{y}): "
                                 f"gpu {got.color[y, x]} oracle {ref.color[y, x]}")
    return ref, got


def random_clip_triangles(n, seed, width, height, z_levels=None, w_range=(1.0, 1.0), spread=1.0,
                          size_px=(4.0, 60.0), alpha=(1.0, 1.0)):
    """n independent triangles given in clip space; positions spread over
    [-spread, spread] in NDC, optional quantised z (for depth ties), random w."""
    u = scenes.splitmix_f32(seed, n * 12).astype(np.float64).reshape(n, 12)
    cx = (u[:, 0] * 2 - 1) * spread
    cy = (u[:, 1] * 2 - 1) * spread
    size = size_px[0] + u[:, 2] * (size_px[1] - size_px[0])
    ang = u[:, 3:6] * 2 * np.pi
    px = cx[:, None] + np.cos(ang) * size[:, None] * 2.0 / width
    py = cy[:, None] + np.sin(ang) * size[:, None] * 2.0 / height
    if z_levels is not None:
        z = np.asarray(z_levels)[(u[:, 6] * len(z_levels)).astype(int) % len(z_levels)][:, None] * np.ones((1, 3))
    else:
        z = u[:, 6:9]
    w = w_range[0] + u[:, 9:12] * (w_range[1] - w_range[0])
    vb = np.zeros((n, 3, 8), np.float32)
    vb[:, :, 0] = px * w
    vb[:, :, 1] = py * w
    vb[:, :, 2] = z * w
    vb[:, :, 3] = w
    col = scenes.splitmix_f32(seed ^ 0xABCDEF, n * 4).reshape(n, 1, 4)
    vb[:, :, 4:8] = col
    vb[:, :, 7] = alpha[0] + col[:, :, 3] * (alpha[1] - alpha[0])
    return vb


def clip_color_draw(vb, **kw):
    return scenes.DrawSpec(program=wgc.PROGRAM_CLIP_COLOR, vertex_data=vb.tobytes(), vertex_stride=32,
                           attributes=[(wgc.FLOAT32X4, 0, 0), (wgc.FLOAT32X4, 16, 1)],
                           count=vb.shape[0] * 3, **kw)


# ---------------------------------------------------------------- workloads

def test_c1_plain_small(gpu_device):
    assert_parity(scenes.config1_plain_small(), gpu_device)


@pytest.mark.parametrize("ff,cull", [(wgc.FRONT_CW, wgc.CULL_BACK), (wgc.FRONT_CW, wgc.CULL_FRONT),
                                     (wgc.FRONT_CCW, wgc.CULL_BACK), (wgc.FRONT_CCW, wgc.CULL_NONE)])
def test_colored_triangle_scenes(gpu_device, ff, cull):
    spec = scenes.colored_triangle(ff, cull)
    spec.name = f"colored_triangle ff={ff} cull={cull}"
    assert_parity(spec, gpu_device)


def test_c2_textured_reference_1080p(gpu_device):
    assert_parity(scenes.config2_textured(extended=False), gpu_device)


def test_c2_textured_extended(gpu_device_ext):
    spec = scenes.config2_textured(extended=True, width=640, height=360)
    assert_parity(spec, gpu_device_ext)


def test_c2_textured_extended_1080p(gpu_device_ext):
    """The benchmark's headline pass at its full size: perspective-correct uv,
    bilinear Repeat sampling of the sRGB texture, 1920x1080."""
    assert_parity(scenes.config2_textured(extended=True), gpu_device_ext)


def test_c2_textured_extended_1080p_wgsl(gpu_device_ext):
    """The same pass with the WGSL forms of its shaders (the drop-in path)."""
    assert_parity(scenes.as_wgsl(scenes.config2_textured(extended=True)), gpu_device_ext)


@pytest.mark.parametrize("make", [scenes.perspective_quad, scenes.bilinear_strip])
def test_extension_closed_form_scenes(gpu_device_ext, make):
    assert_parity(make(), gpu_device_ext)


def test_c3_overdraw_reference_small(gpu_device):
    spec = scenes.config3_overdraw(n=20_000, width=640, height=610)
    assert_parity(spec, gpu_device)


def test_c3_overdraw_blend_small(gpu_device_ext):
    spec = scenes.config3_overdraw(n=20_000, width=640, height=610, blend=True)
    assert_parity(spec, gpu_device_ext)


def test_c3_overdraw_reference_full(gpu_device):
    assert_parity(scenes.config3_overdraw(), gpu_device)


def test_c3_overdraw_blend_full(gpu_device_ext):
    """200k triangles over 2560x2440 blended src-over in API order: every
    colour byte equal to the oracle's sequential blend."""
    assert_parity(scenes.config3_overdraw(blend=True), gpu_device_ext)


# ---------------------------------------------------------------- pipeline corners

def test_clipping_all_planes(gpu_device):
    # triangles spread to 2x the viewport, w in [0.3, 3], z in [-0.5, 1.5]:
    # crosses every clip plane, including near (z < 0) and far (z > w)
    vb = random_clip_triangles(4000, 91, 320, 240, w_range=(0.3, 3.0), spread=2.0, size_px=(5.0, 400.0))
    vb[:, :, 2] = (vb[:, :, 2] / vb[:, :, 3] * 2.0 - 0.5) * vb[:, :, 3]
    spec = scenes.PassSpec(320, 240, has_depth=True, clear_depth=1.0,
                           draws=[clip_color_draw(vb, depth=(wgc.LESS, True))], name="clipping_all_planes")
    assert_parity(spec, gpu_device)


def test_big_triangles_and_overlap(gpu_device):
    # footprints of many tiles go through the big-triangle list
    vb = random_clip_triangles(300, 17, 512, 384, size_px=(100.0, 700.0), z_levels=[0.2, 0.4, 0.6])
    spec = scenes.PassSpec(512, 384, has_depth=True, clear_depth=1.0,
                           draws=[clip_color_draw(vb, depth=(wgc.LESS, True))], name="big_triangles")
    assert_parity(spec, gpu_device)


def _crowded(n=4000, seed=23, **kw):
    """n small triangles around the centre of a 1024x1024 target (4096 tiles
    of 16x16): the four tiles there collect about n/4 entries each, past
    their bin capacity, so the spill list is read."""
    vb = random_clip_triangles(n, seed, 1024, 1024, spread=0.006, size_px=(1.5, 5.0), **kw)
    return vb


def test_crowded_tiles_spill_their_bins(gpu_device):
    vb = _crowded()
    spec = scenes.PassSpec(1024, 1024, has_depth=True, clear_depth=1.0,
                           draws=[clip_color_draw(vb, depth=(wgc.LESS, True))], name="crowded_tiles")
    assert_parity(spec, gpu_device)


def test_crowded_tiles_spill_ordered(gpu_device_ext):
    vb = _crowded(n=3000, seed=29, alpha=(0.2, 0.8))
    spec = scenes.PassSpec(1024, 1024, clear_color=(0.1, 0.2, 0.3, 1.0),
                           draws=[clip_color_draw(vb, blend=wgc.BLEND_ALPHA)], semantics=wgc.SEMANTICS_EXTENDED,
                           name="crowded_tiles_blend")
    assert_parity(spec, gpu_device_ext)


def test_crowded_tiles_spill_batched(gpu_device):
    """The spill list of a batched group holds global tiles: a crowded pass
    next to ordinary ones renders like each pass alone."""
    announce("crowded tiles, batched")
    crowded = scenes.PassSpec(1024, 1024, has_depth=True, clear_depth=1.0,
                              draws=[clip_color_draw(_crowded(seed=31), depth=(wgc.LESS, True))], name="crowded_b")
    others = [scenes.PassSpec(160 + 16 * i, 120, has_depth=True, clear_depth=1.0,
                              draws=[clip_color_draw(random_clip_triangles(500, 40 + i, 160 + 16 * i, 120),
                                                     depth=(wgc.LESS, True))], name=f"other_{i}") for i in range(3)]
    specs = [others[0], crowded, others[1], others[2]]
    res, total = runner.render_batch(specs, gpu_device)
    want = 0
    for s, r in zip(specs, res):
        ref = bridge.render(s)
        want += ref.fragments
        assert np.array_equal(r.color, ref.color), s.name
        assert np.array_equal(r.depth.view(np.uint32), ref.depth.view(np.uint32)), s.name
    assert total == want


@pytest.mark.parametrize("compare", list(range(1, 9)))
@pytest.mark.parametrize("write", [True, False])
def test_depth_compare_and_ties(gpu_device, compare, write):
    # quantised depths give many equal-depth ties; a random initial depth
    # buffer (LoadOp::Load) makes every compare function take both branches
    w, h = 160, 120
    vb = random_clip_triangles(3000, 5 + compare, w, h, z_levels=[0.25, 0.5, 0.75], size_px=(3.0, 30.0))
    init = np.where(scenes.splitmix_f32(77, w * h) < 0.5, 0.5, 0.625).astype(np.float32)
    spec = scenes.PassSpec(w, h, clear_color=(0.2, 0.2, 0.2, 1.0), has_depth=True, clear_depth=None,
                           depth_init=init.tobytes(),
                           draws=[clip_color_draw(vb, depth=(compare, write))],
                           name=f"depth compare={compare} write={write}")
    assert_parity(spec, gpu_device)


def test_multi_draw_load_ops_viewport_scissor_instances(gpu_device):
    w, h = 200, 150
    sphere = scenes.config1_plain_small()
    d0 = sphere.draws[0]
    d0.viewport = (10.0, 5.0, 120.0, 120.0, 0.0, 1.0)
    d0.scissor = (20, 10, 150, 100)
    # u16 indices with a base vertex, two instances
    idx32 = np.frombuffer(d0.index_data, np.uint32)
    d1 = scenes.DrawSpec(**{**d0.__dict__})
    d1.index_data = (idx32.astype(np.int64) - 0).astype(np.uint16).tobytes()
    d1.index_format = 2
    d1.viewport = (60.0, 20.0, 130.0, 110.0, 0.0, 1.0)
    d1.scissor = None
    d1.instance_count = 2
    d1.depth = (wgc.LESS_EQUAL, True)
    vb = random_clip_triangles(500, 23, w, h, size_px=(4.0, 40.0))
    d2 = clip_color_draw(vb, depth=(wgc.GREATER_EQUAL, False))
    init_c = (scenes.splitmix_u64(5, w * h) & np.uint64(0xFFFFFFFF)).astype(np.uint32).tobytes()
    init_d = scenes.splitmix_f32(6, w * h).tobytes()
    spec = scenes.PassSpec(w, h, clear_color=None, color_init=init_c, has_depth=True, clear_depth=None,
                           depth_init=init_d, draws=[d0, d1, d2], name="multi_draw")
    assert_parity(spec, gpu_device)


def test_base_vertex_offset(gpu_device):
    spec = scenes.config1_plain_small()
    d = spec.draws[0]
    # prepend 10 junk vertices and draw with base_vertex = 10 (index.rs:144-207)
    junk = np.full((10, 7), 1e9, np.float32).tobytes()
    d.vertex_data = junk + d.vertex_data
    d.base_vertex = 10
    spec.name = "base_vertex"
    assert_parity(spec, gpu_device)


def test_index_out_of_range_is_an_error(gpu_device):
    announce("index_out_of_range")
    spec = scenes.config1_plain_small()
    idx = np.frombuffer(spec.draws[0].index_data, np.uint32).copy()
    idx[7] = 100000
    spec.draws[0].index_data = idx.tobytes()
    with pytest.raises(bridge.OracleError):
        bridge.render(spec)
    with pytest.raises(wgc.WgcError):
        runner.render(spec, gpu_device)
    # the device stays usable
    assert_parity(scenes.config1_plain_small(), gpu_device)


def test_resubmission_is_idempotent(gpu_device):
    announce("resubmission")
    spec = scenes.config1_plain_small()
    pp = runner.PreparedPass(gpu_device, spec)
    pp.submit()
    a = pp.result()
    for _ in range(3):
        pp.submit()
    b = pp.result()
    assert np.array_equal(a.color, b.color) and np.array_equal(a.depth, b.depth)
    assert a.fragments == b.fragments
