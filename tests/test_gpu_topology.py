"""Points, lines, line strips, triangle strips and primitive restart on the
GPU against the CPU oracle (coverage, depth bits and colour bytes), through
the C ABI, with the built-in programs and their WGSL forms."""
import copy

import numpy as np
import pytest

import wgc
from oracle import bridge
from wgc import runner, scenes

from test_gpu_parity import announce, assert_parity, clip_color_draw, random_clip_triangles

pytestmark = pytest.mark.gpu

POINTS, LINES, LINE_STRIP, TRIANGLES, TRIANGLE_STRIP = range(5)


def vertices(n, seed, w, h, spread=1.2, z_levels=None, w_range=(1.0, 1.0)):
    """n clip-space vertices (pos vec4 + colour vec4) spread over the target."""
    vb = random_clip_triangles((n + 2) // 3, seed, w, h, spread=spread, size_px=(10.0, 200.0),
                               z_levels=z_levels, w_range=w_range)
    return vb.reshape(-1, 8)[:n].copy()


def draw(vb, topology, **kw):
    d = clip_color_draw(vb.reshape(-1, 1, 8), **kw)
    d.count = vb.shape[0] if "count" not in kw else kw["count"]
    d.topology = topology
    return d


def both(spec, dev):
    """The built-in program and its WGSL form."""
    assert_parity(spec, dev)
    assert_parity(scenes.as_wgsl(spec), dev)


def test_colored_triangle_lines(gpu_device):
    """colored_triangle.rs:188-198: LineStrip over vertices 0..4."""
    spec = scenes.colored_triangle(wgc.FRONT_CW, wgc.CULL_NONE, vertices=4, topology=LINE_STRIP)
    spec.name = "colored_triangle_lines"
    both(spec, gpu_device)


@pytest.mark.parametrize("topology", [POINTS, LINES, LINE_STRIP], ids=["points", "lines", "line_strip"])
def test_points_and_lines_with_depth_and_clipping(gpu_device, topology):
    w, h = 320, 240
    vb = vertices(3000, 11 + topology, w, h, spread=1.6, z_levels=[0.25, 0.5, 0.75], w_range=(0.5, 2.0))
    # some vertices beyond near / far so that lines clip there as well
    vb[::7, 2] = -0.3 * vb[::7, 3]
    vb[3::11, 2] = 1.4 * vb[3::11, 3]
    spec = scenes.PassSpec(w, h, has_depth=True, clear_depth=1.0,
                           draws=[draw(vb, topology, depth=(wgc.LESS, True))], name=f"topology {topology}")
    both(spec, gpu_device)


@pytest.mark.parametrize("compare,write", [(wgc.LESS_EQUAL, True), (wgc.GREATER, False), (wgc.ALWAYS, True),
                                           (wgc.NOT_EQUAL, True)])
def test_line_depth_modes_and_load(gpu_device, compare, write):
    """Every resolve kind of the point/line path, including the ordered one
    (NotEqual + write), over a loaded depth buffer with ties."""
    w, h = 160, 120
    vb = vertices(1200, 29, w, h, z_levels=[0.5, 0.625])
    init = np.where(scenes.splitmix_f32(78, w * h) < 0.5, 0.5, 0.625).astype(np.float32)
    spec = scenes.PassSpec(w, h, clear_color=(0.2, 0.2, 0.2, 1.0), has_depth=True, clear_depth=None,
                           depth_init=init.tobytes(), draws=[draw(vb, LINES, depth=(compare, write))],
                           name=f"lines compare={compare} write={write}")
    both(spec, gpu_device)


def test_blended_lines_and_points(gpu_device_ext):
    """src-over blending takes the ordered point/line path (extension)."""
    w, h = 128, 96
    vb = vertices(600, 31, w, h)
    vb[:, 7] = 0.5
    d1 = draw(vb, LINE_STRIP, blend=wgc.BLEND_ALPHA)
    d2 = draw(vertices(300, 37, w, h), POINTS, blend=wgc.BLEND_ALPHA)
    spec = scenes.PassSpec(w, h, draws=[d1, d2], semantics=wgc.SEMANTICS_EXTENDED, name="blended lines")
    both(spec, gpu_device_ext)


@pytest.mark.parametrize("kind", ["blend", "not_equal"])
def test_ordered_lines_over_many_tiles(gpu_device_ext, kind):
    """The ordered point/line path over a target of 20x15 screen tiles:
    long lines (up to 700 px, many leaving the target and clipped) cross
    many tiles and overlap each other, so every pixel must see its
    fragments in API order whichever tile applies them."""
    w, h = 640, 480
    u = scenes.splitmix_f32(4242, 1500 * 4).reshape(1500, 4).astype(np.float64)
    vb = np.zeros((3000, 8), np.float32)
    vb[0::2, 0] = u[:, 0] * 3.0 - 1.5
    vb[0::2, 1] = u[:, 1] * 3.0 - 1.5
    vb[1::2, 0] = u[:, 2] * 3.0 - 1.5
    vb[1::2, 1] = u[:, 3] * 3.0 - 1.5
    vb[:, 2] = np.where(np.arange(3000) % 3 == 0, 0.5, 0.25)
    vb[:, 3] = 1.0
    vb[:, 4:8] = scenes.splitmix_f32(4243, 3000 * 4).reshape(3000, 4)
    vb[:, 7] = 0.4
    pts = vertices(400, 4244, w, h)
    pts[:, 7] = 0.6
    if kind == "blend":
        draws = [draw(vb, LINES, blend=wgc.BLEND_ALPHA), draw(pts, POINTS, blend=wgc.BLEND_ALPHA)]
        spec = scenes.PassSpec(w, h, draws=draws, semantics=wgc.SEMANTICS_EXTENDED, name="ordered lines, blend")
    else:
        init = np.where(scenes.splitmix_f32(4245, w * h) < 0.5, 0.5, 0.25).astype(np.float32)
        draws = [draw(vb, LINES, depth=(wgc.NOT_EQUAL, True)), draw(pts, POINTS, depth=(wgc.NOT_EQUAL, True))]
        spec = scenes.PassSpec(w, h, has_depth=True, clear_depth=None, depth_init=init.tobytes(), draws=draws,
                               semantics=wgc.SEMANTICS_EXTENDED, name="ordered lines, not-equal + write")
    both(spec, gpu_device_ext)


def test_triangle_strip(gpu_device):
    w, h = 256, 192
    vb = vertices(2000, 41, w, h, spread=1.5, w_range=(0.4, 2.0))
    spec = scenes.PassSpec(w, h, has_depth=True, clear_depth=1.0,
                           draws=[draw(vb, TRIANGLE_STRIP, depth=(wgc.LESS, True), cull_mode=wgc.CULL_BACK)],
                           name="triangle strip")
    both(spec, gpu_device)


@pytest.mark.parametrize("fmt", [2, 4])
@pytest.mark.parametrize("topology", [TRIANGLE_STRIP, LINE_STRIP], ids=["triangle_strip", "line_strip"])
def test_primitive_restart(gpu_device, fmt, topology):
    """Index MAX restarts the strip (index.rs:72-106, ResetTriStripIter):
    segments of random length, runs of separators, and a separator-free
    stretch longer than the separator summary's chunk."""
    w, h = 256, 192
    n = 1500
    vb = vertices(n, 43 + fmt, w, h, z_levels=[0.3, 0.6])
    rng = np.random.default_rng(fmt * 10 + topology)
    sep = 0xFFFF if fmt == 2 else 0xFFFFFFFF
    idx = []
    v = 0
    while len(idx) < 4000:
        if 1000 < len(idx) < 1700:
            idx.append(v % n)            # a long strip without separators
        elif rng.random() < 0.12:
            idx.extend([sep] * int(rng.integers(1, 3)))
        else:
            idx.append(v % n)
        v += 1
    d = draw(vb, topology, depth=(wgc.LESS, True))
    d.index_data = np.asarray(idx, np.uint16 if fmt == 2 else np.uint32).tobytes()
    d.index_format = fmt
    d.strip_index_format = fmt
    d.count = len(idx)
    spec = scenes.PassSpec(w, h, has_depth=True, clear_depth=1.0, draws=[d], name=f"restart fmt={fmt}")
    both(spec, gpu_device)


def test_instanced_points_and_lines_with_viewport_and_scissor(gpu_device):
    w, h = 200, 150
    vb = vertices(400, 47, w, h)
    d1 = draw(vb, LINES, depth=(wgc.LESS_EQUAL, True))
    d1.instance_count = 3
    d1.viewport = (20.0, 10.0, 150.0, 120.0, 0.0, 1.0)
    d1.scissor = (30, 20, 120, 90)
    d2 = draw(vb, POINTS, depth=(wgc.LESS_EQUAL, True))
    d2.instance_count = 2
    spec = scenes.PassSpec(w, h, has_depth=True, clear_depth=1.0, draws=[d1, d2], name="instanced lines")
    both(spec, gpu_device)


def test_blended_lines_with_viewport_scissor_and_instances(gpu_device_ext):
    """The ordered point/line path under a viewport and a scissor rect that
    cut the screen tiles it walks (fragments outside the scissor are
    skipped within a tile's run of steps), with instancing: every pixel's
    blends in API order."""
    w, h = 300, 200
    vb = vertices(800, 53, w, h, spread=1.4)
    vb[:, 7] = 0.45
    d1 = draw(vb, LINE_STRIP, blend=wgc.BLEND_ALPHA)
    d1.instance_count = 2
    d1.viewport = (15.0, 25.0, 260.0, 150.0, 0.0, 1.0)
    d1.scissor = (37, 41, 190, 101)
    d2 = draw(vertices(500, 59, w, h), POINTS, blend=wgc.BLEND_ALPHA)
    d2.scissor = (5, 70, 250, 60)
    spec = scenes.PassSpec(w, h, draws=[d1, d2], semantics=wgc.SEMANTICS_EXTENDED, name="blended lines, scissor")
    both(spec, gpu_device_ext)


def test_strip_restart_format_mismatch_is_an_error(gpu_device):
    announce("restart format mismatch")
    w, h = 64, 64
    vb = vertices(30, 53, w, h)
    d = draw(vb, TRIANGLE_STRIP)
    d.index_data = np.arange(30, dtype=np.uint32).tobytes()
    d.index_format = 4
    d.strip_index_format = 2
    d.count = 30
    spec = scenes.PassSpec(w, h, draws=[d])
    with pytest.raises(bridge.OracleError):
        bridge.render(spec)
    with pytest.raises(wgc.WgcError):
        runner.render(spec, gpu_device)
