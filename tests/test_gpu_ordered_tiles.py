"""The ordered raster (blending, and overwrite without depth, applied in API
order) for every raster tile shape: single draws, a crowded tile whose keys
come from the spill list, and ordered passes in the batched form next to
order-independent ones.  Every pass must equal the oracle's render."""
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


@pytest.fixture
def tile_cfg(gpu_device, gpu_device_ext, request):
    cfg = request.param
    for dev in (gpu_device, gpu_device_ext):
        dev.set_option(wgc.OPTION_RASTER_TILE, cfg)
    yield cfg
    for dev in (gpu_device, gpu_device_ext):
        dev.set_option(wgc.OPTION_RASTER_TILE, 0)


def triangles(n, seed, width, height, spread=1.0, size_px=(4.0, 60.0), alpha=(0.2, 0.8)):
    u = scenes.splitmix_f32(seed, n * 8).astype(np.float64).reshape(n, 8)
    cx, cy = (u[:, 0] * 2 - 1) * spread, (u[:, 1] * 2 - 1) * spread
    size = size_px[0] + u[:, 2] * (size_px[1] - size_px[0])
    ang = u[:, 3:6] * 2 * np.pi
    vb = np.zeros((n, 3, 8), np.float32)
    vb[:, :, 0] = cx[:, None] + np.cos(ang) * size[:, None] * 2.0 / width
    vb[:, :, 1] = cy[:, None] + np.sin(ang) * size[:, None] * 2.0 / height
    vb[:, :, 2] = 0.5
    vb[:, :, 3] = 1.0
    vb[:, :, 4:7] = u[:, None, 6:7] * np.array([1.0, 0.5, 0.25])
    vb[:, :, 7] = (alpha[0] + u[:, 7] * (alpha[1] - alpha[0]))[:, None]
    return vb


def blend_pass(vb, width, height, name, blend=wgc.BLEND_ALPHA):
    d = scenes.DrawSpec(program=wgc.PROGRAM_CLIP_COLOR, vertex_data=vb.tobytes(), vertex_stride=32,
                        attributes=[(wgc.FLOAT32X4, 0, 0), (wgc.FLOAT32X4, 16, 1)], count=vb.shape[0] * 3,
                        blend=blend)
    return scenes.PassSpec(width, height, clear_color=(0.1, 0.2, 0.3, 1.0), draws=[d],
                           semantics=wgc.SEMANTICS_EXTENDED, name=name)


def assert_same(spec, got, ref, per_pass_count=True):
    if per_pass_count:
        assert got.fragments == ref.fragments, spec.name
    if not np.array_equal(got.color, ref.color):
        diff = np.argwhere((got.color != ref.color).any(-1))
        y, x = diff[0]
        raise AssertionError(f"{spec.name}: colour differs at {len(diff)} px, first ({x},{y}): "
                             f"gpu {got.color[y, x]} oracle {ref.color[y, x]}")


TILES = sorted(wgc.RASTER_TILES)


@pytest.mark.parametrize("tile_cfg", TILES, indirect=True)
def test_overdraw_blend_per_tile_shape(gpu_device_ext, tile_cfg):
    spec = scenes.config3_overdraw(n=20_000, width=640, height=610, blend=True)
    announce(f"overdraw blend, raster cfg {tile_cfg}")
    assert_same(spec, runner.render(spec, gpu_device_ext), bridge.render(spec))


@pytest.mark.parametrize("tile_cfg", TILES, indirect=True)
def test_overdraw_in_api_order_per_tile_shape(gpu_device, tile_cfg):
    spec = scenes.config3_overdraw(n=20_000, width=640, height=610)
    announce(f"overdraw in API order, raster cfg {tile_cfg}")
    assert_same(spec, runner.render(spec, gpu_device), bridge.render(spec))


@pytest.mark.parametrize("tile_cfg", TILES, indirect=True)
def test_crowded_ordered_tile_per_tile_shape(gpu_device_ext, tile_cfg):
    """2500 small triangles on the four central 16x16 tiles of a 1024x1024
    target: past every bin's capacity, so most keys come from the spill
    list, and under the 4096 keys an ordered tile sorts at once."""
    spec = blend_pass(triangles(2500, 77, 1024, 1024, spread=0.006, size_px=(1.5, 5.0)), 1024, 1024, "crowded")
    announce(f"crowded ordered tile, raster cfg {tile_cfg}")
    assert_same(spec, runner.render(spec, gpu_device_ext), bridge.render(spec))


@pytest.mark.parametrize("tile_cfg", TILES, indirect=True)
def test_ordered_passes_batched_per_tile_shape(gpu_device_ext, tile_cfg):
    """Blended passes of several sizes (ragged edge tiles) between
    order-independent depth-tested ones, in one submission."""
    announce(f"ordered passes batched, raster cfg {tile_cfg}")
    specs = []
    for i in range(6):
        w, h = 100 + 37 * i, 60 + 23 * i
        if i % 2:
            specs.append(blend_pass(triangles(800 + 300 * i, 300 + i, w, h), w, h, f"blend_{i}"))
        else:
            d = scenes.DrawSpec(program=wgc.PROGRAM_CLIP_COLOR,
                                vertex_data=triangles(600, 400 + i, w, h, alpha=(1.0, 1.0)).tobytes(),
                                vertex_stride=32, attributes=[(wgc.FLOAT32X4, 0, 0), (wgc.FLOAT32X4, 16, 1)],
                                count=1800, depth=(wgc.LESS_EQUAL, True))
            specs.append(scenes.PassSpec(w, h, has_depth=True, clear_depth=1.0, draws=[d],
                                         semantics=wgc.SEMANTICS_EXTENDED, name=f"depth_{i}"))
    res, total = runner.render_batch(specs, gpu_device_ext)
    want = 0
    for s, r in zip(specs, res):
        ref = bridge.render(s)
        want += ref.fragments
        assert_same(s, r, ref, per_pass_count=False)   # the batch reports one total
        if ref.depth is not None:
            assert np.array_equal(r.depth.view(np.uint32), ref.depth.view(np.uint32)), s.name
    assert total == want


def one_tile(vb, cx=136, cy=136, size=256):
    """Moves triangles generated around the centre of a size x size target
    to pixel (cx, cy), well inside one 16x16 tile."""
    vb = vb.copy()
    vb[:, :, 0] += (cx + 0.5) / size * 2 - 1
    vb[:, :, 1] -= (cy + 0.5) / size * 2 - 1
    return vb


@pytest.mark.parametrize("n", [4096, 4097, 9000])
def test_ordered_tile_beyond_one_sort(gpu_device_ext, n):
    """More blended triangles on one tile than one sort holds: the tile
    takes them in chunks of 4096 in API order (n = 4096 is one full chunk,
    4097 leaves one key for a second)."""
    vb = one_tile(triangles(n, 500 + n, 256, 256, spread=0.01, size_px=(1.5, 3.0)))
    spec = blend_pass(vb, 256, 256, f"one tile, {n} blended triangles")
    announce(spec.name)
    assert_same(spec, runner.render(spec, gpu_device_ext), bridge.render(spec))


def test_not_equal_write_beyond_one_sort(gpu_device_ext):
    """NotEqual + depth write (applied in API order) with 7000 triangles on
    one tile over four depth levels."""
    vb = one_tile(triangles(7000, 611, 256, 256, spread=0.01, size_px=(1.5, 3.0), alpha=(1.0, 1.0)))
    u = scenes.splitmix_f32(612, vb.shape[0])
    vb[:, :, 2] = (np.array([0.2, 0.4, 0.6, 0.8], np.float32)[(u * 4).astype(int) % 4])[:, None]
    d = scenes.DrawSpec(program=wgc.PROGRAM_CLIP_COLOR, vertex_data=vb.tobytes(), vertex_stride=32,
                        attributes=[(wgc.FLOAT32X4, 0, 0), (wgc.FLOAT32X4, 16, 1)], count=vb.shape[0] * 3,
                        depth=(wgc.NOT_EQUAL, True))
    spec = scenes.PassSpec(256, 256, has_depth=True, clear_depth=0.5, draws=[d], semantics=wgc.SEMANTICS_EXTENDED,
                           name="not_equal_write_one_tile")
    announce(spec.name)
    ref = bridge.render(spec)
    got = runner.render(spec, gpu_device_ext)
    assert_same(spec, got, ref)
    assert np.array_equal(got.depth.view(np.uint32), ref.depth.view(np.uint32))


def test_ordered_tile_beyond_one_sort_batched(gpu_device_ext):
    """The crowded blended pass between two ordinary ones in one submission
    (its spill entries are global tiles of the group)."""
    announce("one tile beyond one sort, batched")
    specs = [blend_pass(triangles(500, 700, 96, 80), 96, 80, "small_a"),
             blend_pass(one_tile(triangles(6000, 701, 256, 256, spread=0.01, size_px=(1.5, 3.0))), 256, 256, "crowded"),
             blend_pass(triangles(700, 702, 128, 64), 128, 64, "small_b")]
    res, total = runner.render_batch(specs, gpu_device_ext)
    want = 0
    for s, r in zip(specs, res):
        ref = bridge.render(s)
        want += ref.fragments
        assert_same(s, r, ref, per_pass_count=False)
    assert total == want
