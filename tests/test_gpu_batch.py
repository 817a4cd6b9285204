"""The batched form: many independent passes in one submission share their
launches (launched directly, or replayed as a hipGraph when that option is
set).  Every pass must equal the oracle's render of it alone, for every
raster tile shape."""
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


def small_batch(n=20, seed=0x5EED_0004):
    specs = []
    for i, (w, h, tris, tex) in enumerate(scenes.batched_sizes(seed, n)):
        # keep the oracle quick: targets up to 256 px, up to 5k triangles
        w, h, tris = max(16, w // 4), max(16, h // 4), min(tris, 5000.0)
        specs.append(scenes.batched_pass(seed, i, w, h, tris, tex))
    return specs


def check_against_oracle(specs, results, total):
    want_total = 0
    for s, r in zip(specs, results):
        ref = bridge.render(s)
        want_total += ref.fragments
        assert np.array_equal(r.color, ref.color), s.name
        assert np.array_equal(r.depth.view(np.uint32), ref.depth.view(np.uint32)), s.name
    assert total == want_total


@pytest.mark.parametrize("cfg", sorted(wgc.RASTER_TILES))
def test_batch_matches_oracle_per_pass(gpu_device, cfg):
    announce(f"batch raster cfg {cfg}")
    gpu_device.set_option(wgc.OPTION_RASTER_TILE, cfg)
    try:
        specs = small_batch()
        res, total = runner.render_batch(specs, gpu_device)
        check_against_oracle(specs, res, total)
    finally:
        gpu_device.set_option(wgc.OPTION_RASTER_TILE, 0)


def test_graph_replay_and_batching_switch(gpu_device):
    announce("graph replay")
    specs = small_batch(12, seed=99)
    graphs = gpu_device.get_option(wgc.OPTION_GRAPHS)
    gpu_device.set_option(wgc.OPTION_GRAPHS, 1)
    try:
        pp = runner.PreparedPasses(gpu_device, specs)
        pp.submit()
        first = pp.results()
        for _ in range(3):   # the second captures the graph, the others replay it
            pp.submit()
        again = pp.results()
        for a, b in zip(first, again):
            assert np.array_equal(a.color, b.color) and np.array_equal(a.depth, b.depth)
        # the same passes one launch set at a time, no graph
        gpu_device.set_option(wgc.OPTION_BATCHING, 0)
        gpu_device.set_option(wgc.OPTION_GRAPHS, 0)
        pp.submit()
        plain = pp.results()
    finally:
        gpu_device.set_option(wgc.OPTION_BATCHING, 1)
        gpu_device.set_option(wgc.OPTION_GRAPHS, graphs)
    for a, b in zip(first, plain):
        assert np.array_equal(a.color, b.color) and np.array_equal(a.depth, b.depth)


def test_plan_rebuilt_between_unpolled_submissions(gpu_device):
    """Switching the raster tile rebuilds a command buffer's plan while
    submissions that replay its graph are still queued: the old graph and
    draw tables are let go only after those have executed."""
    announce("plan rebuilt between unpolled submissions")
    specs = small_batch(12, seed=5)
    graphs = gpu_device.get_option(wgc.OPTION_GRAPHS)
    gpu_device.set_option(wgc.OPTION_GRAPHS, 1)
    pp = runner.PreparedPasses(gpu_device, specs)
    try:
        pp.submit()
        pp.submit()          # captures the graph
        pp.submit()          # replays it
        gpu_device.set_option(wgc.OPTION_RASTER_TILE, 3)
        pp.submit()          # new plan; the old one is still queued
        gpu_device.set_option(wgc.OPTION_RASTER_TILE, 0)
        pp.submit()
        res = pp.results()
    finally:
        gpu_device.set_option(wgc.OPTION_RASTER_TILE, 0)
        gpu_device.set_option(wgc.OPTION_GRAPHS, graphs)
    check_against_oracle(specs, res, pp.total_fragments())


def test_batched_clipping_of_every_plane(gpu_device):
    """Passes of triangles that cross every clip plane (near and far
    included), batched into one launch group: the clip kernel's lanes clip
    primitives of different draws, and the primitives whose split paths
    have more than three splits go round again on the whole wave."""
    announce("batched clipping, every plane")
    sys.path.insert(0, str(__import__("pathlib").Path(__file__).parent))
    from test_gpu_parity import clip_color_draw, random_clip_triangles
    specs = []
    for i, (w, h) in enumerate(((160, 120), (96, 200), (240, 64), (128, 128))):
        vb = random_clip_triangles(1500, 300 + i, w, h, w_range=(0.3, 3.0), spread=2.0, size_px=(5.0, 300.0))
        vb[:, :, 2] = (vb[:, :, 2] / vb[:, :, 3] * 2.0 - 0.5) * vb[:, :, 3]
        specs.append(scenes.PassSpec(w, h, has_depth=True, clear_depth=1.0,
                                     draws=[clip_color_draw(vb, depth=(wgc.LESS, True))], name=f"clip_batch_{i}"))
    res, total = runner.render_batch(specs, gpu_device)
    want = 0
    for s, r in zip(specs, res):
        ref = bridge.render(s)
        want += ref.fragments
        assert np.array_equal(r.color, ref.color), s.name
        assert np.array_equal(r.depth.view(np.uint32), ref.depth.view(np.uint32)), s.name
    assert total == want


def test_dependent_passes_stay_ordered(gpu_device):
    """Two passes on the same target: the second loads what the first drew,
    so they must not share a launch group."""
    announce("dependent passes")
    a = scenes.config1_plain_small()
    b = scenes.config1_plain_small(seed=scenes.SEED_BASE + 11)
    b.clear_color = None
    b.clear_depth = None
    b.draws[0].depth = (wgc.LESS_EQUAL, True)
    pp = runner.PreparedPasses(gpu_device, [a])
    # record both passes against one pair of attachments
    enc = gpu_device.create_command_encoder()
    cache = {}
    keep = []
    runner.record_pass(gpu_device, enc, a, pp.res[0], keep, cache)
    runner.record_pass(gpu_device, enc, b, pp.res[0], keep, cache)
    cb = enc.finish()
    gpu_device.submit([cb])
    gpu_device.poll()
    color, depth = pp.res[0].read()
    ref_a = bridge.render(a)
    b2 = scenes.config1_plain_small(seed=scenes.SEED_BASE + 11)
    b2.clear_color = None
    b2.clear_depth = None
    b2.draws[0].depth = (wgc.LESS_EQUAL, True)
    b2.color_init = ref_a.color.tobytes()
    b2.depth_init = ref_a.depth.tobytes()
    ref = bridge.render(b2)
    assert np.array_equal(color, ref.color)
    assert np.array_equal(depth.view(np.uint32), ref.depth.view(np.uint32))


@pytest.fixture(scope="module")
def full_batch():
    """The benchmark's batched workload at its full size: 256 passes, W and H
    log-uniform in [64, 1024], spheres of 100-50k triangles, half textured;
    with the oracle's render of every pass alone."""
    seed = scenes.SEED_BASE + 4
    specs = [scenes.batched_pass(seed, i, *a) for i, a in enumerate(scenes.batched_sizes(seed, 256))]
    refs = [bridge.render(s) for s in specs]
    return specs, refs


def _compare_full(specs, refs, res, total):
    bad = []
    for s, r, ref in zip(specs, res, refs):
        if not (np.array_equal(r.color, ref.color) and np.array_equal(r.depth.view(np.uint32),
                                                                      ref.depth.view(np.uint32))):
            bad.append(s.name)
    assert not bad, f"{len(bad)} passes differ from the oracle: {bad[:8]}"
    assert total == sum(ref.fragments for ref in refs)


@pytest.mark.parametrize("cfg", sorted(wgc.RASTER_TILES))
def test_full_batch_matches_oracle(gpu_device, full_batch, cfg):
    announce(f"full 256-pass batch, raster cfg {cfg}")
    specs, refs = full_batch
    gpu_device.set_option(wgc.OPTION_RASTER_TILE, cfg)
    try:
        pp = runner.PreparedPasses(gpu_device, specs)
        pp.submit()
        res = pp.results()
        total = pp.total_fragments()
        # a second submission renders the same images
        pp.submit()
        again = pp.results()
    finally:
        gpu_device.set_option(wgc.OPTION_RASTER_TILE, 0)
    _compare_full(specs, refs, res, total)
    _compare_full(specs, refs, again, total)


def test_full_batch_autotuned(gpu_device, full_batch):
    """The shape the autotuner picks for the full batch renders every pass
    like the oracle."""
    announce("full 256-pass batch, autotuned")
    from wgc import autotune
    specs, refs = full_batch
    pp = runner.PreparedPasses(gpu_device, specs)
    try:
        autotune.tune(pp, reps=2)
        pp.submit()
        res = pp.results()
        _compare_full(specs, refs, res, pp.total_fragments())
    finally:
        gpu_device.set_option(wgc.OPTION_RASTER_TILE, 0)


def test_autotuner_picks_a_shape(gpu_device):
    announce("autotune")
    from wgc import autotune
    specs = small_batch(8, seed=7)
    pp = runner.PreparedPasses(gpu_device, specs)
    timings = autotune.tune(pp, reps=3)
    try:
        assert set(timings) == set(wgc.RASTER_TILES)
        best = gpu_device.get_option(wgc.OPTION_RASTER_TILE)
        assert timings[best] == min(timings.values())
        pp.submit()
        res = pp.results()
        check_against_oracle(specs, res, pp.total_fragments())
    finally:
        gpu_device.set_option(wgc.OPTION_RASTER_TILE, 0)
