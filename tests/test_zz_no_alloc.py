"""The allocation-free submission (wgc_queue_submit_noalloc) through the C ABI.

It must render exactly what wgc_queue_submit renders: two calls in a row,
the second with a smaller pass, and a command buffer of several passes.  A
pass whose scratch exceeds the static array is refused before anything runs
and leaves the device usable.
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


def _render(pp, noalloc):
    dev = pp.device
    if noalloc:
        dev.submit_noalloc([pp.command_buffer])
    else:
        dev.submit([pp.command_buffer])
    dev.poll()
    return pp.results(), dev.last_pass_stats().fragments


def _same(a, b):
    for x, y in zip(a, b):
        assert np.array_equal(x.color, y.color)
        if x.depth is not None:
            assert np.array_equal(x.depth.view(np.uint32), y.depth.view(np.uint32))


def test_noalloc_equals_submit_big_then_small(gpu_device, gpu_device_ext):
    # first call: the 1080p textured pass (extended semantics)
    announce("noalloc: 1080p extended pass")
    big = scenes.config2_textured(extended=True)
    pp_big = runner.PreparedPass(gpu_device_ext, big)
    want, want_frags = _render(pp_big, noalloc=False)
    got, got_frags = _render(pp_big, noalloc=True)
    _same(got, want)
    assert got_frags == want_frags
    ref = bridge.render(big)
    assert ref.fragments == got_frags
    assert np.array_equal(got[0].color, ref.color)
    assert np.array_equal(got[0].depth.view(np.uint32), ref.depth.view(np.uint32))
    # second call, smaller: the 256x256 sphere on the other device object
    # (the same static array on this GPU, left as the first call left it)
    announce("noalloc: 256x256 pass after the 1080p pass")
    small = scenes.config1_plain_small()
    pp_small = runner.PreparedPass(gpu_device, small)
    got, got_frags = _render(pp_small, noalloc=True)
    want, want_frags = _render(pp_small, noalloc=False)
    _same(got, want)
    assert got_frags == want_frags
    ref = bridge.render(small)
    assert np.array_equal(got[0].color, ref.color) and ref.fragments == got_frags


def test_noalloc_several_passes_and_resubmission(gpu_device):
    """A command buffer of independent passes (batched by wgc_queue_submit,
    one launch group per pass here), submitted twice without a wait."""
    announce("noalloc: 8 passes, twice")
    seed = scenes.SEED_BASE + 4
    specs = []
    for i, (w, h, tris, tex) in enumerate(scenes.batched_sizes(seed, 8)):
        specs.append(scenes.batched_pass(seed, i, max(16, w // 2), max(16, h // 2), min(tris, 8000.0), tex))
    pp = runner.PreparedPasses(gpu_device, specs)
    want, want_frags = _render(pp, noalloc=False)
    dev = pp.device
    dev.submit_noalloc([pp.command_buffer])
    dev.submit_noalloc([pp.command_buffer])
    dev.poll()
    got = pp.results()
    _same(got, want)
    assert dev.last_pass_stats().fragments == want_frags


def test_noalloc_ordered_points_and_lines(gpu_device_ext):
    """Blended lines and points (the ordered point/line path, whose set-up
    primitives are laid out in the static array too), submitted twice."""
    announce("noalloc: blended lines and points")
    w, h = 320, 240
    u = scenes.splitmix_f32(9100, 600 * 4).reshape(600, 4)
    vb = np.zeros((1200, 8), np.float32)
    vb[0::2, 0], vb[0::2, 1] = u[:, 0] * 2.4 - 1.2, u[:, 1] * 2.4 - 1.2
    vb[1::2, 0], vb[1::2, 1] = u[:, 2] * 2.4 - 1.2, u[:, 3] * 2.4 - 1.2
    vb[:, 2], vb[:, 3] = 0.5, 1.0
    vb[:, 4:8] = scenes.splitmix_f32(9101, 1200 * 4).reshape(1200, 4)
    vb[:, 7] = 0.5
    draws = []
    for topology, count in ((1, 1200), (0, 300)):   # lines, then points
        draws.append(scenes.DrawSpec(program=wgc.PROGRAM_CLIP_COLOR, vertex_data=vb.tobytes(), vertex_stride=32,
                                     attributes=[(wgc.FLOAT32X4, 0, 0), (wgc.FLOAT32X4, 16, 1)],
                                     blend=wgc.BLEND_ALPHA, count=count, topology=topology))
    spec = scenes.PassSpec(w, h, draws=draws, semantics=wgc.SEMANTICS_EXTENDED, name="noalloc blended lines")
    pp = runner.PreparedPass(gpu_device_ext, spec)
    want, want_frags = _render(pp, noalloc=False)
    for _ in range(2):
        pp.reset_attachments()
        got, got_frags = _render(pp, noalloc=True)
        _same(got, want)
        assert got_frags == want_frags
    ref = bridge.render(spec)
    assert np.array_equal(got[0].color, ref.color) and ref.fragments == got_frags


def test_noalloc_refuses_what_does_not_fit(gpu_device):
    """200k triangles over 2560x2440 need more scratch than the static array
    holds: refused with WGC_ERR_OUT_OF_MEMORY, nothing launched, and the
    device renders the next pass normally."""
    announce("noalloc: oversized pass refused")
    pp = runner.PreparedPass(gpu_device, scenes.config3_overdraw())
    with pytest.raises(wgc.WgcError) as ei:
        gpu_device.submit_noalloc([pp.command_buffer])
    assert ei.value.code == wgc.ERR_OUT_OF_MEMORY
    assert "scratch" in str(ei.value)
    small = scenes.config1_plain_small()
    pp2 = runner.PreparedPass(gpu_device, small)
    got, frags = _render(pp2, noalloc=True)
    ref = bridge.render(small)
    assert np.array_equal(got[0].color, ref.color) and frags == ref.fragments
