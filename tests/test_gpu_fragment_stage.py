"""The fragment stage's ordering rules on the GPU against the oracle
(fragment.rs:127-214, 427-488; state.rs:536-541): WGSL fragment shaders
that discard, write @builtin(frag_depth) before or after their colour, read
@builtin(primitive_index), under @early_depth_test(force) or a conservative
(allowed) early test, for several depth compare functions, over a random
initial depth buffer with many depth ties."""
import sys

import numpy as np
import pytest

import wgc
from oracle import bridge
from wgc import runner, scenes
from wgc.scenes import (EARLY_ALLOW, EARLY_FORCE, EARLY_NONE, FS_DEPTH_ALPHA, FS_DEPTH_CONST, FS_DEPTH_FIRST,
                        FS_DISCARD, FS_PRIM_RED)

pytestmark = pytest.mark.gpu


def announce(name):
    print(f"\n[gpu case] {name}", flush=True)
    sys.stdout.flush()


def check(spec, dev_ref, dev_ext):
    announce(spec.name)
    dev = dev_ext if spec.semantics == wgc.SEMANTICS_EXTENDED else dev_ref
    ref = bridge.render(spec, frag_counts=True)
    got = runner.render(spec, dev, fragment_counts=True)
    assert got.fragments == ref.fragments
    assert np.array_equal(got.frag_count, ref.frag_count), "coverage"
    if ref.depth is not None:
        gd, rd = got.depth.view(np.uint32), ref.depth.view(np.uint32)
        if not np.array_equal(gd, rd):
            y, x = np.argwhere(gd != rd)[0]
            raise AssertionError(f"depth differs at {int((gd != rd).sum())} px, first ({x},{y}): "
                                 f"gpu {got.depth[y, x]!r} oracle {ref.depth[y, x]!r}")
    if not np.array_equal(got.color, ref.color):
        diff = np.argwhere((got.color != ref.color).any(-1))
        y, x = diff[0]
        raise AssertionError(f"colour differs at {len(diff)} px, first ({x},{y}): gpu {got.color[y, x]} "
                             f"oracle {ref.color[y, x]}")
    return ref


FLAGS = [FS_DISCARD, FS_DEPTH_ALPHA | FS_DEPTH_FIRST, FS_DEPTH_CONST | FS_DEPTH_FIRST, FS_DEPTH_ALPHA, FS_PRIM_RED,
         FS_DISCARD | FS_DEPTH_ALPHA | FS_DEPTH_FIRST | FS_PRIM_RED]
DEPTH = [(wgc.LESS, True), (wgc.GREATER_EQUAL, True), (wgc.ALWAYS, True), (wgc.LESS, False)]


@pytest.mark.parametrize("early", [EARLY_NONE, EARLY_FORCE, EARLY_ALLOW])
@pytest.mark.parametrize("flags", FLAGS)
@pytest.mark.parametrize("depth", DEPTH)
def test_fragment_stage_triangles(gpu_device, gpu_device_ext, flags, early, depth):
    spec = scenes.fragment_test_scene(flags, early, depth[0], depth[1])
    check(spec, gpu_device, gpu_device_ext)


@pytest.mark.parametrize("topology", [0, 1, 2])
@pytest.mark.parametrize("flags,early", [(FS_DISCARD, EARLY_FORCE), (FS_DEPTH_ALPHA | FS_DEPTH_FIRST, EARLY_NONE),
                                         (FS_PRIM_RED, EARLY_ALLOW)])
def test_fragment_stage_points_and_lines(gpu_device, gpu_device_ext, topology, flags, early):
    spec = scenes.fragment_test_scene(flags, early, wgc.LESS_EQUAL, True, n=400, topology=topology)
    check(spec, gpu_device, gpu_device_ext)


@pytest.mark.parametrize("flags,early", [(FS_DISCARD, EARLY_FORCE), (FS_DEPTH_CONST | FS_DEPTH_FIRST, EARLY_ALLOW)])
def test_fragment_stage_with_blending(gpu_device, gpu_device_ext, flags, early):
    spec = scenes.fragment_test_scene(flags, early, wgc.LESS_EQUAL, True, blend=True)
    check(spec, gpu_device, gpu_device_ext)


@pytest.mark.parametrize("instances", [1, 3])
def test_primitive_index_with_culling_and_instances(gpu_device, gpu_device_ext, instances):
    """Back faces culled but counted, numbering from 0 in every instance."""
    spec = scenes.fragment_test_scene(FS_PRIM_RED, EARLY_NONE, wgc.LESS_EQUAL, True, n=900, cull_mode=wgc.CULL_BACK,
                                      instances=instances)
    ref = check(spec, gpu_device, gpu_device_ext)
    reds = set(np.unique(ref.color[..., 0]).tolist())
    assert len(reds) >= 8   # all eight residues appear (and the clear colour)


def test_fragment_stage_batched_with_plain_passes(gpu_device):
    """Passes of these pipelines next to ordinary passes in one submission
    (the batched form groups ordered and order-independent passes apart)."""
    announce("fragment stage batched")
    specs = []
    for i, (flags, early) in enumerate([(FS_DISCARD, EARLY_FORCE), (0, EARLY_NONE), (FS_DEPTH_ALPHA | FS_DEPTH_FIRST,
                                                                                      EARLY_ALLOW), (FS_PRIM_RED, 0)]):
        specs.append(scenes.fragment_test_scene(flags, early, wgc.LESS, True, seed=scenes.SEED_BASE + 40 + i))
    res, total = runner.render_batch(specs, gpu_device)
    want = 0
    for s, r in zip(specs, res):
        ref = bridge.render(s)
        want += ref.fragments
        assert np.array_equal(r.color, ref.color), s.name
        assert np.array_equal(r.depth.view(np.uint32), ref.depth.view(np.uint32)), s.name
    assert total == want
