"""Closed forms that pin the oracle's fragment stage (fragment.rs:127-214,
427-488): shader-written depth and where the late test sees it, the early
depth test of @early_depth_test (force / allow) with discard, and
@builtin(primitive_index).  Each check compares two oracle renders whose
relation follows from the reference's rules alone, or a render with values
those rules fix, so a mistake in the order of test, shader and write fails
one of them."""
import numpy as np
import pytest

import wgc
from oracle import bridge
from wgc import scenes
from wgc.scenes import (EARLY_ALLOW, EARLY_FORCE, EARLY_NONE, FS_DEPTH_ALPHA, FS_DEPTH_CONST, FS_DEPTH_FIRST,
                        FS_DISCARD, FS_PRIM_RED)


def scene(flags, early, compare=wgc.LESS, write=True, **kw):
    return scenes.fragment_test_scene(flags, early, compare, write, **kw)


def flat_z(spec, z):
    """The same draw with every vertex at depth z (ties everywhere)."""
    vb = np.frombuffer(spec.draws[0].vertex_data, np.float32).copy().reshape(-1, 8)
    vb[:, 2] = z * vb[:, 3]
    spec.draws[0].vertex_data = vb.tobytes()
    return spec


def covered(spec):
    return bridge.render(spec, frag_counts=True).frag_count > 0


def reversed_without_depth(spec):
    """The same triangles in reverse order with no depth attachment: each
    pixel shows the fragment that is first in the original order."""
    vb = np.frombuffer(spec.draws[0].vertex_data, np.float32).reshape(-1, 3, 8)[::-1].copy()
    spec.draws[0].vertex_data = vb.tobytes()
    spec.draws[0].depth = None
    spec.has_depth = False
    return spec


def test_constant_frag_depth_before_colour_is_tested_and_written():
    """frag_depth = 0.3125 as the member before the colour: with Less over a
    cleared buffer every covered pixel ends at exactly 0.3125 whatever the
    interpolated depth, and every later fragment ties and fails, so the
    colour is the first covering fragment's: the image of the same triangles
    drawn in reverse order without a depth test."""
    for z in (0.6, 0.2):   # interpolated depths behind and in front of 0.3125
        a = flat_z(scene(FS_DEPTH_CONST | FS_DEPTH_FIRST, EARLY_NONE), z)
        a.depth_init, a.clear_depth = None, 1.0
        ra = bridge.render(a)
        cov = covered(a)
        assert cov.any()
        assert (ra.depth[cov] == np.float32(0.3125)).all() and (ra.depth[~cov] == 1.0).all()
        rb = bridge.render(reversed_without_depth(flat_z(scene(0, EARLY_NONE), z)))
        assert np.array_equal(ra.color, rb.color)


def test_frag_depth_after_colour_changes_nothing():
    """A frag_depth member after the colour is read after the late test ran:
    the render equals the one without it."""
    for depth in (FS_DEPTH_CONST, FS_DEPTH_ALPHA):
        for compare in (wgc.LESS, wgc.GREATER_EQUAL, wgc.ALWAYS):
            a, b = bridge.render(scene(depth, EARLY_NONE, compare)), bridge.render(scene(0, EARLY_NONE, compare))
            assert np.array_equal(a.color, b.color)
            assert np.array_equal(a.depth.view(np.uint32), b.depth.view(np.uint32))


def test_frag_depth_from_alpha_differs_when_tested():
    """The same shader with its depth member first does change the result
    (the test is not vacuous), and the depths it leaves are alpha values."""
    a = bridge.render(scene(FS_DEPTH_ALPHA | FS_DEPTH_FIRST, EARLY_NONE, wgc.ALWAYS))
    b = bridge.render(scene(0, EARLY_NONE, wgc.ALWAYS))
    assert not np.array_equal(a.depth, b.depth)
    spec = scene(0, EARLY_NONE, wgc.ALWAYS)
    vb = np.frombuffer(spec.draws[0].vertex_data, np.float32).reshape(-1, 8)
    alphas = set(np.unique(vb[:, 7]).tolist())
    init = set(np.unique(np.frombuffer(spec.depth_init, np.float32)).tolist())
    # every vertex of a triangle shares its alpha, so the interpolated alpha
    # is that value up to the rounding of the barycentric sum (a few ulp of
    # the barycentrics, which cancel near the edges)
    got = np.unique(a.depth)
    assert all(v in init or min(abs(v - x) for x in alphas) < 1e-4 for v in got.tolist())


@pytest.mark.parametrize("compare", [wgc.LESS, wgc.LESS_EQUAL, wgc.GREATER, wgc.ALWAYS])
def test_forced_early_test_with_every_fragment_discarded(compare):
    """@early_depth_test(force) and a shader that discards every fragment:
    the colour is untouched and the depth is what the same draw without
    discard leaves (the early test wrote it before the shader ran)."""
    a = scene(FS_DISCARD, EARLY_FORCE, compare, alpha=(0.1, 0.45))
    b = scene(0, EARLY_NONE, compare, alpha=(0.1, 0.45))
    ra, rb = bridge.render(a), bridge.render(b)
    clear = bridge.render(scenes.fragment_test_scene(0, EARLY_NONE, compare, True, n=0))
    assert np.array_equal(ra.color, clear.color)
    assert np.array_equal(ra.depth.view(np.uint32), rb.depth.view(np.uint32))
    assert not np.array_equal(rb.depth, clear.depth)


def test_discard_without_early_test_writes_nothing():
    a = bridge.render(scene(FS_DISCARD, EARLY_NONE, alpha=(0.1, 0.45)))
    clear = bridge.render(scenes.fragment_test_scene(0, EARLY_NONE, wgc.LESS, True, n=0))
    assert np.array_equal(a.color, clear.color) and np.array_equal(a.depth, clear.depth)


@pytest.mark.parametrize("compare", [wgc.LESS, wgc.LESS_EQUAL, wgc.GREATER_EQUAL, wgc.ALWAYS])
def test_forced_early_test_without_discard_is_the_late_test(compare):
    """Without discard or shader depth, testing before or after the shader
    is the same thing."""
    a, b = bridge.render(scene(0, EARLY_FORCE, compare)), bridge.render(scene(0, EARLY_NONE, compare))
    assert np.array_equal(a.color, b.color) and np.array_equal(a.depth, b.depth)


def test_allowed_early_test_with_strict_compare_never_writes_colour():
    """Allow runs the test twice: with Less + write the early test stores
    the fragment's depth, so the late test (z < z) fails and no colour is
    ever written; the depth is the single-test result.  With LessEqual the
    second test passes and the render equals the single-test one."""
    a, b = bridge.render(scene(0, EARLY_ALLOW, wgc.LESS)), bridge.render(scene(0, EARLY_NONE, wgc.LESS))
    clear = bridge.render(scenes.fragment_test_scene(0, EARLY_NONE, wgc.LESS, True, n=0))
    assert np.array_equal(a.color, clear.color)
    assert np.array_equal(a.depth, b.depth)
    a, b = bridge.render(scene(0, EARLY_ALLOW, wgc.LESS_EQUAL)), bridge.render(scene(0, EARLY_NONE, wgc.LESS_EQUAL))
    assert np.array_equal(a.color, b.color) and np.array_equal(a.depth, b.depth)
    # without depth writes the stored depth never changes: one test or two agree
    a, b = bridge.render(scene(0, EARLY_ALLOW, wgc.LESS, False)), bridge.render(scene(0, EARLY_NONE, wgc.LESS, False))
    assert np.array_equal(a.color, b.color)


def _grid_triangles(nx, ny, w, h, flip_every_other):
    """One small triangle in the middle of each cell of an nx x ny grid,
    counter-clockwise, every other one clockwise when asked."""
    tris = []
    for k in range(nx * ny):
        cx = ((k % nx) + 0.5) / nx * 2 - 1
        cy = ((k // nx) + 0.5) / ny * 2 - 1
        r = 0.6 / max(nx, ny)
        v = [(cx - r, cy - r), (cx + r, cy - r), (cx, cy + r)]
        if flip_every_other and k % 2:
            v = [v[0], v[2], v[1]]
        tris.append(v)
    vb = np.zeros((len(tris), 3, 8), np.float32)
    for k, t in enumerate(tris):
        for j, (x, y) in enumerate(t):
            vb[k, j] = (x, y, 0.5, 1.0, 0.0, 0.5, 0.25, 1.0)
    return vb


@pytest.mark.parametrize("instances", [1, 3])
def test_primitive_index_counts_culled_primitives(instances):
    """Red = (primitive_index % 8) * 0.125: the triangle of grid cell k shows
    trunc(255 * 0.125 * (k % 8)); back faces (every other cell) are culled
    and still counted; every instance numbers its primitives from 0."""
    nx, ny, w, h = 8, 4, 128, 64
    vb = _grid_triangles(nx, ny, w, h, flip_every_other=True)
    spec = scenes.fragment_test_scene(FS_PRIM_RED, EARLY_NONE, None, n=1, width=w, height=h)
    d = spec.draws[0]
    d.vertex_data = vb.tobytes()
    d.count = vb.shape[0] * 3
    d.cull_mode = wgc.CULL_BACK
    d.instance_count = instances
    d.interp = wgc.INTERP_FLAT          # the other channels exact: the first vertex's values
    d.wgsl = scenes.fragment_test_wgsl(FS_PRIM_RED, EARLY_NONE, wgc.INTERP_FLAT)
    r = bridge.render(spec)
    for k in range(nx * ny):
        cx, cy = int(((k % nx) + 0.5) * w / nx), int(h - ((k // nx) + 0.5) * h / ny)
        px = r.color[cy, cx]
        if k % 2:
            assert tuple(px) == (25, 51, 76, 255)      # culled: the clear colour
        else:
            assert px[0] == int(np.float32(255.0) * np.float32((k % 8) * 0.125)), (k, px)
            assert tuple(px[1:]) == (127, 63, 255)
