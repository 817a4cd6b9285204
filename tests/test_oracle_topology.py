"""The oracle's point, line and strip paths against the reference's own
known-answer tests and closed forms (CPU only).

KATs: util/bresenham.rs:160-220 (four lines), clipper.rs:420-564 (the
Cohen-Sutherland cases), primitive.rs:501-509 (the WebGPU strip example).
Closed forms: strip restart segments, line windows, pixel sets of axis-aligned
lines and of points, and invariants of Bresenham on all short lines.
"""
import itertools

import numpy as np
import pytest

from oracle import bridge
from wgc import scenes

TOPO_POINTS, TOPO_LINES, TOPO_LINE_STRIP, TOPO_TRIANGLES, TOPO_TRIANGLE_STRIP = range(5)


# ---------------------------------------------------------------- bresenham.rs:160-220

@pytest.mark.parametrize("start,end,want", [
    ((0, 1), (6, 4), [(0, 1), (1, 1), (2, 2), (3, 2), (4, 3), (5, 3)]),      # test_wp_example
    ((6, 4), (0, 1), [(6, 4), (5, 4), (4, 3), (3, 3), (2, 2), (1, 2)]),      # test_inverse_wp
    ((2, 3), (5, 3), [(2, 3), (3, 3), (4, 3)]),                              # test_straight_hline
    ((2, 3), (2, 6), [(2, 3), (2, 4), (2, 5)]),                              # test_straight_vline
])
def test_bresenham_reference_kats(start, end, want):
    pts, _ = bridge.bresenham(start, end)
    assert pts == want


def test_bresenham_invariants_on_all_short_lines():
    """Every line between points of a 9x9 grid: as many points as the major
    delta, starting at `start`, 8-connected, one step along the major axis
    each, less than a pixel from the ideal line (the crate's integer variant
    starts its error term at dy - dx, so it truncates rather than rounds), and
    t = the running binary32 sum of 1/major (bresenham.rs:129-150)."""
    grid = list(itertools.product(range(9), range(9)))
    for s in grid:
        for e in grid:
            pts, t = bridge.bresenham(s, e)
            dx, dy = e[0] - s[0], e[1] - s[1]
            major = max(abs(dx), abs(dy))
            assert len(pts) == major, (s, e)
            if not major:
                continue
            assert pts[0] == s
            for (x0, y0), (x1, y1) in zip(pts, pts[1:]):
                assert max(abs(x1 - x0), abs(y1 - y0)) == 1
            for k, (x, y) in enumerate(pts):
                if abs(dx) >= abs(dy):
                    assert x == s[0] + k * np.sign(dx)
                    ideal = s[1] + dy * (x - s[0]) / dx
                    assert abs(y - ideal) < 1.0
                else:
                    assert y == s[1] + k * np.sign(dy)
                    ideal = s[0] + dx * (y - s[1]) / dy
                    assert abs(x - ideal) < 1.0
            dt = np.float32(1.0) / np.float32(major)
            acc = np.float32(0.0)
            for k in range(major):
                assert t[k] == acc
                acc = np.float32(acc + dt)


# ---------------------------------------------------------------- clipper.rs:420-564

def close(a, b):
    return np.allclose(np.asarray(a, np.float32), np.asarray(b, np.float32), rtol=0, atol=1.2e-7)


def test_line_trivially_inside():
    a, b = [-0.5, -0.5, 0.5, 1.0], [0.5, 0.5, 0.5, 1.0]
    (p0, p1), al = bridge.clip_line(a, b)
    assert close(p0, a) and close(p1, b) and al == (0.0, 1.0)


@pytest.mark.parametrize("a,b", [
    ([-1.5, -1.5, 0.5, 1.0], [-1.5, 1.5, 0.5, 1.0]), ([1.5, -1.5, 0.5, 1.0], [1.5, 1.5, 0.5, 1.0]),
    ([-1.5, -1.5, 0.5, 1.0], [1.5, -1.5, 0.5, 1.0]), ([-1.5, 1.5, 0.5, 1.0], [1.5, 1.5, 0.5, 1.0]),
    # line_non_trivially_outside
    ([-10.0, 0.0, 0.5, 1.0], [0.0, -10.0, 0.5, 1.0]), ([10.0, 0.0, 0.5, 1.0], [0.0, -10.0, 0.5, 1.0]),
    ([-10.0, 0.0, 0.5, 1.0], [0.0, 10.0, 0.5, 1.0]), ([10.0, 0.0, 0.5, 1.0], [0.0, 10.0, 0.5, 1.0]),
])
def test_line_outside(a, b):
    assert bridge.clip_line(a, b) is None


@pytest.mark.parametrize("a,c", [
    ([-1.5, -0.5, 0.5, 1.0], [-1.0, -0.25, 0.5, 1.0]),   # left
    ([1.5, -0.5, 0.5, 1.0], [1.0, 0.0, 0.5, 1.0]),       # right
    ([-0.5, -1.5, 0.5, 1.0], [-0.25, -1.0, 0.5, 1.0]),   # bottom
    ([-0.5, 1.5, 0.5, 1.0], [0.0, 1.0, 0.5, 1.0]),       # top
    ([0.5, -0.5, 1.5, 1.0], [0.5, 0.0, 1.0, 1.0]),       # back
    ([0.5, -0.5, -0.5, 1.0], [0.5, 0.0, 0.0, 1.0]),      # front
])
def test_clip_line_against_each_plane(a, c):
    b = [0.5, 0.5, 0.5, 1.0]
    (p0, p1), _ = bridge.clip_line(a, b)
    assert close(p0, c) and close(p1, b)
    (p0, p1), _ = bridge.clip_line(b, a)
    assert close(p0, b) and close(p1, c)


def test_clip_line_second_update_reads_the_updated_first_point():
    """clipper.rs:343-356 lerps the second end point between the already
    clipped first point and the second point with the second alpha.  With both
    ends outside (left and right) the far end is therefore not on the plane."""
    a, b = [-3.0, 0.0, 0.5, 1.0], [3.0, 0.0, 0.5, 1.0]
    (p0, p1), (a0, a1) = bridge.clip_line(a, b)
    assert close([a0, a1], [1.0 / 3.0, 2.0 / 3.0])
    assert close(p0, [-1.0, 0.0, 0.5, 1.0])
    # the reference's second update: lerp(p0_updated, b, 2/3) = -1/3 + 2 = 5/3
    assert abs(p1[0] - 5.0 / 3.0) < 1e-6


# ---------------------------------------------------------------- primitive.rs

def test_triangle_strip_reference_kat():
    """primitive.rs:501-509: the example from the WebGPU specification."""
    assert bridge.assemble(TOPO_TRIANGLE_STRIP, 6) == [[0, 1, 2], [2, 1, 3], [2, 3, 4], [4, 3, 5]]


def test_list_and_line_strip_assembly():
    assert bridge.assemble(TOPO_POINTS, 3) == [[0], [1], [2]]
    assert bridge.assemble(TOPO_LINES, 5) == [[0, 1], [2, 3]]              # remainder dropped
    assert bridge.assemble(TOPO_LINE_STRIP, 4) == [[0, 1], [1, 2], [2, 3]]
    assert bridge.assemble(TOPO_TRIANGLES, 7) == [[0, 1, 2], [3, 4, 5]]
    assert bridge.assemble(TOPO_TRIANGLE_STRIP, 2) == []


@pytest.mark.parametrize("fmt", [2, 4])
def test_strip_restart(fmt):
    sep = 0xFFFF if fmt == 2 else 0xFFFFFFFF
    idx = [0, 1, 2, 3, sep, 4, 5, sep, sep, 6, 7, 8, sep, 9]
    tris = bridge.assemble(TOPO_TRIANGLE_STRIP, len(idx), idx, fmt, restart=True)
    # segments [0..4), [5..7), [9..12), [13]; winding restarts with each segment
    assert tris == [[0, 1, 2], [2, 1, 3], [9, 10, 11]]
    lines = bridge.assemble(TOPO_LINE_STRIP, len(idx), idx, fmt, restart=True)
    assert lines == [[0, 1], [1, 2], [2, 3], [5, 6], [9, 10], [10, 11]]
    # without restart the maximum is an ordinary index
    assert len(bridge.assemble(TOPO_TRIANGLE_STRIP, len(idx), idx, fmt)) == len(idx) - 2
    with pytest.raises(bridge.OracleError):
        bridge.assemble(TOPO_TRIANGLES, len(idx), idx, fmt, restart=True)


# ---------------------------------------------------------------- rendered closed forms

def clip_draw(vb, topology, **kw):
    from wgc import FLOAT32X4, PROGRAM_CLIP_COLOR
    return scenes.DrawSpec(program=PROGRAM_CLIP_COLOR, vertex_data=np.asarray(vb, np.float32).tobytes(),
                           vertex_stride=32, attributes=[(FLOAT32X4, 0, 0), (FLOAT32X4, 16, 1)],
                           count=len(vb), topology=topology, **kw)


def test_horizontal_line_covers_its_row_end_exclusive():
    """16x16 target: a line between pixel corners (2, 5) and (11, 5) covers
    x = 2..10 of row 5 (the end is exclusive) with colour lerped along it."""
    w = h = 16
    def clip(px, py):
        return [px / w * 2 - 1, 1 - py / h * 2, 0.5, 1.0]
    vb = [clip(2, 5) + [1, 0, 0, 1], clip(11, 5) + [0, 0, 1, 1]]
    spec = scenes.PassSpec(w, h, draws=[clip_draw(vb, TOPO_LINES)])
    r = bridge.render(spec, frag_counts=True)
    covered = np.argwhere(r.frag_count > 0)
    assert sorted((int(x), int(y)) for y, x in covered) == [(x, 5) for x in range(2, 11)]
    red = r.color[5, 2:11, 0].astype(int)
    assert red[0] == 255 and (np.diff(red) <= 0).all()


def test_points_land_on_their_pixels():
    w = h = 8
    vb = [[(x + 0.5) / w * 2 - 1, 1 - (y + 0.5) / h * 2, 0.5, 1.0, 0, 1, 0, 1] for x, y in [(1, 2), (6, 7), (0, 0)]]
    spec = scenes.PassSpec(w, h, draws=[clip_draw(vb, TOPO_POINTS)])
    r = bridge.render(spec, frag_counts=True)
    assert sorted((int(x), int(y)) for y, x in np.argwhere(r.frag_count > 0)) == [(0, 0), (1, 2), (6, 7)]
    assert (r.color[2, 1] == [0, 255, 0, 255]).all()


def test_strip_equals_its_triangle_list():
    """A triangle strip renders like the list of its assembled triangles."""
    rng = np.random.default_rng(3)
    v = rng.uniform(-1, 1, (12, 8)).astype(np.float32)
    v[:, 2] = rng.uniform(0.1, 0.9, 12)
    v[:, 3] = 1.0
    tris = bridge.assemble(TOPO_TRIANGLE_STRIP, 12)
    lst = np.concatenate([v[t] for t in tris])
    a = bridge.render(scenes.PassSpec(64, 64, draws=[clip_draw(v, TOPO_TRIANGLE_STRIP)]))
    b = bridge.render(scenes.PassSpec(64, 64, draws=[clip_draw(lst, TOPO_TRIANGLES)]))
    assert np.array_equal(a.color, b.color)
