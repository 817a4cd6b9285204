"""Pins the oracle's coverage and interpolation where the mathematics fixes
the answer: an independent per-pixel restatement of the scanline rule,
brute-forced over every small integer triangle, and closed-form scenes.
"""
import itertools
import struct

import numpy as np

from oracle import bridge
from wgc import scenes
import wgc


def _trunc_div(a, b):
    """Rust isize `/`: truncates toward zero."""
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b >= 0) else -q


def closed_form_pixels(tri):
    """Per-row span rule written from scanline.rs:14-198 as a closed form:
    sort by y (compare-and-swap (0,1),(0,2),(1,2)); a single row when all y
    are equal; rows [a.y, b.y) from edges AC/AB; rows [b.y, c.y] from AC/BC
    unless BC is horizontal (then row c.y is not emitted)."""
    v = [list(tri[0:2]), list(tri[2:4]), list(tri[4:6])]
    for l, r in ((0, 1), (0, 2), (1, 2)):
        if v[l][1] > v[r][1]:
            v[l], v[r] = v[r], v[l]
    (ax, ay), (bx, by), (cx, cy) = v
    out = set()
    if cy == ay:
        for x in range(min(ax, bx, cx), max(ax, bx, cx) + 1):
            out.add((x, ay))
        return out
    for y in range(ay, cy + 1):
        i = y - ay
        x1 = ax + _trunc_div((cx - ax) * i, cy - ay)
        if by == ay or y >= by:
            if cy == by:
                continue
            x2 = bx + _trunc_div((cx - bx) * (y - by), cy - by)
        else:
            x2 = ax + _trunc_div((bx - ax) * i, by - ay)
        for x in range(min(x1, x2), max(x1, x2) + 1):
            out.add((x, y))
    return out


def test_scanline_matches_closed_form_exhaustive_small():
    pts = [(x, y) for x in range(4) for y in range(4)]
    for a, b, c in itertools.product(pts, repeat=3):
        tri = [*a, *b, *c]
        got = bridge.scanline_pixels(tri)
        assert len(got) == len(set(got)), tri  # each pixel at most once per triangle
        assert set(got) == closed_form_pixels(tri), tri


def test_scanline_matches_closed_form_random():
    rng = np.random.default_rng(3)
    for _ in range(3000):
        tri = [int(v) for v in rng.integers(0, 64, size=6)]
        assert set(bridge.scanline_pixels(tri)) == closed_form_pixels(tri), tri


def test_axis_aligned_right_triangle_staircase():
    # (0,0) (8,0) (0,8): rows 0..7 from AC/AB, BC horizontal? no: sorted a=(0,0)
    # b=(8,0) c=(0,8) -> start in the second half; x1 = 0, x2 = 8 - y
    got = set(bridge.scanline_pixels([0, 0, 8, 0, 0, 8]))
    expect = {(x, y) for y in range(9) for x in range(0, 8 - y + 1)}
    assert got == expect


def _quad_spec(w, h):
    verts = np.array([[-1, -1, 0.5, 1], [1, -1, 0.5, 1], [1, 1, 0.5, 1],
                      [-1, -1, 0.5, 1], [1, 1, 0.5, 1], [-1, 1, 0.5, 1]], np.float32)
    col = np.tile(np.array([[0.25, 0.5, 0.75, 1.0]], np.float32), (6, 1))
    vb = np.concatenate([verts, col], axis=1)
    d = scenes.DrawSpec(program=wgc.PROGRAM_CLIP_COLOR, vertex_data=vb.tobytes(), vertex_stride=32,
                        attributes=[(wgc.FLOAT32X4, 0, 0), (wgc.FLOAT32X4, 16, 1)], count=6)
    return scenes.PassSpec(w, h, clear_color=(0, 0, 0, 0), draws=[d])


def test_full_viewport_quad_covers_every_pixel():
    w, h = 37, 23
    r = bridge.render(_quad_spec(w, h), frag_counts=True)
    fc = r.frag_count
    assert (fc >= 1).all()
    # the shared diagonal is emitted by both triangles
    assert (fc == 2).sum() > 0
    assert r.fragments == int(fc.sum())
    # interpolating a constant with corner barycentrics moves it by ulps:
    # 0.25, 0.5, 0.75 stay clear of a unorm8 step, 1.0 may truncate to 254
    assert (r.color[..., :3] == np.array([63, 127, 191], np.uint8)).all()
    assert np.isin(r.color[..., 3], [254, 255]).all()


def test_depth_is_interpolated_ndc_z():
    # a triangle with constant z: every covered pixel gets exactly that depth
    verts = np.array([[-0.9, -0.9, 0.375, 1], [0.9, -0.9, 0.375, 1], [0.0, 0.9, 0.375, 1]], np.float32)
    col = np.ones((3, 4), np.float32)
    vb = np.concatenate([verts, col], axis=1)
    d = scenes.DrawSpec(program=wgc.PROGRAM_CLIP_COLOR, vertex_data=vb.tobytes(), vertex_stride=32,
                        attributes=[(wgc.FLOAT32X4, 0, 0), (wgc.FLOAT32X4, 16, 1)], count=3,
                        depth=(wgc.LESS, True))
    spec = scenes.PassSpec(64, 64, has_depth=True, clear_depth=1.0, draws=[d])
    r = bridge.render(spec, frag_counts=True)
    cov = r.frag_count > 0
    assert cov.sum() > 0
    # lambda sums to 1 only up to rounding, so z*sum(lambda) may move by an ulp
    assert np.all(np.abs(r.depth[cov] - 0.375) <= 4 * np.spacing(np.float32(0.375)))
    assert (r.depth[~cov] == 1.0).all()


def test_colored_triangle_golden_equalities():
    """tests/reference/*.png (LFS oids): colored_triangle equals
    colored_triangle_draw_backwards_no_cull; colored_triangle_cull_front equals
    colored_triangle_draw_backwards, and that image is the clear colour."""
    base = bridge.render(scenes.colored_triangle(wgc.FRONT_CW, wgc.CULL_BACK))
    bw_nc = bridge.render(scenes.colored_triangle(wgc.FRONT_CCW, wgc.CULL_NONE))
    cf = bridge.render(scenes.colored_triangle(wgc.FRONT_CW, wgc.CULL_FRONT))
    bw = bridge.render(scenes.colored_triangle(wgc.FRONT_CCW, wgc.CULL_BACK))
    assert np.array_equal(base.color, bw_nc.color)
    assert np.array_equal(cf.color, bw.color)
    assert (cf.color == np.array([0, 0, 0, 255], np.uint8)).all()
    assert base.fragments == 512 * 256  # half the 512x512 target


def test_back_face_culling_is_winding_swap():
    spec = scenes.config1_plain_small()
    a = bridge.render(spec)
    spec.draws[0].front_face = wgc.FRONT_CW
    spec.draws[0].cull_mode = wgc.CULL_FRONT
    b = bridge.render(spec)
    assert np.array_equal(a.color, b.color) and np.array_equal(a.depth, b.depth)


def test_src_over_closed_form():
    # one opaque-ish triangle over a known clear colour
    verts = np.array([[-1, -1, 0, 1], [3, -1, 0, 1], [-1, 3, 0, 1]], np.float32)
    col = np.tile(np.array([[1.0, 0.0, 0.0, 0.25]], np.float32), (3, 1))
    vb = np.concatenate([verts, col], axis=1)
    d = scenes.DrawSpec(program=wgc.PROGRAM_CLIP_COLOR, vertex_data=vb.tobytes(), vertex_stride=32,
                        attributes=[(wgc.FLOAT32X4, 0, 0), (wgc.FLOAT32X4, 16, 1)], count=3,
                        blend=wgc.BLEND_ALPHA)
    spec = scenes.PassSpec(8, 8, clear_color=(0.0, 0.0, 1.0, 1.0), draws=[d],
                           semantics=wgc.SEMANTICS_EXTENDED)
    r = bridge.render(spec, frag_counts=True)
    # rgb = s*a + d*(1-a), a = sa + da*(1-sa), truncated to u8
    f = np.float32
    exp = [int(f(f(1) * f(0.25)) * f(255)), 0, int(f(f(1) * f(0.75)) * f(255)), 255]
    once = r.frag_count == 1
    assert once.any()
    assert (r.color[once] == np.array(exp, np.uint8)).all()



def test_perspective_correct_closed_form():
    """Extension: along an edge from clip (x=-1, w=1) to (x=4, w=4) the clip
    parameter s at NDC n is (n+1)/(5-3n); a varying that is 0 on the left and
    1 on the right must come out as s, not as the screen-linear (n+1)/2."""
    spec = scenes.perspective_quad()
    W = spec.width
    r = bridge.render(spec, frag_counts=True)
    x = np.arange(W, dtype=np.float64)
    n = 2.0 * x / W - 1.0                      # pixel corner in NDC
    s = (n + 1.0) / (5.0 - 3.0 * n)
    exp = np.floor(255.0 * s)
    got = r.color[:, :, 0].astype(np.int64)
    assert (r.frag_count > 0).all()
    assert np.all(np.abs(got - exp[None, :]) <= 1)
    # and it is far from screen-linear interpolation in the middle
    assert abs(int(got[0, W // 2]) - int(255 * 0.5)) > 40


def test_bilinear_repeat_closed_form():
    """Extension: bilinear filtering with the u*W - 0.5 texel-centre
    convention and Repeat wrap on a 2x2 texture (texel values 200 and 40 in
    red), sampled along u = x/W at the pixel corners of a screen-aligned quad
    (w = 1, so linear = perspective)."""
    spec = scenes.bilinear_strip()
    W = spec.width
    r = bridge.render(spec, frag_counts=True)
    u = np.arange(W, dtype=np.float64) / W
    fu = u * 2.0 - 0.5
    i0 = np.floor(fu)
    a = fu - i0
    t = np.array([200.0, 40.0]) / 255.0
    c0 = t[(i0.astype(np.int64)) % 2]
    c1 = t[(i0.astype(np.int64) + 1) % 2]
    exp = np.floor(255.0 * (c0 * (1 - a) + c1 * a))
    got = r.color[:, :, 0].astype(np.int64)
    assert (r.frag_count > 0).all()
    assert np.all(np.abs(got - exp[None, :]) <= 1)
    # u = 0 lies between the last and the first texel (wrapped): the mean
    assert abs(int(got[0, 0]) - 120) <= 1
