"""Pins the CPU oracle against the reference's own known-answer tests.

Every vector below is data from the reference's #[cfg(test)] modules (cited
per test); the oracle is called through its C entry points.
"""
import math

import numpy as np
import pytest

from oracle import bridge


# wgpu-cpu/src/util/sort.rs:61-72 (bubblesort3_bug) and :74-134 (random_bubblesorts)
def test_bubblesort3_bug():
    assert bridge.bubblesort3([230, 84, 75]) == [75, 84, 230]


SORT_PAIRS = [
    ([847, 895, 730], [730, 847, 895]), ([407, 622, 50], [50, 407, 622]),
    ([498, 698, 242], [242, 498, 698]), ([729, 619, 845], [619, 729, 845]),
    ([78, 484, 855], [78, 484, 855]), ([856, 128, 44], [44, 128, 856]),
    ([64, 618, 512], [64, 512, 618]), ([132, 858, 663], [132, 663, 858]),
    ([122, 182, 486], [122, 182, 486]), ([521, 930, 464], [464, 521, 930]),
    ([536, 956, 417], [417, 536, 956]), ([440, 364, 860], [364, 440, 860]),
    ([46, 855, 920], [46, 855, 920]), ([226, 163, 389], [163, 226, 389]),
    ([369, 992, 285], [285, 369, 992]), ([823, 464, 856], [464, 823, 856]),
    ([847, 1000, 179], [179, 847, 1000]), ([144, 7, 213], [7, 144, 213]),
    ([839, 82, 108], [82, 108, 839]), ([36, 649, 856], [36, 649, 856]),
    ([508, 363, 603], [363, 508, 603]), ([252, 34, 29], [29, 34, 252]),
    ([274, 629, 663], [274, 629, 663]), ([900, 439, 406], [406, 439, 900]),
    ([231, 941, 884], [231, 884, 941]), ([98, 524, 25], [25, 98, 524]),
    ([714, 63, 201], [63, 201, 714]), ([587, 748, 250], [250, 587, 748]),
    ([569, 602, 123], [123, 569, 602]), ([896, 104, 31], [31, 104, 896]),
    ([525, 463, 510], [463, 510, 525]), ([290, 324, 18], [18, 290, 324]),
    ([822, 770, 850], [770, 822, 850]), ([330, 592, 965], [330, 592, 965]),
    ([489, 70, 491], [70, 489, 491]), ([166, 434, 779], [166, 434, 779]),
    ([430, 435, 367], [367, 430, 435]), ([580, 220, 810], [220, 580, 810]),
    ([807, 391, 254], [254, 391, 807]), ([214, 297, 524], [214, 297, 524]),
    ([494, 777, 964], [494, 777, 964]), ([452, 409, 62], [62, 409, 452]),
    ([676, 631, 494], [494, 631, 676]), ([297, 97, 938], [97, 297, 938]),
    ([381, 788, 572], [381, 572, 788]), ([330, 192, 138], [138, 192, 330]),
    ([354, 229, 575], [229, 354, 575]), ([290, 110, 807], [110, 290, 807]),
    ([258, 313, 183], [183, 258, 313]), ([253, 103, 815], [103, 253, 815]),
]


def test_bubblesort3_random():
    # the reference's test sorts the *expected* triple again and compares;
    # sorting the unsorted input must give it as well
    for unsorted, expected in SORT_PAIRS:
        assert bridge.bubblesort3(expected) == expected
        assert bridge.bubblesort3(unsorted) == expected


# wgpu-cpu/src/util/scanline.rs:206-275: regression triangles must not panic
def test_scanline_regressions_do_not_panic():
    px = bridge.scanline_pixels([0, 0, 120, 0, 119, 26])
    assert len(px) > 0
    px = bridge.scanline_pixels([256, 230, 257, 84, 358, 75])
    assert len(px) > 0
    assert all(x >= 0 for x, _ in px)


# wgpu-cpu/src/render_pass/clipper.rs:823-892
def test_clip_passes_through_inside():
    tri = [(0.0, 0.5, 0.0, 1.0), (-0.5, -0.5, 0.0, 1.0), (0.5, -0.5, 0.0, 1.0)]
    out = bridge.clip_triangle(tri)
    assert out.shape[0] == 1
    assert np.array_equal(out[0, :, :4], np.array(tri, np.float32))
    assert np.array_equal(out[0, :, 4:], np.eye(3, dtype=np.float32))


def test_clip_rejects_outside():
    tri = [(-2.0, 0.5, 0.0, 1.0), (-2.5, -0.5, 0.0, 1.0), (-1.5, -0.5, 0.0, 1.0)]
    assert bridge.clip_triangle(tri).shape[0] == 0


def test_clip_one_vertex_outside():
    tri = [(0.0, 1.5, 0.0, 1.0), (-0.5, 0.5, 0.0, 1.0), (0.5, 0.5, 0.0, 1.0)]
    out = bridge.clip_triangle(tri)[:, :, :4]
    expected = np.array([
        [(-0.25, 1.0, 0.0, 1.0), (-0.5, 0.5, 0.0, 1.0), (0.25, 1.0, 0.0, 1.0)],
        [(-0.5, 0.5, 0.0, 1.0), (0.5, 0.5, 0.0, 1.0), (0.25, 1.0, 0.0, 1.0)],
    ], np.float32)
    assert out.shape == expected.shape
    assert np.allclose(out, expected, atol=1e-6)


def test_clip_one_vertex_inside():
    tri = [(0.0, -0.5, 0.0, 1.0), (-0.5, -1.5, 0.0, 1.0), (0.5, -1.5, 0.0, 1.0)]
    out = bridge.clip_triangle(tri)[:, :, :4]
    expected = np.array([[(0.0, -0.5, 0.0, 1.0), (-0.25, -1.0, 0.0, 1.0), (0.25, -1.0, 0.0, 1.0)]],
                        np.float32)
    assert out.shape == expected.shape
    assert np.allclose(out, expected, atol=1e-6)


def test_clip_barycentrics_reproduce_positions():
    # invariant: each clipped vertex is the barycentric combination of the
    # original vertices (the clipper lerps positions and barycentrics alike)
    rng = np.random.default_rng(7)
    for _ in range(300):
        tri = rng.uniform(-3, 3, size=(3, 4)).astype(np.float32)
        tri[:, 3] = rng.uniform(0.5, 3, size=3)
        out = bridge.clip_triangle(tri.tolist())
        for sub in out:
            for v in sub:
                p = v[4:7] @ tri
                assert np.allclose(p, v[:4], atol=2e-5 * (1 + np.abs(tri).max()))
                assert abs(v[4:7].sum() - 1.0) < 1e-5
                w = v[3]
                assert v[0] >= -w - 1e-4 and v[0] <= w + 1e-4
                assert v[1] >= -w - 1e-4 and v[1] <= w + 1e-4
                assert v[2] >= -1e-4 and v[2] <= w + 1e-4


# naga-cranelift/src/tests.rs:1287-1315 (matrix_vector_product)
def test_matrix_vector_product():
    cols = [(1, 1, 1, 1), (0, 2, 0, 2), (3, 2, 1, 0), (1, 2, 3, 4)]
    assert bridge.mat4_mul_vec4(cols, (1, 2, 3, 4)) == [14.0, 19.0, 16.0, 21.0]


def test_matrix_vector_product_is_column_sum_without_fma():
    # ((c0*v0 + c1*v1) + c2*v2) + c3*v3 in binary32, each op rounded once
    rng = np.random.default_rng(11)
    for _ in range(200):
        m = rng.normal(size=(4, 4)).astype(np.float32)
        v = rng.normal(size=4).astype(np.float32)
        r = bridge.mat4_mul_vec4(m.tolist(), v.tolist())
        exp = m[0] * v[0]
        exp = (exp + m[1] * v[1]).astype(np.float32)
        exp = (exp + m[2] * v[2]).astype(np.float32)
        exp = (exp + m[3] * v[3]).astype(np.float32)
        assert np.array_equal(np.array(r, np.float32), exp)


# wgpu-cpu/src/render_pass/primitive.rs:508-531 (triangle_front_face) through
# the cull decision of a full render: Cw triangle with front=Cw is visible
def test_triangle_front_face_through_culling():
    from wgc import scenes, FRONT_CW, FRONT_CCW, CULL_BACK, CULL_FRONT
    r1 = bridge.render(scenes.colored_triangle(FRONT_CW, CULL_BACK))
    r2 = bridge.render(scenes.colored_triangle(FRONT_CCW, CULL_BACK))
    assert r1.fragments > 0 and r2.fragments == 0


# texture.rs:376-379: (v * 255).clamp(0, 255) as u8 truncates
@pytest.mark.parametrize("v,expected", [(0.0, 0), (1.0, 255), (0.5, 127), (0.999, 254), (-1.0, 0),
                                        (2.0, 255), (float("nan"), 0), (1 / 255, 1), (0.00392, 0)])
def test_unorm8_truncates(v, expected):
    assert bridge.f32_to_unorm8(v) == expected


# binding.rs:150-164: rem_euclid then round(x * (size - 1)), half away from zero
@pytest.mark.parametrize("x,mode,size,expected", [
    (0.0, 1, 1024, 0), (1.0, 1, 1024, 0), (0.5, 1, 1024, 512), (-0.25, 1, 4, 2),
    (1.25, 1, 4, 1), (0.5, 1, 2, 1), (0.5, 0, 4, 2), (1.5, 0, 4, 3), (-3.0, 0, 4, 0),
    (1.5, 2, 4, 2), (0.5, 2, 4, 2), (-1e-9, 1, 8, 7), (float("nan"), 0, 8, 0),
])
def test_texel_coordinate(x, mode, size, expected):
    assert bridge.texel_coordinate(x, mode, size) == expected


def test_to_raster_closed_form():
    # raster.rs:145-160 on a 512x512 viewport
    vp = (0.0, 0.0, 512.0, 512.0, 0.0, 1.0)
    fb, fr = bridge.to_raster(vp, (-1.0, -1.0, 0.25, 1.0))
    assert fb == (0, 512) and fr == (0.0, 512.0, 0.25, 1.0)
    fb, fr = bridge.to_raster(vp, (0.0, 1.0, 0.5, 2.0))
    assert fb == (256, 128) and fr == (256.0, 128.0, 0.25, 0.5)
    # negative viewport coordinates clamp to 0 (max(0.0))
    fb, fr = bridge.to_raster(vp, (-2.0, 0.0, 0.0, 1.0))
    assert fb == (0, 256) and fr[0] == 0.0


def test_srgb_table_endpoints():
    t = bridge.srgb_table()
    assert t[0] == 0.0 and t[255] == 1.0
    assert abs(t[10] - (10 / 255) / 12.92) < 1e-7
    assert np.all(np.diff(t) > 0)


# sRGB -> linear values of IEC 61966-2-1 for 8-bit codes, as published in
# sRGB conversion tables (7 decimals).  Codes 10 and 11 sit on either side of
# the 0.04045 threshold; the others are in the power branch, where an
# exponent of 2.2 (128 -> 0.2158605 becomes 0.2195) or a shifted threshold
# fails.
SRGB_PUBLISHED = {10: 0.0030353, 11: 0.0033465, 64: 0.0512695, 128: 0.2158605,
                  188: 0.5028865, 200: 0.5775804, 255: 1.0}


@pytest.mark.parametrize("code,linear", sorted(SRGB_PUBLISHED.items()))
def test_srgb_table_published_values(code, linear):
    t = bridge.srgb_table()
    # 7-decimal rounding (5e-8) plus one f32 ulp at 1.0 (6e-8)
    assert abs(float(t[code]) - linear) <= 1.1e-7, (code, float(t[code]), linear)


def test_srgb_table_inverts_to_its_codes():
    """The encoding direction of the same standard (linear -> sRGB:
    12.92 L below 0.0031308, else 1.055 L^(1/2.4) - 0.055) maps every table
    entry back to its own code: a different formula, so a mistake in the
    decoding branch or its threshold shows up as a code that does not come
    back."""
    t = bridge.srgb_table().astype(np.float64)
    enc = np.where(t <= 0.0031308, 12.92 * t, 1.055 * np.power(t, 1.0 / 2.4) - 0.055)
    back = np.round(enc * 255.0).astype(int)
    assert np.array_equal(back, np.arange(256))
    # and the round trip is tight, not merely within half a code
    assert np.abs(enc * 255.0 - np.arange(256)).max() < 1e-4
