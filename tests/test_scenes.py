"""Host-side scene generation: seeded, deterministic, and of the sizes the
workloads name."""
import numpy as np

from wgc import scenes


def test_splitmix_reference_values():
    # SplitMix64 seeded with 0: first outputs (Vigna's reference generator)
    u = scenes.splitmix_u64(0, 3)
    assert [int(x) for x in u] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_config_sizes():
    s1 = scenes.config1_plain_small()
    d = s1.draws[0]
    assert len(d.vertex_data) // 28 == 561 and d.count == 960 * 3
    s2 = scenes.config2_textured()
    d = s2.draws[0]
    assert len(d.vertex_data) // 24 == 33153 and d.count == 65024 * 3
    assert (s2.width, s2.height) == (1920, 1080)
    assert d.texture[0] == 1024 and len(d.texture[2]) == 1024 * 1024 * 4


def test_generation_is_deterministic():
    a = scenes.config3_overdraw(n=1000, width=320, height=305)
    b = scenes.config3_overdraw(n=1000, width=320, height=305)
    assert a.draws[0].vertex_data == b.draws[0].vertex_data
    v = np.frombuffer(a.draws[0].vertex_data, np.float32).reshape(-1, 8)
    assert (v[:, 7] >= 0.1).all() and (v[:, 7] <= 0.9).all()


def test_checker_texture():
    t = np.frombuffer(scenes.checker_texture(64, 32, 16, 5), np.uint8).reshape(64, 64, 4)
    assert (t[..., 3] == 255).all()
    assert t[0:32, 0:32, 0].min() >= 239 and t[0:32, 32:64, 0].max() <= 16


def test_serving_trace_is_poisson_and_sorted():
    from wgc import serving
    tr = serving.make_trace(0x5EED_0005, rate=2000.0, seconds=2.0, n_meshes=8)
    t = np.array([r.arrival for r in tr])
    assert np.all(np.diff(t) >= 0) and t[-1] <= 2.0
    # 4000 expected arrivals; Poisson sd = 63
    assert abs(len(tr) - 4000) < 5 * 63
    gaps = np.diff(t)
    assert abs(gaps.mean() - 1 / 2000.0) < 0.1 / 2000.0
    assert all(64 <= r.width <= 1024 and 64 <= r.height <= 1024 for r in tr)
    assert {r.mesh for r in tr} == set(range(8))
