"""The C ABI driven from C++ alone (wgpu-cpu_amd/harness/colored_triangle.cpp,
built by __graft_entry__.build): the reference's colored_triangle scenes
(wgpu-cpu-tests/src/tests/colored_triangle.rs) rendered from the WGSL shader,
copied to a buffer and read through a mapping, equal the oracle's render."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import wgc
from oracle import bridge
from wgc import scenes

from test_gpu_parity import announce

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
HARNESS = ROOT / "wgpu-cpu_amd" / "harness" / "colored_triangle"
SHADER = ROOT / "wgpu-cpu_amd" / "shaders" / "colored_triangle.wgsl"


@pytest.mark.parametrize("topology,vertices,ff,cull", [
    (3, 3, wgc.FRONT_CCW, wgc.CULL_NONE),
    (3, 3, wgc.FRONT_CW, wgc.CULL_FRONT),
    (3, 3, wgc.FRONT_CCW, wgc.CULL_BACK),
    (2, 4, wgc.FRONT_CW, wgc.CULL_NONE),      # colored_triangle_lines
], ids=["plain", "cull_front", "backwards", "lines"])
def test_cpp_harness_matches_the_oracle(tmp_path, topology, vertices, ff, cull):
    announce(f"C++ harness topology={topology} ff={ff} cull={cull}")
    if not HARNESS.exists():
        pytest.fail("the C++ harness is not built (run __graft_entry__.build())")
    out = tmp_path / "target.rgba"
    subprocess.run([str(HARNESS), str(SHADER), str(out), str(topology), str(vertices), str(ff), str(cull)],
                   check=True, timeout=120)
    got = np.frombuffer(out.read_bytes(), np.uint8).reshape(512, 512, 4)
    ref = bridge.render(scenes.colored_triangle(ff, cull, vertices=vertices, topology=topology))
    assert np.array_equal(got, ref.color)
    assert got.any()
