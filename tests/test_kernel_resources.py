"""Register and scratch use of the built kernels (code-object metadata of
this build's gfx950 objects).

The kernels keep their per-lane state in registers: a kernel with scratch
memory beyond a few spilled registers has regressed (for instance a call
that the compiler did not inline copies the 1 KiB of draw parameters, which
live in the kernel-argument segment, to scratch memory in every wave).
Skipped when the library was not built here or the ROCm tools are absent."""
import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
BUILD = ROOT / "wgpu-cpu_amd" / "wgc" / "_build"
LLVM = Path("/opt/rocm/llvm/bin")
MAX_SCRATCH = 128   # bytes per lane: a few spilled registers at most


def kernel_metadata(obj: Path, tmp: Path):
    fat = tmp / (obj.stem + ".fatbin")
    co = tmp / (obj.stem + ".co")
    subprocess.run([str(LLVM / "llvm-objcopy"), f"--dump-section=.hip_fatbin={fat}", str(obj), str(tmp / "x.o")],
                   check=True, capture_output=True)
    subprocess.run([str(LLVM / "clang-offload-bundler"), "--unbundle", "--type=o", f"--input={fat}",
                    "--targets=hipv4-amdgcn-amd-amdhsa--gfx950", f"--output={co}"], check=True, capture_output=True)
    notes = subprocess.run([str(LLVM / "llvm-readelf"), "--notes", str(co)], check=True, capture_output=True,
                           text=True).stdout
    out = {}
    for blk in notes.split("- .agpr_count")[1:]:
        name = re.search(r"\.name:\s+(\S+)", blk).group(1)
        get = lambda k: int(re.search(r"\." + k + r":\s+(\d+)", blk).group(1))
        out[name] = {"scratch": get("private_segment_fixed_size"), "vgpr": get("vgpr_count"),
                     "spill": get("vgpr_spill_count")}
    return out


def objects():
    return sorted(BUILD.glob("kernels*.o"))


@pytest.mark.skipif(not objects() or not (LLVM / "llvm-readelf").exists(), reason="library not built here")
def test_kernels_keep_their_state_in_registers(tmp_path):
    seen = 0
    bad = []
    for obj in objects():
        for name, m in kernel_metadata(obj, tmp_path).items():
            if not name.startswith("_ZN3wgc"):
                continue
            seen += 1
            if m["scratch"] > MAX_SCRATCH:
                bad.append((obj.name, name, m))
    assert seen > 50, seen   # every unit's kernels were read
    assert not bad, bad


@pytest.mark.skipif(not objects() or not (LLVM / "llvm-readelf").exists(), reason="library not built here")
def test_setup_and_clip_kernels_fit_their_occupancy(tmp_path):
    """Setup and the clipper keep several waves per SIMD: the clipper follows
    one split path per lane (one triangle in registers, not the seven of a
    walk over the split tree), so the clip kernel and the setup kernel that
    clips in place stay within 128 VGPRs (four waves per SIMD); setup
    without the clipper within 96."""
    seen = 0
    for obj in objects():
        for name, m in kernel_metadata(obj, tmp_path).items():
            if "k_setup" in name or "k_clip" in name:
                seen += 1
                clips = "k_clip" in name or re.search(r"k_setupILi\d+ELb[01]ELb1E", name) is not None
                assert m["vgpr"] <= (128 if clips else 96), (obj.name, name, m)
    assert seen >= 20, seen
