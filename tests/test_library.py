"""The C ABI library: it loads without a GPU and exports every entry point
that include/*.h declares; the Python mirror binds each of them."""
import re
from pathlib import Path

import pytest

import wgc
from wgc import _build

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(wgc_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


@pytest.fixture(scope="module")
def lib():
    if not wgc.library_path().exists():
        _build.build_library()
    return wgc.load_library()


def test_header_declares_the_boundary():
    names = declared_functions()
    assert "wgc_queue_submit" in names and "wgc_render_pass_draw_indexed" in names
    assert len(names) >= 40


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_mirror_binds_every_declared_symbol():
    assert set(declared_functions()) == set(wgc.SIGNATURES)


def test_abi_version_and_error_channel(lib):
    assert lib.wgc_abi_version() == 1
    # a call that fails before touching HIP sets the thread-local message
    rc = lib.wgc_device_create(None, None)
    assert rc == -1
    assert b"null" in lib.wgc_last_error()
