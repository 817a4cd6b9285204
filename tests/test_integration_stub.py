"""The Rust FFI stub in INTEGRATION.md against include/wgc.h.

A Rust adapter copies that stub: a struct with a missing or reordered field,
or a function with a different parameter list, would make libwgc read the
wrong bytes.  Both texts are parsed here and compared: every C structure has
a #[repr(C)] struct with the same fields in the same order and mapped
types, and every declared function has an extern "C" declaration with the
same parameters and return type."""
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent

C_TO_RUST = {"void": "c_void", "char": "c_char", "uint8_t": "u8", "uint32_t": "u32", "int32_t": "i32",
             "uint64_t": "u64", "size_t": "usize", "float": "f32", "double": "f64"}


def rust_of(ctype: str) -> str:
    """Rust spelling of a C type: pointers become *const / *mut (const
    applying to the pointee), scalars their fixed-width names; typedef'd
    handles, structures and the callback keep their names."""
    const = bool(re.search(r"\bconst\b", ctype))
    core = re.sub(r"\bconst\b", "", ctype).strip()
    stars = core.count("*")
    core = core.replace("*", "").strip()
    out = C_TO_RUST.get(core, core)
    for i in range(stars):   # innermost pointer first: `const` qualifies the pointee
        out = ("*const " if const and i == 0 else "*mut ") + out
    return out


def norm(t: str) -> str:
    return re.sub(r"\s+", " ", t.strip())


def header():
    text = re.sub(r"/\*.*?\*/", "", (ROOT / "include" / "wgc.h").read_text(), flags=re.S)
    structs = {}
    for m in re.finditer(r"typedef struct (\w+) \{(.*?)\} \1;", text, flags=re.S):
        fields = []
        for decl in m.group(2).split(";"):
            decl = norm(decl)
            if not decl:
                continue
            const, base, stars, names = re.match(r"(const )?(\w+) ?(\**) ?(.*)$", decl).groups()
            for n in names.split(","):
                n = n.strip()
                extra = "*" * (len(n) - len(n.lstrip("*")))
                n = n.lstrip("*").strip()
                arr = re.match(r"(\w+)\[(\d+)\]$", n)
                t = rust_of((const or "") + base + " " + stars + extra)
                if arr:
                    n, t = arr.group(1), f"[{t}; {arr.group(2)}]"
                fields.append((n, t))
        structs[m.group(1)] = fields
    funcs = {}
    for m in re.finditer(r"([\w \*\n]+?)\b(wgc_\w+)\s*\(([^)]*)\)\s*;", text):
        ret, name, params = norm(m.group(1)), m.group(2), norm(m.group(3))
        ps = []
        if params != "void":
            for p in params.split(","):
                ty, n, arr = re.match(r"(.*?)(\w+)(\[\d+\])?$", p.strip()).groups()
                ps.append((n, rust_of(ty + (" *" if arr else ""))))
        funcs[name] = (ps, None if ret == "void" else rust_of(ret))
    handles = dict(re.findall(r"typedef struct (\w+_T) \*(\w+);", text))
    return structs, funcs, handles


def stub():
    text = (ROOT / "INTEGRATION.md").read_text()
    m = re.search(r"```rust\n(// wgc-sys/src/lib\.rs.*?)```", text, flags=re.S)
    assert m, "INTEGRATION.md holds no wgc-sys stub"
    code = re.sub(r"//[^\n]*", "", m.group(1))
    structs = {}
    for sm in re.finditer(r"#\[repr\(C\)\] pub struct (\w+) \{(.*?)\}", code, flags=re.S):
        body = norm(sm.group(2))
        if body.startswith("_p:"):
            continue
        fields = []
        for f in re.findall(r"pub (\w+): ([^,]+?(?:\[[^\]]*\])?)\s*(?:,|$)", body):
            fields.append((f[0], norm(f[1])))
        structs[sm.group(1)] = fields
    funcs = {}
    ext = re.search(r'extern "C" \{(.*)\}', code, flags=re.S)
    assert ext, "no extern block"
    for fm in re.finditer(r"pub fn (\w+)\((.*?)\)\s*(?:->\s*([^;]+))?;", ext.group(1), flags=re.S):
        params = norm(fm.group(2))
        ps = []
        if params:
            for p in params.split(","):
                p = p.strip()
                if p:
                    n, t = p.split(":", 1)
                    ps.append((n.strip(), norm(t)))
        funcs[fm.group(1)] = (ps, norm(fm.group(3)) if fm.group(3) else None)
    handles = set(re.findall(r"pub type (\w+) = \*mut (\w+_T);", code))
    return structs, funcs, handles, code


def test_rust_of_maps_pointers():
    assert rust_of("const char *") == "*const c_char"
    assert rust_of("void **") == "*mut *mut c_void"
    assert rust_of("const WgcCommandBuffer *") == "*const WgcCommandBuffer"
    assert rust_of("uint64_t *") == "*mut u64"


def test_stub_structs_match_the_header():
    h_structs, _, _ = header()
    r_structs, _, _, _ = stub()
    assert set(h_structs) <= set(r_structs), sorted(set(h_structs) - set(r_structs))
    for name, fields in h_structs.items():
        assert r_structs[name] == fields, (name, r_structs[name], fields)


def test_stub_functions_match_the_header():
    _, h_funcs, _ = header()
    _, r_funcs, _, _ = stub()
    missing = sorted(set(h_funcs) - set(r_funcs))
    assert not missing, missing
    extra = sorted(set(r_funcs) - set(h_funcs))
    assert not extra, extra
    for name, (params, ret) in h_funcs.items():
        assert r_funcs[name] == (params, ret), (name, r_funcs[name], (params, ret))


def test_stub_handles_and_callback():
    _, _, h_handles = header()
    _, _, r_handles, code = stub()
    assert {(v, k) for k, v in h_handles.items()} == r_handles
    assert re.search(r'pub type WgcBufferMapCallback = Option<unsafe extern "C" fn\(status: i32, '
                     r'userdata: \*mut c_void\)>;', code)


@pytest.mark.parametrize("mutation", ["drop_field", "swap_params"])
def test_checker_catches_drift(mutation, monkeypatch):
    """The comparison fails on the two mistakes it exists for."""
    text = (ROOT / "INTEGRATION.md").read_text()
    if mutation == "drop_field":
        bad = text.replace("    pub strip_index_format: u32,\n", "", 1)
    else:
        bad = text.replace("pub fn wgc_buffer_read(buffer: WgcBuffer, offset: u64, dst: *mut c_void, size: u64)",
                           "pub fn wgc_buffer_read(buffer: WgcBuffer, dst: *mut c_void, offset: u64, size: u64)", 1)
    assert bad != text
    real = Path.read_text
    monkeypatch.setattr(Path, "read_text",
                        lambda self, *a, **k: bad if self.name == "INTEGRATION.md" else real(self, *a, **k))
    with pytest.raises(AssertionError):
        test_stub_structs_match_the_header()
        test_stub_functions_match_the_header()
