"""ctypes binding of the CPU parity oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by the tests, the smoke check and the
CPU-baseline leg of the benchmark.  It executes a ``wgc.scenes.PassSpec``
with the C restatement of the reference's render-pass executor.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from dataclasses import dataclass
from pathlib import Path
from typing import Optional

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

OR_MAX_VB = 4
OR_MAX_ATTR = 8


class OrVertexBuffer(C.Structure):
    _fields_ = [("data", C.c_void_p), ("size", C.c_uint64), ("stride", C.c_uint32),
                ("step_instance", C.c_uint32)]


class OrAttribute(C.Structure):
    _fields_ = [("buffer", C.c_uint32), ("offset", C.c_uint32), ("format", C.c_uint32)]


class OrDraw(C.Structure):
    _fields_ = [
        ("program", C.c_uint32), ("front_face_cw", C.c_uint32), ("cull_mode", C.c_uint32),
        ("depth_enabled", C.c_uint32), ("depth_compare", C.c_uint32), ("depth_write", C.c_uint32),
        ("interp", C.c_uint32), ("blend", C.c_uint32), ("bilinear", C.c_uint32),
        ("srgb_decode", C.c_uint32), ("n_attr", C.c_uint32), ("attr", OrAttribute * OR_MAX_ATTR),
        ("n_vb", C.c_uint32), ("vb", OrVertexBuffer * OR_MAX_VB),
        ("index_data", C.c_void_p), ("index_size", C.c_uint64), ("index_format", C.c_uint32),
        ("uniform", C.c_void_p), ("texture", C.c_void_p),
        ("tex_w", C.c_uint32), ("tex_h", C.c_uint32), ("tex_srgb", C.c_uint32),
        ("addr_u", C.c_uint32), ("addr_v", C.c_uint32),
        ("viewport", C.c_float * 6), ("scissor", C.c_uint32 * 4),
        ("indexed", C.c_uint32), ("first", C.c_uint32), ("count", C.c_uint32),
        ("base_vertex", C.c_int32), ("first_instance", C.c_uint32), ("instance_count", C.c_uint32),
        ("topology", C.c_uint32), ("strip_restart", C.c_uint32),
        ("fs_flags", C.c_uint32), ("fs_early", C.c_uint32),
    ]


class OrPass(C.Structure):
    _fields_ = [
        ("width", C.c_uint32), ("height", C.c_uint32), ("color", C.c_void_p),
        ("color_format", C.c_uint32), ("color_clear", C.c_uint32), ("clear_color", C.c_double * 4),
        ("depth", C.c_void_p), ("depth_clear", C.c_uint32), ("clear_depth", C.c_float),
        ("n_draws", C.c_uint32), ("draws", C.POINTER(OrDraw)), ("frag_count", C.c_void_p),
        ("fragments", C.c_uint64), ("error", C.c_int), ("message", C.c_char * 256),
    ]


_lib: Optional[C.CDLL] = None


def build() -> Path:
    """Compiles liboracle.so with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True, capture_output=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        L.oracle_execute_pass.argtypes = [C.POINTER(OrPass)]
        L.oracle_execute_pass.restype = C.c_int
        L.oracle_bubblesort3_i64.argtypes = [C.POINTER(C.c_int64)]
        L.oracle_bubblesort3_i64.restype = None
        L.oracle_scanline_pixels.argtypes = [C.POINTER(C.c_uint32), C.c_void_p, C.c_int64]
        L.oracle_scanline_pixels.restype = C.c_int64
        L.oracle_clip_triangle.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_int]
        L.oracle_clip_triangle.restype = C.c_int
        L.oracle_to_raster.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_float),
                                       C.POINTER(C.c_uint32), C.POINTER(C.c_float)]
        L.oracle_to_raster.restype = None
        L.oracle_f32_to_unorm8.argtypes = [C.c_float]
        L.oracle_f32_to_unorm8.restype = C.c_uint8
        L.oracle_texel_coordinate.argtypes = [C.c_float, C.c_uint32, C.c_uint32]
        L.oracle_texel_coordinate.restype = C.c_uint32
        L.oracle_srgb_table.argtypes = []
        L.oracle_srgb_table.restype = C.POINTER(C.c_float)
        L.oracle_mat4_mul_vec4.argtypes = [C.POINTER(C.c_float)] * 3
        L.oracle_mat4_mul_vec4.restype = None
        L.oracle_bresenham.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.c_void_p, C.c_void_p,
                                       C.c_int64]
        L.oracle_bresenham.restype = C.c_int64
        L.oracle_clip_line.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                       C.POINTER(C.c_float)]
        L.oracle_clip_line.restype = C.c_int
        L.oracle_assemble.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p,
                                      C.c_int64]
        L.oracle_assemble.restype = C.c_int64
        _lib = L
    return _lib


@dataclass
class OracleResult:
    color: Optional[np.ndarray]
    depth: Optional[np.ndarray]
    fragments: int
    frag_count: Optional[np.ndarray]


class OracleError(RuntimeError):
    pass


def render(spec, frag_counts: bool = False) -> OracleResult:
    """Executes a PassSpec on the CPU oracle (single thread)."""
    from wgc import BGRA8_UNORM, BGRA8_UNORM_SRGB, FILTER_LINEAR, SEMANTICS_EXTENDED, BLEND_ALPHA

    L = lib()
    w, h = spec.width, spec.height
    keep = []
    color = None
    if spec.has_color:
        color = (np.frombuffer(spec.color_init, np.uint8).copy() if spec.color_init is not None
                 else np.zeros(w * h * 4, np.uint8))
    depth = None
    if spec.has_depth:
        depth = (np.frombuffer(spec.depth_init, np.float32).copy() if spec.depth_init is not None
                 else np.zeros(w * h, np.float32))
    fc = np.zeros(w * h, np.uint32) if frag_counts else None
    ext = spec.semantics == SEMANTICS_EXTENDED
    draws = (OrDraw * max(1, len(spec.draws)))()
    cur_vp = (0.0, 0.0, float(w), float(h), 0.0, 1.0)
    cur_sc = (0, 0, w, h)
    for i, d in enumerate(spec.draws):
        o = draws[i]
        o.program = d.program
        o.front_face_cw = d.front_face
        o.cull_mode = d.cull_mode
        if d.depth is not None:
            o.depth_enabled = 1
            o.depth_compare = d.depth[0]
            o.depth_write = 1 if d.depth[1] else 0
        o.interp = d.interp
        o.blend = 1 if (ext and d.blend == BLEND_ALPHA and spec.has_color) else 0
        o.bilinear = 1 if (ext and d.filter == FILTER_LINEAR) else 0
        o.srgb_decode = 1 if ext else 0
        if d.vertex_data:
            vbuf = (C.c_uint8 * len(d.vertex_data)).from_buffer_copy(d.vertex_data)
            keep.append(vbuf)
            o.n_vb = 1
            o.vb[0] = OrVertexBuffer(C.addressof(vbuf), len(d.vertex_data), d.vertex_stride, 0)
            n_attr = 0
            for fmt, off, loc in d.attributes:
                o.attr[loc] = OrAttribute(0, off, fmt)
                n_attr = max(n_attr, loc + 1)
            o.n_attr = n_attr
        if d.instance_buffer is not None:
            data, stride, attrs = d.instance_buffer
            ibuf_i = (C.c_uint8 * len(data)).from_buffer_copy(data)
            keep.append(ibuf_i)
            slot = o.n_vb
            o.vb[slot] = OrVertexBuffer(C.addressof(ibuf_i), len(data), stride, 1)
            o.n_vb = slot + 1
            for fmt, off, loc in attrs:
                o.attr[loc] = OrAttribute(slot, off, fmt)
                o.n_attr = max(o.n_attr, loc + 1)
        if d.index_data is not None:
            ibuf = (C.c_uint8 * len(d.index_data)).from_buffer_copy(d.index_data)
            keep.append(ibuf)
            o.index_data = C.addressof(ibuf)
            o.index_size = len(d.index_data)
            o.index_format = d.index_format
            o.indexed = 1
        if d.uniform is not None:
            ubuf = (C.c_uint8 * len(d.uniform)).from_buffer_copy(d.uniform)
            keep.append(ubuf)
            o.uniform = C.addressof(ubuf)
        if d.texture is not None:
            tw, th, tb, srgb = d.texture
            tbuf = (C.c_uint8 * len(tb)).from_buffer_copy(tb)
            keep.append(tbuf)
            o.texture = C.addressof(tbuf)
            o.tex_w, o.tex_h, o.tex_srgb = tw, th, 1 if srgb else 0
            o.addr_u, o.addr_v = d.address_u, d.address_v
        # viewport and scissor are sticky pass state (state.rs:211-217): a
        # draw without its own keeps what an earlier draw of the pass set
        if d.viewport is not None:
            cur_vp = d.viewport
        if d.scissor is not None:
            cur_sc = d.scissor
        for k in range(6):
            o.viewport[k] = cur_vp[k]
        for k in range(4):
            o.scissor[k] = cur_sc[k]
        o.first, o.count = d.first, d.count
        o.topology = d.topology
        if d.strip_index_format:
            # state.rs:304-312: strip_index_format must equal the bound index format
            if d.index_data is not None and d.strip_index_format != d.index_format:
                raise OracleError("strip_index_format differs from the index format")
            o.strip_restart = 1
        o.fs_flags, o.fs_early = d.fs_flags, d.fs_early
        o.base_vertex = d.base_vertex
        o.first_instance, o.instance_count = d.first_instance, d.instance_count
    p = OrPass()
    p.width, p.height = w, h
    if color is not None:
        p.color = color.ctypes.data
        p.color_format = spec.color_format
        p.color_clear = 1 if spec.clear_color is not None else 0
        for k, v in enumerate(spec.clear_color or (0, 0, 0, 0)):
            p.clear_color[k] = v
    if depth is not None:
        p.depth = depth.ctypes.data
        p.depth_clear = 1 if spec.clear_depth is not None else 0
        p.clear_depth = spec.clear_depth if spec.clear_depth is not None else 0.0
    p.n_draws = len(spec.draws)
    p.draws = C.cast(draws, C.POINTER(OrDraw))
    if fc is not None:
        p.frag_count = fc.ctypes.data
    rc = L.oracle_execute_pass(C.byref(p))
    if rc != 0:
        raise OracleError(p.message.decode())
    return OracleResult(color.reshape(h, w, 4) if color is not None else None,
                        depth.reshape(h, w) if depth is not None else None, int(p.fragments),
                        fc.reshape(h, w) if fc is not None else None)


# ---- pieces for the known-answer tests ----

def bubblesort3(v):
    arr = (C.c_int64 * 3)(*v)
    lib().oracle_bubblesort3_i64(arr)
    return list(arr)


def scanline_pixels(tri):
    t = (C.c_uint32 * 6)(*tri)
    n = lib().oracle_scanline_pixels(t, None, 0)
    if n < 0:
        raise OracleError("scanline panicked")
    out = np.zeros(2 * max(n, 1), np.uint32)
    lib().oracle_scanline_pixels(t, out.ctypes.data, n)
    return [tuple(x) for x in out[:2 * n].reshape(-1, 2)]


def clip_triangle(verts):
    inp = (C.c_float * 12)(*[float(x) for v in verts for x in v])
    out = np.zeros(64 * 3 * 7, np.float32)
    n = lib().oracle_clip_triangle(inp, out.ctypes.data, 64)
    return out[:n * 21].reshape(n, 3, 7)


def to_raster(viewport, clip):
    vp = (C.c_float * 6)(*viewport)
    cp = (C.c_float * 4)(*clip)
    fb = (C.c_uint32 * 2)()
    fr = (C.c_float * 4)()
    lib().oracle_to_raster(vp, cp, fb, fr)
    return (fb[0], fb[1]), tuple(fr)


def f32_to_unorm8(v: float) -> int:
    return lib().oracle_f32_to_unorm8(v)


def texel_coordinate(x: float, mode: int, size: int) -> int:
    return lib().oracle_texel_coordinate(x, mode, size)


def mat4_mul_vec4(m_cols, v):
    """m_cols: 4 columns of 4 floats."""
    m = (C.c_float * 16)(*[float(x) for col in m_cols for x in col])
    vv = (C.c_float * 4)(*[float(x) for x in v])
    r = (C.c_float * 4)()
    lib().oracle_mat4_mul_vec4(m, vv, r)
    return list(r)


def srgb_table() -> np.ndarray:
    p = lib().oracle_srgb_table()
    return np.ctypeslib.as_array(p, shape=(256,)).copy()


def bresenham(start, end):
    """util/bresenham.rs: [(x, y), ...] and [t, ...] from start (inclusive) to end (exclusive)."""
    a = (C.c_uint32 * 2)(*start)
    b = (C.c_uint32 * 2)(*end)
    n = lib().oracle_bresenham(a, b, None, None, 0)
    xy = np.zeros(2 * max(n, 1), np.uint32)
    t = np.zeros(max(n, 1), np.float32)
    lib().oracle_bresenham(a, b, xy.ctypes.data, t.ctypes.data, n)
    return [tuple(int(v) for v in xy[2 * i:2 * i + 2]) for i in range(n)], t[:n]


def clip_line(a, b):
    """clipper.rs Cohen-Sutherland: None or ((p0, p1), (alpha0, alpha1))."""
    fa = (C.c_float * 4)(*a)
    fb = (C.c_float * 4)(*b)
    out = (C.c_float * 8)()
    al = (C.c_float * 2)()
    if not lib().oracle_clip_line(fa, fb, out, al):
        return None
    return (list(out[:4]), list(out[4:])), (al[0], al[1])


def assemble(topology, count, indices=None, index_format=4, restart=False):
    """primitive.rs assembly: index positions of every primitive."""
    raw = None
    keep = None
    if indices is not None:
        keep = np.asarray(indices, np.uint16 if index_format == 2 else np.uint32).tobytes()
        raw = C.c_char_p(keep)
    n = lib().oracle_assemble(topology, 1 if restart else 0, index_format, raw, count, None, 0)
    if n < 0:
        raise OracleError("strip_index_format set, but not a strip topology")
    nv = 1 if topology == 0 else (2 if topology in (1, 2) else 3)
    out = np.zeros(max(n, 1) * nv, np.uint32)
    lib().oracle_assemble(topology, 1 if restart else 0, index_format, raw, count, out.ctypes.data, n)
    return [list(int(v) for v in out[i * nv:(i + 1) * nv]) for i in range(n)]
