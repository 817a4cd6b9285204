"""Render-pass descriptions and the seeded synthetic scenes of the workloads.

A :class:`PassSpec` is a complete, API-level description of one render pass
(attachments, load ops, and a list of draws with their pipeline state and
resource bytes).  ``wgc.runner.render`` executes it through libwgc; the
parity tests hand the same description to the CPU oracle.

Scene generation is host-side numpy; both executors consume the same bytes,
so nothing here needs to be bit-reproducible across machines except through
its seeds.  PRNG: SplitMix64, ``f32 = (u >> 40) * 2^-24``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

from . import (ALWAYS, BLEND_ALPHA, BLEND_REPLACE, CLAMP_TO_EDGE, CULL_BACK, CULL_NONE, FILTER_LINEAR,
               FILTER_NEAREST, FLOAT32X2, FLOAT32X3, FLOAT32X4, FRONT_CCW, FRONT_CW, INTERP_LINEAR,
               INTERP_PERSPECTIVE, LESS, PROGRAM_CLIP_COLOR, PROGRAM_COLORED_TRIANGLE, PROGRAM_MVP_COLOR,
               PROGRAM_MVP_TEXTURE, REPEAT, RGBA8_UNORM, RGBA8_UNORM_SRGB, SEMANTICS_EXTENDED,
               SEMANTICS_REFERENCE, STEP_VERTEX)

SHADER_DIR = __import__("pathlib").Path(__file__).resolve().parent.parent / "shaders"
_SHADER_FILES = {PROGRAM_MVP_COLOR: "mvp_color.wgsl", PROGRAM_MVP_TEXTURE: "mvp_texture.wgsl",
                 PROGRAM_COLORED_TRIANGLE: "colored_triangle.wgsl", PROGRAM_CLIP_COLOR: "clip_color.wgsl"}


def wgsl_source(program: int, interp: int = INTERP_LINEAR, color_components: int = 0) -> str:
    """The WGSL module of a built-in program (wgpu-cpu_amd/shaders), with the
    @interpolate attribute for `interp` and a vec3 / vec4 colour input."""
    src = (SHADER_DIR / _SHADER_FILES[program]).read_text()
    mode = {INTERP_LINEAR: "linear", INTERP_PERSPECTIVE: "perspective"}.get(interp, "flat")
    src = src.replace("@interpolate(linear)", f"@interpolate({mode})")
    if program in (PROGRAM_MVP_COLOR, PROGRAM_CLIP_COLOR) and color_components in (3, 4):
        if color_components == 4 and program == PROGRAM_MVP_COLOR:
            src = src.replace("@location(1) color: vec3f", "@location(1) color: vec4f")
            src = src.replace("@interpolate(" + mode + ") color: vec3f", "@interpolate(" + mode + ") color: vec4f")
            src = src.replace("vec4f(f.color.x, f.color.y, f.color.z, 1.0)", "f.color")
        if color_components == 3 and program == PROGRAM_CLIP_COLOR:
            src = src.replace("@location(1) color: vec4f", "@location(1) color: vec3f")
            src = src.replace("@interpolate(" + mode + ") color: vec4f", "@interpolate(" + mode + ") color: vec3f")
            src = src.replace("return f.color;", "return vec4f(f.color.x, f.color.y, f.color.z, 1.0);")
    return src


FS_DISCARD, FS_DEPTH_ALPHA, FS_DEPTH_CONST, FS_DEPTH_FIRST, FS_PRIM_RED = 1, 2, 4, 8, 16
EARLY_NONE, EARLY_FORCE, EARLY_ALLOW = 0, 1, 2


def fragment_test_wgsl(flags: int, early: int, interp: int = INTERP_LINEAR) -> str:
    """WGSL of a fragment test program: clip-space position and a vec4
    colour c passed through; the fragment stage discards when c.w < 0.5
    (FS_DISCARD), writes @builtin(frag_depth) = c.w or 0.3125
    (FS_DEPTH_ALPHA / FS_DEPTH_CONST) as a member before (FS_DEPTH_FIRST) or
    after the colour, puts f32(primitive_index % 8u) * 0.125 in red
    (FS_PRIM_RED), and carries @early_depth_test(force) or
    @early_depth_test(less_equal) (a conservative depth: the early test is
    allowed and the late one kept)."""
    mode = {INTERP_LINEAR: "linear", INTERP_PERSPECTIVE: "perspective"}.get(interp, "flat")
    depth = flags & (FS_DEPTH_ALPHA | FS_DEPTH_CONST)
    members = ["@location(0) color: vec4f"]
    if depth:
        d = "@builtin(frag_depth) depth: f32"
        members = [d] + members if flags & FS_DEPTH_FIRST else members + [d]
    attr = {EARLY_NONE: "", EARLY_FORCE: "@early_depth_test(force)\n",
            EARLY_ALLOW: "@early_depth_test(less_equal)\n"}[early]
    params = "f: VsOut" + (", @builtin(primitive_index) pi: u32" if flags & FS_PRIM_RED else "")
    body = []
    if flags & FS_DISCARD:
        body.append("    if (f.color.w < 0.5) {\n        discard;\n    }")
    red = "f32(pi % 8u) * 0.125" if flags & FS_PRIM_RED else "f.color.x"
    body.append(f"    let c = vec4f({red}, f.color.y, f.color.z, f.color.w);")
    if depth:
        dv = "f.color.w" if flags & FS_DEPTH_ALPHA else "0.3125"
        args = ["c"]
        args = ([dv] + args) if flags & FS_DEPTH_FIRST else (args + [dv])
        body.append(f"    return FsOut({', '.join(args)});")
        ret = "FsOut"
    else:
        body.append("    return c;")
        ret = "@location(0) vec4f"
    fs_struct = ("struct FsOut {\n    " + ",\n    ".join(members) + ",\n}\n\n") if depth else ""
    return ("struct VsIn {\n    @location(0) position: vec4f,\n    @location(1) color: vec4f,\n}\n\n"
            "struct VsOut {\n    @builtin(position) position: vec4f,\n"
            f"    @location(0) @interpolate({mode}) color: vec4f,\n}}\n\n" + fs_struct +
            "@vertex\nfn vs_main(v: VsIn) -> VsOut {\n    return VsOut(v.position, v.color);\n}\n\n"
            f"@fragment\n{attr}fn fs_main({params}) -> {ret} {{\n" + "\n".join(body) + "\n}\n")


def as_wgsl(spec: "PassSpec") -> "PassSpec":
    """A copy of `spec` whose draws run the WGSL modules of their programs."""
    import copy
    out = copy.deepcopy(spec)
    for d in out.draws:
        ncol = 0
        attrs = list(d.attributes) + (list(d.instance_buffer[2]) if d.instance_buffer else [])
        for fmt, _off, loc in attrs:
            if loc == 1:
                ncol = {FLOAT32X3: 3, FLOAT32X4: 4}.get(fmt, 0)
        if d.fs_flags or d.fs_early:
            d.wgsl = fragment_test_wgsl(d.fs_flags, d.fs_early, d.interp)
        else:
            d.wgsl = wgsl_source(d.program, d.interp, ncol)
    out.name = (spec.name or "scene") + " (wgsl)"
    return out


GOLDEN = np.uint64(0x9E3779B97F4A7C15)
SEED_BASE = 0x5EED_0000


@dataclass
class DrawSpec:
    program: int
    interp: int = INTERP_LINEAR
    vertex_data: bytes = b""
    vertex_stride: int = 0
    attributes: List[Tuple[int, int, int]] = field(default_factory=list)  # (format, offset, location)
    index_data: Optional[bytes] = None
    index_format: int = 4
    uniform: Optional[bytes] = None            # 64 B column-major mat4x4f
    texture: Optional[Tuple[int, int, bytes, bool]] = None  # (w, h, rgba bytes, srgb)
    address_u: int = REPEAT
    address_v: int = REPEAT
    filter: int = FILTER_NEAREST
    front_face: int = FRONT_CCW
    cull_mode: int = CULL_NONE
    depth: Optional[Tuple[int, bool]] = None  # (compare, write)
    blend: int = BLEND_REPLACE
    viewport: Optional[Tuple[float, float, float, float, float, float]] = None
    scissor: Optional[Tuple[int, int, int, int]] = None
    first: int = 0
    count: int = 0
    base_vertex: int = 0
    first_instance: int = 0
    instance_count: int = 1
    wgsl: Optional[str] = None    # WGSL source replacing the built-in program (same interface)
    topology: int = 3             # wgpu::PrimitiveTopology: 0 points, 1 lines, 2 line strip, 3 triangles, 4 strip
    strip_index_format: int = 0   # 0: none; 2 / 4: the index MAX value restarts strips
    # a second vertex buffer stepped per instance: (bytes, stride, [(format, offset, location)])
    instance_buffer: Optional[Tuple[bytes, int, List[Tuple[int, int, int]]]] = None
    # fragment test programs (with PROGRAM_CLIP_COLOR's vertex stage): FS_* bits and the
    # @early_depth_test mode (0 none, 1 force, 2 allow); WGSL text from fragment_test_wgsl
    fs_flags: int = 0
    fs_early: int = 0

    @property
    def indexed(self) -> bool:
        return self.index_data is not None


@dataclass
class PassSpec:
    width: int
    height: int
    color_format: int = RGBA8_UNORM
    has_color: bool = True
    clear_color: Optional[Tuple[float, float, float, float]] = (0.0, 0.0, 0.0, 1.0)  # None: LoadOp::Load
    color_init: Optional[bytes] = None
    has_depth: bool = False
    clear_depth: Optional[float] = 1.0
    depth_init: Optional[bytes] = None
    draws: List[DrawSpec] = field(default_factory=list)
    semantics: int = SEMANTICS_REFERENCE
    name: str = ""


# ------------------------------------------------------------------
# PRNG
# ------------------------------------------------------------------

def _mix(z: np.ndarray) -> np.ndarray:
    z = z.astype(np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def splitmix_u64(seed: int, n: int) -> np.ndarray:
    """The first n outputs of SplitMix64 started at `seed`."""
    with np.errstate(over="ignore"):
        ctr = np.uint64(seed) + GOLDEN * np.arange(1, n + 1, dtype=np.uint64)
    return _mix(ctr)


def splitmix_f32(seed: int, n: int) -> np.ndarray:
    u = splitmix_u64(seed, n)
    return ((u >> np.uint64(40)).astype(np.float64) * 2.0 ** -24).astype(np.float32)


def sub_seed(seed: int, index: int) -> int:
    return int(_mix(np.array([seed ^ index], dtype=np.uint64))[0])


# ------------------------------------------------------------------
# geometry
# ------------------------------------------------------------------

def perspective_view(fov_y_deg: float, aspect: float, near: float, far: float, eye_z: float) -> bytes:
    """proj * view for a camera at (0, 0, eye_z) looking at the origin, +y up,
    WebGPU depth range [0, 1]; column-major f32 (64 bytes)."""
    f = 1.0 / math.tan(math.radians(fov_y_deg) / 2.0)
    proj = np.array([[f / aspect, 0, 0, 0],
                     [0, f, 0, 0],
                     [0, 0, far / (near - far), near * far / (near - far)],
                     [0, 0, -1, 0]], dtype=np.float64)
    view = np.eye(4)
    view[2, 3] = -eye_z
    m = (proj @ view).astype(np.float32)
    return m.T.copy().tobytes()   # column-major


def uv_sphere(segments: int, rings: int):
    """Positions (N,3) f64, uv (N,2), u32 indices of a UV sphere; seam and
    poles duplicated, pole bands as single fans; outward faces wind
    counter-clockwise when seen from outside."""
    r = np.arange(rings + 1, dtype=np.float64)
    s = np.arange(segments + 1, dtype=np.float64)
    theta = math.pi * r / rings
    phi = 2.0 * math.pi * s / segments
    st, ct = np.sin(theta)[:, None], np.cos(theta)[:, None]
    pos = np.stack([st * np.cos(phi)[None, :], np.broadcast_to(ct, (rings + 1, segments + 1)),
                    -st * np.sin(phi)[None, :]], axis=-1).reshape(-1, 3)
    uv = np.stack(np.meshgrid(s / segments, r / rings, indexing="xy"), axis=-1).reshape(-1, 2)
    W = segments + 1
    ri, si = np.meshgrid(np.arange(rings), np.arange(segments), indexing="ij")
    a = ri * W + si
    b = a + 1
    c = (ri + 1) * W + si
    dd = c + 1
    # per quad, in order: the upper triangle (absent on the top band), then
    # the lower one (absent on the bottom band)
    up = np.stack([a, c, b], axis=-1)
    lo = np.stack([b, c, dd], axis=-1)
    has_up = (ri != 0)[..., None]
    has_lo = (ri != rings - 1)[..., None]
    tri = np.concatenate([up[..., None, :], lo[..., None, :]], axis=-2)   # (R, S, 2, 3)
    keep = np.concatenate([has_up, has_lo], axis=-1)                      # (R, S, 2)
    idx = tri[np.broadcast_to(keep, tri.shape[:-1])]
    return pos, uv, idx.astype(np.uint32).reshape(-1)


def _value_noise(p: np.ndarray, seed: int, lattice: float = 8.0) -> np.ndarray:
    """3D value noise in [0,1): trilinear with smoothstep weights over an
    integer lattice of `lattice` cells per unit; lattice values hash the
    packed integer coordinates with SplitMix64."""
    q = p * lattice
    i0 = np.floor(q).astype(np.int64)
    f = q - i0
    w = f * f * (3.0 - 2.0 * f)

    def val(ix, iy, iz):
        key = ((ix & 0x1FFFFF).astype(np.uint64) << np.uint64(42)) | \
              ((iy & 0x1FFFFF).astype(np.uint64) << np.uint64(21)) | (iz & 0x1FFFFF).astype(np.uint64)
        with np.errstate(over="ignore"):
            h = _mix(key ^ np.uint64(seed))
        return (h >> np.uint64(40)).astype(np.float64) * 2.0 ** -24

    out = np.zeros(len(p))
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                wx = w[:, 0] if dx else 1 - w[:, 0]
                wy = w[:, 1] if dy else 1 - w[:, 1]
                wz = w[:, 2] if dz else 1 - w[:, 2]
                out += wx * wy * wz * val(i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz)
    return out


def checker_texture(size: int, cell: int, noise: int, seed: int) -> bytes:
    """RGBA8 checker (255 / 0) of `cell`-pixel cells plus uniform integer
    noise in [-noise, noise] per channel, clamped; alpha 255."""
    y, x = np.mgrid[0:size, 0:size]
    base = (((x // cell) + (y // cell)) % 2 == 0).astype(np.int32) * 255
    rnd = splitmix_u64(seed, size * size * 3).reshape(size, size, 3)
    nz = (rnd % np.uint64(2 * noise + 1)).astype(np.int32) - noise
    rgb = np.clip(base[..., None] + nz, 0, 255).astype(np.uint8)
    rgba = np.concatenate([rgb, np.full((size, size, 1), 255, np.uint8)], axis=-1)
    return rgba.tobytes()


# ------------------------------------------------------------------
# workloads
# ------------------------------------------------------------------

def config1_plain_small(seed: int = SEED_BASE + 1) -> PassSpec:
    """UV sphere 32x16 (960 triangles, u32 indices), r = 1, per-vertex colours
    uniform in [0,1), camera fov 60 at distance 3, 256x256 Rgba8Unorm,
    Depth32Float Less, back-face culling, flat clear colour."""
    pos, _, idx = uv_sphere(32, 16)
    col = splitmix_f32(seed, len(pos) * 3).reshape(-1, 3)
    vb = np.zeros((len(pos), 7), np.float32)
    vb[:, 0:3] = pos
    vb[:, 3] = 1.0
    vb[:, 4:7] = col
    draw = DrawSpec(program=PROGRAM_MVP_COLOR, vertex_data=vb.tobytes(), vertex_stride=28,
                    attributes=[(FLOAT32X4, 0, 0), (FLOAT32X3, 16, 1)], index_data=idx.tobytes(),
                    uniform=perspective_view(60.0, 1.0, 0.1, 10.0, 3.0), front_face=FRONT_CCW,
                    cull_mode=CULL_BACK, depth=(LESS, True), count=len(idx))
    return PassSpec(256, 256, RGBA8_UNORM, clear_color=(0.1, 0.2, 0.3, 1.0), has_depth=True,
                    clear_depth=1.0, draws=[draw], name="c1_plain_small")


def displaced_sphere(segments: int, rings: int, seed: int, amplitude: float = 0.05):
    pos, uv, idx = uv_sphere(segments, rings)
    n = _value_noise(pos, seed)
    pos = pos * (1.0 + amplitude * (2.0 * n - 1.0))[:, None]
    return pos, uv, idx


def config2_textured(seed: int = SEED_BASE + 2, extended: bool = False, width: int = 1920,
                     height: int = 1080, segments: int = 256, rings: int = 128,
                     texture: Optional[bytes] = None) -> PassSpec:
    """Sphere 256x128 (65,024 triangles) radially displaced by seeded value
    noise (amplitude 0.05), f32x2 UVs; 1024x1024 Rgba8UnormSrgb checker (32 px
    cells) plus +-16 noise, Repeat addressing; 1920x1080 Rgba8Unorm target,
    Depth32Float Less.  `extended` selects perspective-correct varyings,
    bilinear filtering and sRGB decode (the reference supports none of them);
    otherwise linear varyings and the reference's nearest sampling."""
    pos, uv, idx = displaced_sphere(segments, rings, sub_seed(seed, 1))
    vb = np.zeros((len(pos), 6), np.float32)
    vb[:, 0:3] = pos
    vb[:, 3] = 1.0
    vb[:, 4:6] = uv
    tex = texture if texture is not None else checker_texture(1024, 32, 16, sub_seed(seed, 2))
    draw = DrawSpec(program=PROGRAM_MVP_TEXTURE,
                    interp=INTERP_PERSPECTIVE if extended else INTERP_LINEAR,
                    vertex_data=vb.tobytes(), vertex_stride=24,
                    attributes=[(FLOAT32X4, 0, 0), (FLOAT32X2, 16, 1)], index_data=idx.tobytes(),
                    uniform=perspective_view(60.0, width / height, 0.1, 10.0, 3.0),
                    texture=(1024, 1024, tex, True), address_u=REPEAT, address_v=REPEAT,
                    filter=FILTER_LINEAR if extended else FILTER_NEAREST,
                    front_face=FRONT_CCW, cull_mode=CULL_BACK, depth=(LESS, True), count=len(idx))
    return PassSpec(width, height, RGBA8_UNORM, clear_color=(0.0, 0.0, 0.0, 1.0), has_depth=True,
                    clear_depth=1.0, draws=[draw],
                    semantics=SEMANTICS_EXTENDED if extended else SEMANTICS_REFERENCE,
                    name="c2_textured_" + ("ext" if extended else "ref"))


def perspective_quad(width: int = 64, height: int = 8) -> PassSpec:
    """Full-target quad in clip space, left edge w = 1, right edge w = 4
    (x = +-w, y = +-w, z = w/2); red varies 0 -> 1 from left to right with
    perspective-correct interpolation (extension)."""
    L, R = 1.0, 4.0
    verts = np.array([[-L, -L, 0.5 * L, L], [R, -R, 0.5 * R, R], [R, R, 0.5 * R, R],
                      [-L, -L, 0.5 * L, L], [R, R, 0.5 * R, R], [-L, L, 0.5 * L, L]], np.float32)
    cl, cr = [0, 0, 0, 1], [1, 0, 0, 1]
    cols = np.array([cl, cr, cr, cl, cr, cl], np.float32)
    vb = np.concatenate([verts, cols], axis=1)
    d = DrawSpec(program=PROGRAM_CLIP_COLOR, interp=INTERP_PERSPECTIVE, vertex_data=vb.tobytes(),
                 vertex_stride=32, attributes=[(FLOAT32X4, 0, 0), (FLOAT32X4, 16, 1)], count=6)
    return PassSpec(width, height, RGBA8_UNORM, clear_color=(0, 0, 0, 0), draws=[d],
                    semantics=SEMANTICS_EXTENDED, name="perspective_quad")


def bilinear_strip(width: int = 32, height: int = 4) -> PassSpec:
    """Screen-aligned quad with u = 0 -> 1 left to right over a 2x2 texture
    whose columns hold red 200 and 40, bilinear + Repeat (extension)."""
    texels = np.array([[[200, 0, 0, 255], [40, 0, 0, 255]],
                       [[200, 0, 0, 255], [40, 0, 0, 255]]], np.uint8)
    verts = np.array([[-1, -1, 0.5, 1], [1, -1, 0.5, 1], [1, 1, 0.5, 1],
                      [-1, -1, 0.5, 1], [1, 1, 0.5, 1], [-1, 1, 0.5, 1]], np.float32)
    uv = np.array([[0, 1], [1, 1], [1, 0], [0, 1], [1, 0], [0, 0]], np.float32)
    vb = np.concatenate([verts, uv], axis=1)
    ident = np.eye(4, dtype=np.float32).T.tobytes()
    d = DrawSpec(program=PROGRAM_MVP_TEXTURE, interp=INTERP_PERSPECTIVE, vertex_data=vb.tobytes(),
                 vertex_stride=24, attributes=[(FLOAT32X4, 0, 0), (FLOAT32X2, 16, 1)], count=6,
                 uniform=ident, texture=(2, 2, texels.tobytes(), False), address_u=REPEAT, address_v=REPEAT,
                 filter=FILTER_LINEAR)
    return PassSpec(width, height, RGBA8_UNORM, clear_color=(0, 0, 0, 0), draws=[d],
                    semantics=SEMANTICS_EXTENDED, name="bilinear_strip")


def addressing_quad(address_u: int, address_v: int, extended: bool = False, bilinear: bool = False,
                    width: int = 96, height: int = 64, seed: int = SEED_BASE + 21) -> PassSpec:
    """Screen-filling quad (two triangles, identity camera) whose uv runs
    from -1.3 to 2.7 across and -0.8 to 1.9 down, over a seeded 8x8 RGBA
    texture: every texel coordinate outside [0, 1) goes through the
    sampler's address mode (binding.rs:150-164) on both axes."""
    rnd = splitmix_u64(seed, 8 * 8 * 4)
    tex = (rnd & np.uint64(0xFF)).astype(np.uint8).reshape(8, 8, 4)
    tex[..., 3] = 255
    u0, u1, v0, v1 = -1.3, 2.7, -0.8, 1.9
    verts = np.array([[-1, -1, 0.5, 1], [1, -1, 0.5, 1], [1, 1, 0.5, 1],
                      [-1, -1, 0.5, 1], [1, 1, 0.5, 1], [-1, 1, 0.5, 1]], np.float32)
    uv = np.array([[u0, v1], [u1, v1], [u1, v0], [u0, v1], [u1, v0], [u0, v0]], np.float32)
    vb = np.concatenate([verts, uv], axis=1)
    ident = np.eye(4, dtype=np.float32).T.tobytes()
    d = DrawSpec(program=PROGRAM_MVP_TEXTURE, interp=INTERP_PERSPECTIVE if extended else INTERP_LINEAR,
                 vertex_data=vb.tobytes(), vertex_stride=24, attributes=[(FLOAT32X4, 0, 0), (FLOAT32X2, 16, 1)],
                 count=6, uniform=ident, texture=(8, 8, tex.tobytes(), False), address_u=address_u,
                 address_v=address_v, filter=FILTER_LINEAR if bilinear else FILTER_NEAREST)
    return PassSpec(width, height, RGBA8_UNORM, clear_color=(0, 0, 0, 0), draws=[d],
                    semantics=SEMANTICS_EXTENDED if extended else SEMANTICS_REFERENCE,
                    name=f"addressing u={address_u} v={address_v} ext={extended} bilinear={bilinear}")


def instanced_triangles(n: int = 300, instances: int = 4, first_instance: int = 0, width: int = 160,
                        height: int = 120, seed: int = SEED_BASE + 22, depth: Optional[Tuple[int, bool]] = None,
                        blend: bool = False) -> PassSpec:
    """n clip-space triangles (positions stepped per vertex, constant z)
    drawn `instances` times; the colour (location 1) comes from a second
    buffer stepped per instance (vertex.rs:191-195), so every instance
    covers the same pixels with its own colour: with Less the first
    instance wins, with LessEqual the last, with src-over all of them in
    order."""
    u = splitmix_f32(seed, n * 6).astype(np.float64).reshape(n, 6)
    cx, cy = u[:, 0] * 1.8 - 0.9, u[:, 1] * 1.8 - 0.9
    size = 0.05 + 0.25 * u[:, 2]
    ang = u[:, 3:4] * 2 * np.pi + np.array([0.0, 2.1, 4.2])[None, :]
    vb = np.zeros((n, 3, 4), np.float32)
    vb[:, :, 0] = cx[:, None] + size[:, None] * np.cos(ang)
    vb[:, :, 1] = cy[:, None] + size[:, None] * np.sin(ang)
    vb[:, :, 2] = 0.25 + 0.5 * u[:, 4:5]
    vb[:, :, 3] = 1.0
    total = first_instance + instances
    cols = splitmix_f32(seed ^ 0x77, total * 4).reshape(total, 4)
    cols[:, 3] = 0.2 + 0.6 * cols[:, 3]
    d = DrawSpec(program=PROGRAM_CLIP_COLOR, vertex_data=vb.tobytes(), vertex_stride=16,
                 attributes=[(FLOAT32X4, 0, 0)], count=3 * n, first_instance=first_instance,
                 instance_count=instances, instance_buffer=(cols.tobytes(), 16, [(FLOAT32X4, 0, 1)]),
                 depth=depth, blend=BLEND_ALPHA if blend else BLEND_REPLACE)
    return PassSpec(width, height, RGBA8_UNORM, clear_color=(0.1, 0.2, 0.3, 1.0), has_depth=depth is not None,
                    clear_depth=1.0, draws=[d], semantics=SEMANTICS_EXTENDED if blend else SEMANTICS_REFERENCE,
                    name=f"instanced first={first_instance} n={instances} depth={depth} blend={blend}")


def fragment_test_scene(flags: int, early: int, compare: Optional[int] = LESS, write: bool = True, n: int = 600,
                        width: int = 160, height: int = 120, seed: int = SEED_BASE + 23, topology: int = 3,
                        z_levels=(0.25, 0.5, 0.75), alpha=(0.1, 0.9), blend: bool = False,
                        cull_mode: int = CULL_NONE, instances: int = 1) -> PassSpec:
    """n overlapping clip-space triangles (or as many points / lines, by
    `topology`) with depths from `z_levels` (ties on purpose) and colour
    alpha uniform in `alpha`, drawn with the fragment test program of
    `flags` / `early` (fragment_test_wgsl) over a random initial depth
    buffer (LoadOp::Load), so that every compare function takes both
    branches.  `compare` None: no depth attachment."""
    u = splitmix_f32(seed, n * 8).astype(np.float64).reshape(n, 8)
    cx, cy = u[:, 0] * 2.2 - 1.1, u[:, 1] * 2.2 - 1.1
    size = 0.04 + 0.3 * u[:, 2]
    nv = {0: 1, 1: 2, 2: 2, 3: 3, 4: 3}[topology]
    ang = u[:, 3:4] * 2 * np.pi + np.array([0.0, 2.1, 4.2])[None, :nv]
    zq = np.asarray(z_levels)[(u[:, 4] * len(z_levels)).astype(int) % len(z_levels)]
    vb = np.zeros((n, nv, 8), np.float32)
    vb[:, :, 0] = cx[:, None] + size[:, None] * np.cos(ang)
    vb[:, :, 1] = cy[:, None] + size[:, None] * np.sin(ang)
    vb[:, :, 2] = zq[:, None]
    vb[:, :, 3] = 1.0
    col = splitmix_f32(seed ^ 0x5A5A, n * 4).reshape(n, 1, 4)
    vb[:, :, 4:8] = col
    vb[:, :, 7] = alpha[0] + col[:, :, 3] * (alpha[1] - alpha[0])
    d = DrawSpec(program=PROGRAM_CLIP_COLOR, vertex_data=vb.tobytes(), vertex_stride=32,
                 attributes=[(FLOAT32X4, 0, 0), (FLOAT32X4, 16, 1)], count=n * nv, topology=topology,
                 depth=None if compare is None else (compare, write),
                 blend=BLEND_ALPHA if blend else BLEND_REPLACE, fs_flags=flags, fs_early=early,
                 cull_mode=cull_mode, instance_count=instances)
    d.wgsl = fragment_test_wgsl(flags, early)
    init = np.where(splitmix_f32(seed ^ 0x1234, width * height) < 0.5, 0.5, 0.625).astype(np.float32)
    return PassSpec(width, height, RGBA8_UNORM, clear_color=(0.1, 0.2, 0.3, 1.0),
                    has_depth=compare is not None, clear_depth=None, depth_init=init.tobytes(), draws=[d],
                    semantics=SEMANTICS_EXTENDED if blend else SEMANTICS_REFERENCE,
                    name=f"fragment test flags={flags} early={early} compare={compare} write={write} "
                         f"topology={topology} blend={blend}")


def config3_overdraw(seed: int = SEED_BASE + 3, blend: bool = False, n: int = 200_000,
                     width: int = 2560, height: int = 2440) -> PassSpec:
    """n screen-space equilateral triangles, centres uniform over the target,
    areas log-normal (median 40 px, sigma 1.0) clamped to [4, 400] px, random
    orientation; f32x4 colours with alpha uniform in [0.1, 0.9]; no depth, no
    culling, non-indexed.  `blend` selects src-over in API order (extension);
    otherwise the reference's overwrite."""
    u = splitmix_f32(seed, n * 9).astype(np.float64).reshape(n, 9)
    cx = u[:, 0] * width
    cy = u[:, 1] * height
    # Box-Muller normal for the log-normal area
    r = np.sqrt(-2.0 * np.log(np.maximum(u[:, 2], 2.0 ** -24)))
    z = r * np.cos(2.0 * math.pi * u[:, 3])
    area = np.clip(40.0 * np.exp(1.0 * z), 4.0, 400.0)
    side = np.sqrt(4.0 * area / math.sqrt(3.0))
    circ = side / math.sqrt(3.0)
    rot = u[:, 4] * 2.0 * math.pi
    ang = rot[:, None] + np.array([0.0, 2.0, 4.0])[None, :] * math.pi / 3.0
    px = cx[:, None] + circ[:, None] * np.cos(ang)
    py = cy[:, None] + circ[:, None] * np.sin(ang)
    ndc_x = px / width * 2.0 - 1.0
    ndc_y = 1.0 - py / height * 2.0
    vb = np.zeros((n, 3, 8), np.float32)
    vb[:, :, 0] = ndc_x
    vb[:, :, 1] = ndc_y
    vb[:, :, 2] = 0.5
    vb[:, :, 3] = 1.0
    vb[:, :, 4] = u[:, 5:6]
    vb[:, :, 5] = u[:, 6:7]
    vb[:, :, 6] = u[:, 7:8]
    vb[:, :, 7] = (0.1 + 0.8 * u[:, 8:9])
    draw = DrawSpec(program=PROGRAM_CLIP_COLOR, vertex_data=vb.tobytes(), vertex_stride=32,
                    attributes=[(FLOAT32X4, 0, 0), (FLOAT32X4, 16, 1)], cull_mode=CULL_NONE,
                    blend=BLEND_ALPHA if blend else BLEND_REPLACE, count=3 * n)
    return PassSpec(width, height, RGBA8_UNORM, clear_color=(0.0, 0.0, 0.0, 1.0), has_depth=False,
                    draws=[draw], semantics=SEMANTICS_EXTENDED if blend else SEMANTICS_REFERENCE,
                    name="c3_overdraw_" + ("blend" if blend else "ref"))


def colored_triangle(front_face: int, cull_mode: int, vertices: int = 3, topology: int = 3) -> PassSpec:
    """wgpu-cpu-tests/src/tests/colored_triangle.rs:11-138: 512x512
    Rgba8UnormSrgb, clear black, Depth32Float Less (clear 1.0), the
    vertex_index-only shader.  colored_triangle_lines (:188-198) is
    LineStrip over vertices 0..4."""
    draw = DrawSpec(program=PROGRAM_COLORED_TRIANGLE, front_face=front_face, cull_mode=cull_mode,
                    depth=(LESS, True), count=vertices, topology=topology)
    return PassSpec(512, 512, RGBA8_UNORM_SRGB, clear_color=(0.0, 0.0, 0.0, 1.0), has_depth=True,
                    clear_depth=1.0, draws=[draw], name="colored_triangle")


def batched_sizes(seed: int, n: int = 256):
    """Per pass of the batched workload: (width, height, triangles, textured),
    W and H log-uniform in [64, 1024], triangles log-uniform in [100, 50k],
    half of the passes textured."""
    u = splitmix_f32(seed, n * 4).astype(np.float64).reshape(n, 4)
    w = np.round(np.exp(np.log(64) + u[:, 0] * (np.log(1024) - np.log(64)))).astype(int)
    h = np.round(np.exp(np.log(64) + u[:, 1] * (np.log(1024) - np.log(64)))).astype(int)
    tris = np.exp(np.log(100) + u[:, 2] * (np.log(50_000) - np.log(100)))
    textured = u[:, 3] < 0.5
    return [(int(w[i]), int(h[i]), float(tris[i]), bool(textured[i])) for i in range(n)]


def sphere_for_triangles(tris: float) -> Tuple[int, int]:
    """Segments = 2 * rings with 2*segments*(rings-1) ~= tris."""
    rings = max(3, int(round(math.sqrt(tris / 4.0))))
    return 2 * rings, rings


_SHARED_TEXTURE = {}


def shared_texture() -> bytes:
    """The textured workload's 1024x1024 checker, generated once and shared
    by every textured pass of the batched and serving workloads."""
    if "tex" not in _SHARED_TEXTURE:
        _SHARED_TEXTURE["tex"] = checker_texture(1024, 32, 16, sub_seed(SEED_BASE + 2, 2))
    return _SHARED_TEXTURE["tex"]


def batched_pass(seed: int, index: int, width: int, height: int, tris: float, textured: bool) -> PassSpec:
    segs, rings = sphere_for_triangles(tris)
    s = sub_seed(seed, index)
    if textured:
        spec = config2_textured(seed=s, width=width, height=height, segments=segs, rings=rings,
                                texture=shared_texture())
    else:
        pos, _, idx = displaced_sphere(segs, rings, sub_seed(s, 1))
        col = splitmix_f32(sub_seed(s, 3), len(pos) * 3).reshape(-1, 3)
        vb = np.zeros((len(pos), 7), np.float32)
        vb[:, 0:3] = pos
        vb[:, 3] = 1.0
        vb[:, 4:7] = col
        draw = DrawSpec(program=PROGRAM_MVP_COLOR, vertex_data=vb.tobytes(), vertex_stride=28,
                        attributes=[(FLOAT32X4, 0, 0), (FLOAT32X3, 16, 1)], index_data=idx.tobytes(),
                        uniform=perspective_view(60.0, width / height, 0.1, 10.0, 3.0),
                        front_face=FRONT_CCW, cull_mode=CULL_BACK, depth=(LESS, True), count=len(idx))
        spec = PassSpec(width, height, RGBA8_UNORM, clear_color=(0.1, 0.2, 0.3, 1.0), has_depth=True,
                        clear_depth=1.0, draws=[draw])
    spec.name = f"batch_{index}"
    return spec


def algorithmic_bytes(spec: PassSpec, count_texture: bool = True) -> int:
    """Bytes a pass must move at minimum: final colour and depth stores, the
    vertex and index reads, the 64 B uniform and the texture footprint."""
    b = spec.width * spec.height * ((4 if spec.has_color else 0) + (4 if spec.has_depth else 0))
    for d in spec.draws:
        b += len(d.vertex_data) + (len(d.index_data) if d.index_data else 0)
        b += 64 if d.uniform else 0
        if count_texture and d.texture:
            b += len(d.texture[2])
    return b
