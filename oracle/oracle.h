/*
 * oracle.h -- CPU restatement of wgpu-cpu's render-pass executor.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity oracle: the tests,
 * the smoke check and the CPU-baseline leg of the benchmark call it to check
 * and time the HIP path.  Nothing in the product (libwgc) links, loads or
 * calls it.
 *
 * Every function follows a file of the reference (paths relative to the
 * reference checkout, jgraef/wgpu-cpu):
 *   wgpu-cpu/src/render_pass/state.rs:490-593   draw loop (VS per index, cull)
 *   wgpu-cpu/src/render_pass/index.rs:45-106    index resolution, restart separators
 *   wgpu-cpu/src/render_pass/primitive.rs:206-487  list / strip assembly (TriStripIter,
 *                                                  ResetTriStripIter, line windows)
 *   wgpu-cpu/src/render_pass/clipper.rs:207-407  Cohen-Sutherland line clipper
 *   wgpu-cpu/src/render_pass/raster.rs:184-265   point and line rasterizers
 *   wgpu-cpu/src/util/bresenham.rs:78-153        Bresenham iterator (end exclusive)
 *   wgpu-cpu/src/render_pass/vertex.rs:128-196  vertex attribute fetch
 *   wgpu-cpu/src/render_pass/primitive.rs:159-197  TriFace / winding
 *   wgpu-cpu/src/render_pass/clipper.rs:150-179,625-802  6-plane triangle clipper
 *   wgpu-cpu/src/render_pass/raster.rs:129-160,280-362  to_raster, TriRasterizer
 *   wgpu-cpu/src/util/scanline.rs:14-198         integer scanline spans
 *   wgpu-cpu/src/util/sort.rs:3-22               bubblesort3
 *   wgpu-cpu/src/util/interpolation.rs:84-89,284-297  lerp, Barycentric::interpolate
 *   wgpu-cpu/src/render_pass/fragment.rs:127-214,319-377,427-488  FS, depth, write
 *   wgpu-cpu/src/render_pass/binding.rs:93-164   textureSample (nearest, rem_euclid)
 *   wgpu-cpu/src/texture.rs:170-239,373-411      texel read/write, clears
 *   wgpu-cpu/src/render_pass/mod.rs:434-452      compare functions
 *   naga-cranelift/src/expression/binary.rs:270-341  mat4x4 * vec4 column sum
 *
 * Shader programs: the reference JIT-compiles WGSL.  The oracle carries the
 * handful of WGSL programs the parity scenes use as C functions that perform
 * the same IEEE-754 binary32 operations in the same order as the reference's
 * Cranelift lowering (no fused multiply-add, see the build flags).
 *
 * Extensions (no reference counterpart; the reference has todo!() or ignores
 * the setting): perspective-correct varyings, bilinear filtering, sRGB decode
 * and src-over blending.  They are off unless a draw asks for them and are
 * pinned by closed-form tests only.
 */
#ifndef WGC_ORACLE_H
#define WGC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAX_VB 4
#define OR_MAX_ATTR 8
#define OR_MAX_VAR 16

/* colour target formats (texture.rs:373-411) */
enum { OR_FMT_RGBA8UNORM = 0, OR_FMT_RGBA8UNORM_SRGB = 1, OR_FMT_BGRA8UNORM = 2,
       OR_FMT_BGRA8UNORM_SRGB = 3 };
/* vertex attribute formats: value = number of f32 components */
enum { OR_VFMT_F32 = 1, OR_VFMT_F32X2 = 2, OR_VFMT_F32X3 = 3, OR_VFMT_F32X4 = 4 };
/* built-in WGSL programs (the WGSL text of each lives in the package's shaders/) */
enum { OR_PROG_MVP_COLOR = 1,        /* hello_mesh.wgsl-style: M*pos, colour varying */
       OR_PROG_MVP_TEXTURE = 2,      /* hello_texture.wgsl-style: M*pos, uv, textureSample */
       OR_PROG_COLORED_TRIANGLE = 3, /* colored_triangle.wgsl: vertex_index only */
       OR_PROG_CLIP_COLOR = 4 };     /* clip-space position + colour pass-through */
/* wgpu::CompareFunction discriminants (mod.rs:434-452) */
enum { OR_CMP_NEVER = 1, OR_CMP_LESS = 2, OR_CMP_EQUAL = 3, OR_CMP_LESS_EQUAL = 4,
       OR_CMP_GREATER = 5, OR_CMP_NOT_EQUAL = 6, OR_CMP_GREATER_EQUAL = 7, OR_CMP_ALWAYS = 8 };
enum { OR_ADDR_CLAMP = 0, OR_ADDR_REPEAT = 1, OR_ADDR_MIRROR = 2 };
enum { OR_INTERP_LINEAR = 0, OR_INTERP_PERSPECTIVE = 1, OR_INTERP_FLAT = 2 };
enum { OR_CULL_NONE = 0, OR_CULL_FRONT = 1, OR_CULL_BACK = 2 };
/* wgpu::PrimitiveTopology discriminants (state.rs:434-455) */
enum { OR_TOPO_POINT_LIST = 0, OR_TOPO_LINE_LIST = 1, OR_TOPO_LINE_STRIP = 2, OR_TOPO_TRIANGLE_LIST = 3,
       OR_TOPO_TRIANGLE_STRIP = 4 };

typedef struct {
    const uint8_t *data;   /* buffer slice, offset already applied */
    uint64_t size;
    uint32_t stride;
    uint32_t step_instance; /* 0 = VertexStepMode::Vertex, 1 = Instance */
} OrVertexBuffer;

typedef struct {
    uint32_t buffer;
    uint32_t offset;
    uint32_t format;       /* OR_VFMT_* */
} OrAttribute;

typedef struct {
    /* pipeline */
    uint32_t program;
    uint32_t front_face_cw;  /* 0: FrontFace::Ccw, 1: FrontFace::Cw */
    uint32_t cull_mode;      /* OR_CULL_* */
    uint32_t depth_enabled;  /* pipeline carries a DepthStencilState */
    uint32_t depth_compare;  /* OR_CMP_* */
    uint32_t depth_write;
    uint32_t interp;         /* OR_INTERP_* for the colour / uv varying */
    uint32_t blend;          /* 0 replace (reference), 1 src-over (extension) */
    uint32_t bilinear;       /* 0 reference nearest, 1 bilinear (extension) */
    uint32_t srgb_decode;    /* 0 reference (no decode), 1 decode sRGB textures (extension) */
    uint32_t n_attr;
    OrAttribute attr[OR_MAX_ATTR];   /* indexed by shader location */
    uint32_t n_vb;
    OrVertexBuffer vb[OR_MAX_VB];
    /* index buffer */
    const uint8_t *index_data;
    uint64_t index_size;
    uint32_t index_format;   /* bytes per index: 2 or 4 */
    /* bind group resources */
    const float *uniform;    /* mat4x4f, column major, 64 bytes */
    const uint8_t *texture;  /* RGBA8, rows of tex_w*4 bytes */
    uint32_t tex_w, tex_h, tex_srgb;
    uint32_t addr_u, addr_v; /* OR_ADDR_* */
    /* dynamic state */
    float viewport[6];       /* x, y, w, h, min_depth, max_depth */
    uint32_t scissor[4];     /* x, y, w, h */
    /* the draw call */
    uint32_t indexed;
    uint32_t first, count;   /* vertices or indices */
    int32_t base_vertex;
    uint32_t first_instance, instance_count;
    /* primitive state */
    uint32_t topology;       /* OR_TOPO_* */
    uint32_t strip_restart;  /* strip_index_format set: index MAX separates strips */
    /* fragment test programs: OR_PROG_CLIP_COLOR's vertex stage with a
     * fragment entry point that reads the colour varying c and
     * @builtin(primitive_index), built from the OR_FS_* bits below (the
     * tests hold the WGSL text of each combination) */
    uint32_t fs_flags;
    uint32_t fs_early;       /* @early_depth_test: 0 none, 1 force, 2 allow (fragment.rs:166-194) */
} OrDraw;

/* OR_FS_DISCARD: discard when c.w < 0.5; OR_FS_DEPTH_ALPHA / _CONST: write
 * @builtin(frag_depth) = c.w / 0.3125; OR_FS_DEPTH_FIRST: that output is a
 * member before the @location(0) one (so the late depth test sees it);
 * OR_FS_PRIM_RED: red = f32(primitive_index % 8u) * 0.125. */
enum { OR_FS_DISCARD = 1, OR_FS_DEPTH_ALPHA = 2, OR_FS_DEPTH_CONST = 4, OR_FS_DEPTH_FIRST = 8,
       OR_FS_PRIM_RED = 16 };

typedef struct {
    uint32_t width, height;
    uint8_t *color;          /* may be NULL */
    uint32_t color_format;
    uint32_t color_clear;    /* LoadOp::Clear */
    double clear_color[4];
    float *depth;            /* may be NULL */
    uint32_t depth_clear;
    float clear_depth;
    uint32_t n_draws;
    const OrDraw *draws;
    uint32_t *frag_count;    /* optional: per-pixel count of emitted fragments */
    uint64_t fragments;      /* out */
    int error;               /* out: 0 ok; else the reference would have panicked */
    char message[256];
} OrPass;

int oracle_execute_pass(OrPass *pass);

/* ---- pieces exposed for known-answer tests ---- */

/* util/sort.rs:3-22, compare on ints */
void oracle_bubblesort3_i64(int64_t v[3]);

/* util/scanline.rs:14-198: enumerate pixels of a triangle with integer
 * vertices; writes up to cap (x, y) pairs, returns the total count.  Returns
 * -1 where the reference panics (negative coordinate). */
int64_t oracle_scanline_pixels(const uint32_t tri[6], uint32_t *out_xy, int64_t cap);

/* clipper.rs:766-802: clip one triangle; out receives up to cap triangles of
 * 3 x (x,y,z,w, b0,b1,b2) floats; returns the number of triangles. */
int oracle_clip_triangle(const float in_xyzw[12], float *out, int cap);

/* raster.rs:129-160 */
void oracle_to_raster(const float viewport[6], const float clip[4], uint32_t fb[2], float frag[4]);

/* texture.rs:376-379 */
uint8_t oracle_f32_to_unorm8(float v);

/* binding.rs:150-164 */
uint32_t oracle_texel_coordinate(float x, uint32_t address_mode, uint32_t size);

/* shared sRGB decode table (IEC 61966-2-1), used by the extension path */
const float *oracle_srgb_table(void);

/* naga-cranelift binary.rs:270-341: column-major mat4x4f * vec4f */
void oracle_mat4_mul_vec4(const float m[16], const float v[4], float r[4]);

/* util/bresenham.rs:78-153: the points from start (inclusive) to end
 * (exclusive) and their t; returns the number of points (at most cap written) */
int64_t oracle_bresenham(const uint32_t start[2], const uint32_t end[2], uint32_t *out_xy, float *out_t,
                         int64_t cap);

/* clipper.rs:266-366 (Cohen-Sutherland): returns 1 and the clipped points
 * and their alphas, or 0 when the line is outside */
int oracle_clip_line(const float a[4], const float b[4], float out[8], float alphas[2]);

/* primitive.rs:206-487 with index.rs:72-106: the index positions (relative to
 * the first index) of every primitive of `count` indices, `n` per primitive
 * (1, 2, 3) in assembly order.  `indices` (NULL: non-indexed) are raw values
 * of `index_format` bytes; with `restart` the maximum value separates strips.
 * Returns the number of primitives (at most cap written). */
int64_t oracle_assemble(uint32_t topology, uint32_t restart, uint32_t index_format, const uint8_t *indices,
                        uint32_t count, uint32_t *out_positions, int64_t cap);

#ifdef __cplusplus
}
#endif
#endif
