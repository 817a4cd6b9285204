// wgc_core.h -- device-side core of the MI355X render-pass executor.
//
// Everything here performs the same IEEE-754 binary32 operations, in the
// same order, as the reference's CPU executor, so that coverage, depth and
// colour come out bit-identical.  The translation unit must be compiled
// with -ffp-contract=off (no fused multiply-add: the reference never fuses,
// naga-cranelift/src/expression/binary.rs:299-321) and correctly rounded
// f32 division.
#pragma once

#ifdef __HIPCC_RTC__
// hipRTC provides the HIP device API itself, and the fixed-width integer
// types in its own namespace.
using __hip_internal::int8_t;
using __hip_internal::int16_t;
using __hip_internal::int32_t;
using __hip_internal::int64_t;
using __hip_internal::uint8_t;
using __hip_internal::uint16_t;
using __hip_internal::uint32_t;
using __hip_internal::uint64_t;
#else
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

namespace wgc {

constexpr int kMaxVb = 4;
constexpr int kMaxAttr = 8;
constexpr int kMaxVar = 4;        // user varying floats of the built-in programs
constexpr int kMaxTileDim = 32;   // tiles are 8x8, 16x16 or 32x32 pixels
constexpr int kBatch = 64;        // triangles staged in LDS per raster step
constexpr int kMaxBinTiles = 16;  // larger footprints go to the big-triangle list
constexpr int kOrderedCap = 4096; // keys the ordered (blend) path sorts at once (a tile with more takes several chunks)
constexpr int kClipFloats = 24;   // one clipped sub-triangle: 3 x (xyzw + b0 b1 b2), padded

enum Program : uint32_t {
    PROG_MVP_COLOR = 1,
    PROG_MVP_TEXTURE = 2,
    PROG_COLORED_TRIANGLE = 3,
    PROG_CLIP_COLOR = 4,
    PROG_USER = 5,     // shaders translated from WGSL, compiled with the kernels at pipeline creation
};

enum Interp : uint32_t { INTERP_LINEAR = 0, INTERP_PERSPECTIVE = 1, INTERP_FLAT = 2 };
// wgpu::PrimitiveTopology (state.rs:434-455)
enum Topology : uint32_t {
    TOPO_POINT_LIST = 0, TOPO_LINE_LIST = 1, TOPO_LINE_STRIP = 2, TOPO_TRIANGLE_LIST = 3, TOPO_TRIANGLE_STRIP = 4,
};
constexpr uint32_t kStripChunk = 256;   // index positions per separator summary (restart strips)
constexpr uint32_t kLineSteps = 8192;   // Bresenham steps per line (the fragment key holds 13 bits)
constexpr uint32_t kLpPrimBytes = 192;   // scratch per primitive of an ordered point / line draw
enum Compare : uint32_t {
    CMP_NEVER = 1, CMP_LESS = 2, CMP_EQUAL = 3, CMP_LESS_EQUAL = 4,
    CMP_GREATER = 5, CMP_NOT_EQUAL = 6, CMP_GREATER_EQUAL = 7, CMP_ALWAYS = 8,
};
enum Addr : uint32_t { ADDR_CLAMP = 0, ADDR_REPEAT = 1, ADDR_MIRROR = 2 };

// How a pixel's sequence of fragments is resolved.  All but RES_ORDERED are
// order-independent restatements of the sequential depth test + overwrite:
// the winner is picked by (depth, primitive order) so that fragments of one
// tile may be processed in any order and still give the sequential result.
enum Resolve : uint32_t {
    RES_LAST = 0,        // no depth test: the last fragment in API order wins
    RES_MIN_FIRST = 1,   // Less + write: smallest depth, first in API order
    RES_MIN_LAST = 2,    // LessEqual + write: smallest depth, last in API order
    RES_MAX_FIRST = 3,   // Greater + write
    RES_MAX_LAST = 4,    // GreaterEqual + write
    RES_PASS_LAST = 5,   // compare against the depth before the draw; last passing wins
    RES_ORDERED = 6,     // blending / NotEqual + write: sorted, sequential per pixel
};

enum ErrorBits : uint32_t {
    ERR_INDEX = 1u << 0,          // index out of the bound index buffer
    ERR_VERTEX = 1u << 1,         // vertex fetch out of the bound vertex buffer
    ERR_W_ZERO = 1u << 2,         // clip position w == 0 (reference: expect("w=0"))
    ERR_CLIP_OVERFLOW = 1u << 3,  // clipped sub-triangle buffer full
    ERR_RECORD_OVERFLOW = 1u << 4,
    // 1u << 5 unused (the ordered path has no key limit)
    ERR_COORD_RANGE = 1u << 6,    // framebuffer coordinate >= 2^30
    ERR_PRIM_RANGE = 1u << 7,     // more than 2^25 primitives in a draw
    ERR_SHADER = 1u << 8,         // a shader invocation aborted (not discard): division by zero,
                                  // overflow, index out of bounds (reference: panic)
};

// Pipeline-level attribute description (shader location -> buffer slot).
struct AttrDev {
    uint32_t buffer, offset, ncomp, pad;
};

// Draw-level attribute: the bound buffer slice resolved on the host, so the
// kernels index nothing dynamically.
struct AttrBind {
    const uint8_t *data;  // buffer slice start
    uint64_t size;        // slice size
    uint32_t stride, step_instance, offset, ncomp;
};

// Scratch of one draw.  The tile arrays, the bins, the spill list, the
// active-tile list, the clip list and the group header are slices of arrays
// shared by every draw of a launch group (tile indices in the spill and
// active lists are global: tile0 + tile of the draw); the rest is per draw.
//
// Bins: every tile owns bin_cap slots.  A small record (footprint of at most
// kMaxBinTiles tiles) takes its position in each tile of its footprint from
// the tile's count (one atomic add) and stores its key there; positions at
// or past bin_cap go to the spill list as (global tile, key).  The record
// that takes position 0 of a tile appends the tile to the active list.
struct Scratch {
    uint32_t *tile_count;    // n_tiles (this draw's slice): records binned to the tile
    uint32_t *bins;          // n_tiles x bin_cap keys (this draw's slice)
    uint2 *spill;            // shared: (global tile, key) past a full bin
    uint32_t *active_list;   // shared: global tiles with at least one binned record
    uint32_t *header;        // shared group header, see Header
    uint2 *clip_list;        // shared: (draw, primitive) that need the clipper
    uint4 *big;              // this draw's big-triangle bin records
    uint32_t *tri;           // triangle records of unclipped primitives (kRecWords each, by primitive)
    uint32_t *clip_tri;      // triangle records of clipped sub-triangles (by clip slot)
    float *clip_tris;        // clip_tri_cap x kClipFloats: positions and clip barycentrics
    uint32_t *clip_base;     // per primitive: first sub-triangle slot
    uint32_t *counters;      // see Counter
    unsigned long long *frags;  // kFragSlots fragment counters of the submission (summed on read)
    uint32_t *error;
    uint32_t *frag_px;       // optional per-pixel fragment count
    unsigned long long *gvis;   // split tiles: the merged winners, TS*TS per tile (kept zero between uses)
    uint32_t *tile_done;     // split tiles: their items that have finished (zero region)
    uint32_t tile0;          // global index of the draw's first tile
    uint32_t bin_cap;        // slots per tile bin
    uint32_t split;          // bin entries per work item of a tile (0: one item per tile)
    float split_inv;         // 1 / split
    uint32_t spill_cap, active_cap;
    uint32_t big_cap, clip_list_cap, clip_tri_cap;
    int32_t *strip_sep;      // restart triangle strips: the last separator position up to the end of
                             // each kStripChunk of index positions (-1: none)
};

constexpr int kFragSlots = 64;  // fragment counters, spread so that atomics rarely meet

// Work items of the tile kernels: a tile with more bin entries than one
// item holds is split into items of `split` entries (at most kMaxItems; the
// last one takes the rest, the spill entries and the big triangles).
// Active-list entries are global tile | item << kItemShift.
constexpr uint32_t kMaxItems = 16;
constexpr uint32_t kItemShift = 24;

// Per-draw counters.  The big-triangle footprint union is kept as maxima of
// (65535 - min) and max so that zero-initialised counters work.
enum Counter { C_BIG = 0, C_CLIP_TRI = 1, C_BIG_NX0 = 2, C_BIG_NY0 = 3, C_BIG_X1 = 4, C_BIG_Y1 = 5, C_NUM = 8 };

// Group header words: active tiles, primitives to clip, spilled bin
// entries, big triangles of every draw of the group; then the tile
// kernels' work queues, one ticket counter per XCD, 128 bytes apart
// (H_QUEUE + kQueueStride * xcd).
constexpr uint32_t kQueueStride = 32;
enum Header { H_ACTIVE = 0, H_CLIP = 1, H_SPILL = 2, H_BIG = 3, H_QUEUE = 32, H_NUM = 32 + 8 * 32 };

// Triangle record: everything the raster needs about one (sub-)triangle,
// computed once in setup (the same values the raster would compute).
constexpr int kRecWords = 36;
enum RecWord {
    R_VP = 0,      // vpx0 vpy0 vpx1 vpy1 vpx2 vpy2
    R_K = 6,       // k0 k1 k2 total
    R_Z = 10,      // z0 z1 z2
    R_IW = 13,     // 1/w of the unclipped vertices
    R_SCAN = 16,   // ax ay bx by cx cy (after bubblesort3 by y)
    R_MODE = 22,   // scan mode | wide << 8
    R_KEY = 23,
    R_VAR = 24,    // 3 x kMaxVar varyings of the unclipped vertices
};

// One resource slot of a WGSL program: a buffer (uniform / storage), a
// texture or a sampler, by the slot the translator assigned.
struct ResBind {
    const uint8_t *ptr;      // buffer data or texels
    uint64_t size;           // buffer bytes
    uint32_t w, h, srgb;     // texture size; texels need sRGB decoding
    uint32_t addr_u, addr_v, bilinear;   // sampler
};
constexpr int kMaxRes = 8;

struct DrawParams {
    // pipeline
    uint32_t program, nvar, interp, blend, bilinear, srgb;
    uint32_t front_face_cw, cull_mode;
    uint32_t depth_test, depth_compare, depth_write, resolve;
    // vertex input
    uint32_t n_attr;
    AttrBind attr[kMaxAttr];
    const uint8_t *index_data;
    uint64_t index_size;
    uint32_t index_format;
    int32_t base_vertex;
    uint32_t indexed, first, count, first_instance, instance_count;
    uint32_t nprim, total_prims;
    uint32_t topology, strip_restart;   // Topology; strip_index_format set (index MAX restarts strips)
    // bindings
    const float *uniform;
    const uint8_t *tex;
    uint32_t tex_w, tex_h, tex_srgb, addr_u, addr_v;
    const float *srgb_lut;
    // dynamic state
    float viewport[6];
    uint32_t scissor[4];
    // framebuffer and its tiling (tile = 1 << tile_shift pixels square)
    uint32_t width, height, tiles_x, tiles_y, tile_shift;
    uint8_t *color;
    uint32_t color_bgra;
    float *depth;
    uint32_t init_clear;     // this raster launch applies the LoadOp::Clear values
    uint32_t preclear;       // setup stored them into the whole target: the tile kernels store only
                             // the pixels the draw writes
    uint32_t clear_color_on, clear_depth_on;
    uint32_t clear_rgba;     // packed in memory byte order
    float clear_depth;
    // WGSL programs: resource slots, and per varying word whether it is
    // flat (first vertex) or perspective-corrected (else linear)
    ResBind res[kMaxRes];
    uint32_t var_flat_mask, var_persp_mask;
    // WGSL fragment shaders: @early_depth_test (0 none, 1 force, 2 allow) and
    // whether a frag_depth output replaces the depth of the late test
    uint32_t fs_early, fs_depth_tested;
    const void *user;        // host-side program identity (never read on the device)
    Scratch s;
};

struct VsOut;
struct TriLds;
#ifdef WGC_USER_PROGRAM
// Defined by the pipeline's generated glue, compiled with the kernels:
// returns 0 or the ErrorBits of a failed invocation.
__device__ __forceinline__ uint32_t user_vertex(const DrawParams &d, uint32_t instance, uint32_t vertex, VsOut &o);
// A fragment with interpolated varyings `var` of primitive
// `primitive_index`; `pos(p)` writes its @builtin(position).  Returns 0
// (colour written to `out`, and to *frag_depth the depth the late test uses
// when the shader writes one before its colour), wgsl ABORT_KILL (discard)
// or another abort code.
template <class Pos>
__device__ __forceinline__ uint32_t user_fragment(const DrawParams &d, const Pos &pos, bool front_facing,
                                                  uint32_t primitive_index, const float *var, float out[4],
                                                  float *frag_depth);
#endif

// ------------------------------------------------------------------
// Scalar helpers with Rust semantics
// ------------------------------------------------------------------

// `f as u32`: NaN -> 0, saturating, truncating.
__device__ __forceinline__ uint32_t f32_as_u32(float f)
{
    if (!(f == f)) return 0u;
    if (f <= 0.0f) return 0u;
    if (f >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)f;
}

// texture.rs:376-379: (v * 255).clamp(0, 255) as u8
__device__ __forceinline__ uint32_t f32_to_unorm8(float v)
{
    float x = v * 255.0f;
    if (x < 0.0f) x = 0.0f;
    if (x > 255.0f) x = 255.0f;
    if (!(x == x)) return 0u;
    return (uint32_t)x;
}

// x % m for m in {1, 2}: x/m and m*trunc(x/m) are exact, and so is the
// difference (Sterbenz), i.e. the same value as the IEEE fmod the reference's
// `%` computes (up to the sign of an exact zero, which no later step sees).
__device__ __forceinline__ float rem_pow2(float x, float m)
{
    float q = truncf(x / m);
    return x - m * q;
}

// f32::rem_euclid for m in {1, 2}
__device__ __forceinline__ float rem_euclid_pow2(float x, float m)
{
    float r = rem_pow2(x, m);
    if (r < 0.0f) r = r + m;
    return r;
}

// binding.rs:150-164
__device__ __forceinline__ uint32_t texel_coordinate(float x, uint32_t mode, uint32_t size)
{
    if (mode == ADDR_CLAMP) {
        if (x < 0.0f) x = 0.0f;
        if (x > 1.0f) x = 1.0f;
    } else if (mode == ADDR_REPEAT) {
        x = rem_euclid_pow2(x, 1.0f);
    } else {
        float r = rem_euclid_pow2(x, 2.0f);
        x = (r <= 1.0f) ? r : 2.0f - r;
    }
    return f32_as_u32(roundf(x * (float)(size - 1u)));
}

// render_pass/mod.rs:434-452
__device__ __forceinline__ bool compare(uint32_t fn, float v, float r)
{
    switch (fn) {
    case CMP_NEVER: return false;
    case CMP_LESS: return v < r;
    case CMP_EQUAL: return v == r;
    case CMP_LESS_EQUAL: return v <= r;
    case CMP_GREATER: return v > r;
    case CMP_NOT_EQUAL: return v != r;
    case CMP_GREATER_EQUAL: return v >= r;
    case CMP_ALWAYS: return true;
    default: return false;
    }
}

// ------------------------------------------------------------------
// Vertex stage (vertex.rs:128-196 fetch; the built-in programs)
// ------------------------------------------------------------------

struct VsOut {
    float clip[4];
    float var[kMaxVar];
};

// vertex.rs:191-195: buffer[instance * stride_i + vertex * stride_v + offset..]
// The load is issued whether or not the attribute lies inside the buffer
// (outside, it reads the buffer's start and the result is false): no branch
// on the address, so the loads of every vertex of a primitive are in flight
// together.
template <uint32_t LOC>
__device__ __forceinline__ bool fetch_attr(const DrawParams &d, uint32_t instance, uint32_t vertex,
                                           float *dst, uint32_t want)
{
    if (LOC >= d.n_attr) return false;
    const AttrBind at = d.attr[LOC];
    if (!at.data || at.ncomp < want || at.size < 4ull * at.ncomp) return false;
    uint64_t start = at.step_instance ? (uint64_t)instance * at.stride : (uint64_t)vertex * at.stride;
    uint64_t off = start + at.offset;
    bool in = off + 4ull * at.ncomp <= at.size;
    const float *p = reinterpret_cast<const float *>(at.data + (in ? off : 0ull));
#pragma unroll
    for (uint32_t k = 0; k < (uint32_t)kMaxVar; ++k)
        if (k < want) dst[k] = p[k];
    return in;
}

// naga-cranelift binary.rs:299-321: column sum with separate fmul / fadd
__device__ __forceinline__ void mat4_mul_vec4(const float *m, const float *v, float *r)
{
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        float s = m[0 * 4 + k] * v[0];
        s = s + m[1 * 4 + k] * v[1];
        s = s + m[2 * 4 + k] * v[2];
        s = s + m[3 * 4 + k] * v[3];
        r[k] = s;
    }
}

// The built-in programs' vertex shaders; `m` is the MVP matrix of the two
// mesh programs (read once per primitive, by the caller).
template <int PROG>
__device__ __forceinline__ bool run_vs(const DrawParams &d, uint32_t instance, uint32_t vertex,
                                       VsOut &o, const float *m)
{
    if constexpr (PROG == PROG_MVP_COLOR || PROG == PROG_MVP_TEXTURE) {
        // both attributes are fetched before either is used
        float pos[4];
        o.var[0] = o.var[1] = o.var[2] = o.var[3] = 0.0f;
        bool ok0 = fetch_attr<0>(d, instance, vertex, pos, 4);
        bool ok1 = fetch_attr<1>(d, instance, vertex, o.var, d.nvar);
        mat4_mul_vec4(m, pos, o.clip);
        return ok0 && ok1;
    } else if constexpr (PROG == PROG_CLIP_COLOR) {
        o.var[0] = o.var[1] = o.var[2] = o.var[3] = 0.0f;
        bool ok0 = fetch_attr<0>(d, instance, vertex, o.clip, 4);
        bool ok1 = fetch_attr<1>(d, instance, vertex, o.var, d.nvar);
        return ok0 && ok1;
    } else {
        // colored_triangle.wgsl
        uint32_t vi = vertex % 3u;
        int32_t ix = (int32_t)vi - 1;
        int32_t iy = (int32_t)(vi & 1u) * 2 - 1;
        o.clip[0] = (float)ix;
        o.clip[1] = (float)iy;
        o.clip[2] = 0.0f;
        o.clip[3] = 1.0f;
        o.var[0] = (vi == 0u) ? 1.0f : 0.0f;
        o.var[1] = (vi == 1u) ? 1.0f : 0.0f;
        o.var[2] = (vi == 2u) ? 1.0f : 0.0f;
        o.var[3] = 1.0f;
        return true;
    }
}

// index.rs:45-70 for the N positions i[k]: every index is loaded before
// any is checked (a position outside the index buffer loads its first index
// and fails), so the N loads are in flight together.
template <int N>
__device__ __forceinline__ bool resolve_indices(const DrawParams &d, const uint32_t i[N], uint32_t vertex[N])
{
    if (!d.indexed) {
#pragma unroll
        for (int k = 0; k < N; ++k) vertex[k] = i[k];
        return true;
    }
    if (d.index_size < d.index_format) return false;
    bool ok = true;
    uint32_t idx[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        uint64_t off = (uint64_t)i[k] * d.index_format;
        bool in = off + d.index_format <= d.index_size;
        ok = ok && in;
        const uint8_t *q = d.index_data + (in ? off : 0ull);
        idx[k] = d.index_format == 2u ? (uint32_t)*reinterpret_cast<const uint16_t *>(q)
                                      : *reinterpret_cast<const uint32_t *>(q);
    }
#pragma unroll
    for (int k = 0; k < N; ++k) {
        int64_t v = (int64_t)idx[k] + d.base_vertex;
        ok = ok && v >= 0 && v <= 0xFFFFFFFFll;
        vertex[k] = (uint32_t)v;
    }
    return ok;
}

// The raw index value at position i of the draw (restart separators are
// raw values, index.rs:72-106); false when outside the index buffer.
__device__ __forceinline__ bool raw_index(const DrawParams &d, uint32_t i, uint32_t &idx)
{
    uint64_t off = (uint64_t)i * d.index_format;
    if (off + d.index_format > d.index_size) return false;
    if (d.index_format == 2u) idx = *reinterpret_cast<const uint16_t *>(d.index_data + off);
    else idx = *reinterpret_cast<const uint32_t *>(d.index_data + off);
    return true;
}

// Index positions (relative to `first`) of primitive `prim` of one instance
// in assembly order (primitive.rs:206-487): lists take n consecutive
// positions; strips start at `prim`, odd triangles of a strip swap their
// first two vertices (TriStripGenerator); with restart a primitive that
// would span a separator does not exist, and the parity counts from the
// start of its segment (ResetTriStripIter).  Returns 1, 0 (no primitive) or
// -1 (an index outside the index buffer).
template <int N>
__device__ __forceinline__ int prim_positions(const DrawParams &d, uint32_t prim, uint32_t pos[N])
{
    if (d.topology == TOPO_POINT_LIST || d.topology == TOPO_LINE_LIST || d.topology == TOPO_TRIANGLE_LIST) {
#pragma unroll
        for (int k = 0; k < N; ++k) pos[k] = prim * (uint32_t)N + (uint32_t)k;
        return 1;
    }
#pragma unroll
    for (int k = 0; k < N; ++k) pos[k] = prim + (uint32_t)k;
    bool odd = false;
    if (d.strip_restart && d.indexed) {
        uint32_t sep = d.index_format == 2u ? 0xFFFFu : 0xFFFFFFFFu;
#pragma unroll
        for (int k = 0; k < N; ++k) {
            uint32_t idx;
            if (!raw_index(d, d.first + pos[k], idx)) return -1;
            if (idx == sep) return 0;
        }
        if (N == 3) {
            // segment start: the position after the last separator before prim
            uint32_t c0 = prim / kStripChunk * kStripChunk;
            int64_t last = -1;
            for (int64_t j = (int64_t)prim - 1; j >= (int64_t)c0; --j) {
                uint32_t idx;
                if (!raw_index(d, d.first + (uint32_t)j, idx)) return -1;
                if (idx == sep) {
                    last = j;
                    break;
                }
            }
            if (last < 0 && c0 > 0) last = d.s.strip_sep[c0 / kStripChunk - 1];
            odd = (((int64_t)prim - (last + 1)) & 1) != 0;
        }
    } else {
        odd = (prim & 1u) != 0u;
    }
    if (N == 3 && odd) {
        uint32_t t = pos[0];
        pos[0] = pos[1];
        pos[1] = t;
    }
    return 1;
}

// Runs the vertex shader for the N vertices of primitive `p` (global across
// instances); false when there is no such primitive (a restart separator) or
// on failure, which raises an error bit.
template <int PROG, int N = 3>
__device__ __forceinline__ bool shade_primitive(const DrawParams &d, uint32_t p, VsOut v[N])
{
    uint32_t inst = d.first_instance + p / d.nprim;
    uint32_t prim = p % d.nprim;
    uint32_t pos[N];
    int pr = prim_positions<N>(d, prim, pos);
    if (pr <= 0) {
        if (pr < 0) atomicOr(d.s.error, (uint32_t)ERR_INDEX);
        return false;
    }
    // every index first, then every vertex: no vertex's loads wait for
    // another vertex's
    uint32_t vertex[N], at[N];
#pragma unroll
    for (int k = 0; k < N; ++k) at[k] = d.first + pos[k];
    bool iok = resolve_indices<N>(d, at, vertex);
    if (!iok) {
        atomicOr(d.s.error, (uint32_t)ERR_INDEX);
        return false;
    }
#ifdef WGC_USER_PROGRAM
    if constexpr (PROG == PROG_USER) {
        uint32_t e = 0u;
#pragma unroll
        for (int k = 0; k < N; ++k) e |= user_vertex(d, inst, vertex[k], v[k]);
        if (e) {
            atomicOr(d.s.error, e);
            return false;
        }
        return true;
    }
#endif
    float m[16];
    if constexpr (PROG == PROG_MVP_COLOR || PROG == PROG_MVP_TEXTURE) {
#pragma unroll
        for (int k = 0; k < 16; ++k) m[k] = d.uniform[k];
    }
    bool vok = true;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        bool ok = run_vs<PROG>(d, inst, vertex[k], v[k], m);
        vok = vok && ok;
    }
    if (!vok) {
        atomicOr(d.s.error, (uint32_t)ERR_VERTEX);
        return false;
    }
    return true;
}

// primitive.rs:169-197 + state.rs:542-561.  Returns false when w == 0.
__device__ __forceinline__ bool face_and_cull(const DrawParams &d, const VsOut v[3],
                                              bool &front_facing, bool &culled)
{
    float nd[3][2];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (v[k].clip[3] == 0.0f) return false;
        nd[k][0] = v[k].clip[0] / v[k].clip[3];
        nd[k][1] = v[k].clip[1] / v[k].clip[3];
    }
    float abx = nd[1][0] - nd[0][0], aby = nd[1][1] - nd[0][1];
    float acx = nd[2][0] - nd[0][0], acy = nd[2][1] - nd[0][1];
    float z = abx * acy - aby * acx;
    bool ccw = z > 0.0f;
    front_facing = ccw ? (d.front_face_cw == 0u) : (d.front_face_cw == 1u);
    culled = (d.cull_mode == 1u) ? front_facing : (d.cull_mode == 2u) ? !front_facing : false;
    return true;
}

// ------------------------------------------------------------------
// Clipper (clipper.rs:150-179, 625-802)
// ------------------------------------------------------------------

// nalgebra Vector4::dot with a WebGPU plane: (p0*q0 + p2*q2) + (p1*q1 + p3*q3)
__device__ __forceinline__ float clip_distance(int plane, const float *p)
{
    float q0 = 0.0f, q1 = 0.0f, q2 = 0.0f, q3 = 1.0f;
    switch (plane) {
    case 0: q0 = 1.0f; break;
    case 1: q0 = -1.0f; break;
    case 2: q1 = 1.0f; break;
    case 3: q1 = -1.0f; break;
    case 4: q2 = 1.0f; q3 = 0.0f; break;
    default: q2 = -1.0f; break;
    }
    float a = p[0] * q0;
    float b = p[1] * q1;
    float c = p[2] * q2;
    float e = p[3] * q3;
    a = a + c;
    b = b + e;
    return a + b;
}

// Bit per clip plane that the vertex is outside of (distance < 0).
__device__ __forceinline__ uint32_t outcode(const float *p)
{
    uint32_t c = 0;
#pragma unroll
    for (int pl = 0; pl < 6; ++pl) c |= (clip_distance(pl, p) < 0.0f) ? (1u << pl) : 0u;
    return c;
}

// 0: no vertex outside any plane -- the clipper returns the triangle
// unchanged with identity barycentrics; 2: all three vertices outside the
// first plane (in the clipper's order, clipper.rs:153-163) that any vertex
// is outside of -- the planes before it leave the triangle untouched and
// that plane drops it, so the clipper emits nothing; 1: the clipper must run.
__device__ __forceinline__ int clip_class(const VsOut v[3])
{
    uint32_t c0 = outcode(v[0].clip), c1 = outcode(v[1].clip), c2 = outcode(v[2].clip);
    uint32_t any = c0 | c1 | c2;
    if (any == 0u) return 0;
    uint32_t first = any & (0u - any);
    if ((c0 & c1 & c2 & first) != 0u) return 2;
    return 1;
}

struct CTri {
    float p[3][4];
    float b[3][3];
};

__device__ __forceinline__ void clip_lerp(const CTri &t, int i, int j, float tt, float *p, float *b)
{
    float s = 1.0f - tt;
#pragma unroll
    for (int k = 0; k < 4; ++k) p[k] = t.p[i][k] * s + t.p[j][k] * tt;
#pragma unroll
    for (int k = 0; k < 3; ++k) b[k] = t.b[i][k] * s + t.b[j][k] * tt;
}

__device__ __forceinline__ void set_vert(CTri &o, int slot, const float *p, const float *b)
{
#pragma unroll
    for (int k = 0; k < 4; ++k) o.p[slot][k] = p[k];
#pragma unroll
    for (int k = 0; k < 3; ++k) o.b[slot][k] = b[k];
}

// clipper.rs:625-728
__device__ inline int clip_against_plane(const CTri &t, int plane, CTri out[2])
{
    float dd[3];
    bool o[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        dd[k] = clip_distance(plane, t.p[k]);
        o[k] = dd[k] < 0.0f;
    }
    float p1[4], b1[3], p2[4], b2[3];
#define WGC_CLIP(i, j, P, B) clip_lerp(t, i, j, dd[i] / (dd[i] - dd[j]), P, B)
    if (o[0] && o[1] && o[2]) return 0;
    if (!o[0] && !o[1] && !o[2]) { out[0] = t; return 1; }
    if (o[0] && o[1] && !o[2]) {  // clip(2,0), clip(1,2) -> [A', B', C]
        WGC_CLIP(2, 0, p1, b1); WGC_CLIP(1, 2, p2, b2);
        set_vert(out[0], 0, p1, b1); set_vert(out[0], 1, p2, b2); set_vert(out[0], 2, t.p[2], t.b[2]);
        return 1;
    }
    if (!o[0] && o[1] && o[2]) {  // clip(0,1), clip(2,0) -> [A, B', C']
        WGC_CLIP(0, 1, p1, b1); WGC_CLIP(2, 0, p2, b2);
        set_vert(out[0], 0, t.p[0], t.b[0]); set_vert(out[0], 1, p1, b1); set_vert(out[0], 2, p2, b2);
        return 1;
    }
    if (o[0] && !o[1] && o[2]) {  // clip(0,1), clip(1,2) -> [A', B, C']
        WGC_CLIP(0, 1, p1, b1); WGC_CLIP(1, 2, p2, b2);
        set_vert(out[0], 0, p1, b1); set_vert(out[0], 1, t.p[1], t.b[1]); set_vert(out[0], 2, p2, b2);
        return 1;
    }
    if (o[0] && !o[1] && !o[2]) {  // clip(0,1), clip(0,2) -> [B', B, C'], [B, C, C']
        WGC_CLIP(0, 1, p1, b1); WGC_CLIP(0, 2, p2, b2);
        set_vert(out[0], 0, p1, b1); set_vert(out[0], 1, t.p[1], t.b[1]); set_vert(out[0], 2, p2, b2);
        set_vert(out[1], 0, t.p[1], t.b[1]); set_vert(out[1], 1, t.p[2], t.b[2]); set_vert(out[1], 2, p2, b2);
        return 2;
    }
    if (!o[0] && o[1] && !o[2]) {  // clip(0,1), clip(1,2) -> [A, A', C'], [A, C', C]
        WGC_CLIP(0, 1, p1, b1); WGC_CLIP(1, 2, p2, b2);
        set_vert(out[0], 0, t.p[0], t.b[0]); set_vert(out[0], 1, p1, b1); set_vert(out[0], 2, p2, b2);
        set_vert(out[1], 0, t.p[0], t.b[0]); set_vert(out[1], 1, p2, b2); set_vert(out[1], 2, t.p[2], t.b[2]);
        return 2;
    }
    // !o[0] && !o[1] && o[2]: clip(1,2), clip(0,2) -> [A, B, B'], [A, B', A']
    WGC_CLIP(1, 2, p1, b1); WGC_CLIP(0, 2, p2, b2);
    set_vert(out[0], 0, t.p[0], t.b[0]); set_vert(out[0], 1, t.p[1], t.b[1]); set_vert(out[0], 2, p1, b1);
    set_vert(out[1], 0, t.p[0], t.b[0]); set_vert(out[1], 1, p1, b1); set_vert(out[1], 2, p2, b2);
    return 2;
#undef WGC_CLIP
}

// ------------------------------------------------------------------
// Raster setup (raster.rs:129-160, 280-321; scanline.rs:14-56)
// ------------------------------------------------------------------

struct RasterTri {
    float vpx[3], vpy[3], z[3];
    uint32_t fbx[3], fby[3];
    float total;
};

__device__ __forceinline__ void to_raster(const float *vp, const float *clip, uint32_t &fbx,
                                          uint32_t &fby, float &vx, float &vy, float &nz)
{
    float tx = vp[0] + 0.5f * vp[2];
    float ty = vp[1] + 0.5f * vp[3];
    float sx = 0.5f * vp[2];
    float sy = 0.5f * vp[3];
    sy = sy * -1.0f;
    float nx = clip[0] / clip[3];
    float ny = clip[1] / clip[3];
    nz = clip[2] / clip[3];
    vx = nx * sx + tx;
    vy = ny * sy + ty;
    vx = fmaxf(vx, 0.0f);
    vy = fmaxf(vy, 0.0f);
    fbx = f32_as_u32(vx);
    fby = f32_as_u32(vy);
}

__device__ __forceinline__ float shoelace(float ax, float ay, float bx, float by, float cx, float cy)
{
    float t1 = (by - ay) * (bx + ax);
    float t2 = (cy - by) * (cx + bx);
    float t3 = (ay - cy) * (ax + cx);
    return (t1 + t2) + t3;
}

__device__ __forceinline__ void raster_setup(const DrawParams &d, const float pos[3][4], RasterTri &r)
{
#pragma unroll
    for (int k = 0; k < 3; ++k) to_raster(d.viewport, pos[k], r.fbx[k], r.fby[k], r.vpx[k], r.vpy[k], r.z[k]);
    r.total = shoelace(r.vpx[0], r.vpy[0], r.vpx[1], r.vpy[1], r.vpx[2], r.vpy[2]);
}

// Vertices of the scanline rule after bubblesort3 by y, with the case the
// iterator starts in (scanline.rs:24-53).
struct ScanTri {
    int32_t ax, ay, bx, by, cx, cy;
    int32_t mode;   // 0: single row (ac.y == 0), 1: start in the second half (ab.y == 0), 2: general
    int32_t pad;
};

__device__ __forceinline__ ScanTri scan_setup(const RasterTri &r)
{
    int32_t x[3], y[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) { x[k] = (int32_t)r.fbx[k]; y[k] = (int32_t)r.fby[k]; }
    // sort.rs:3-7: compare_and_swap (0,1), (0,2), (1,2), swap only on Greater
#define WGC_CAS(l, rr) if (y[l] > y[rr]) { int32_t t = y[l]; y[l] = y[rr]; y[rr] = t; t = x[l]; x[l] = x[rr]; x[rr] = t; }
    WGC_CAS(0, 1) WGC_CAS(0, 2) WGC_CAS(1, 2)
#undef WGC_CAS
    ScanTri s;
    s.ax = x[0]; s.ay = y[0]; s.bx = x[1]; s.by = y[1]; s.cx = x[2]; s.cy = y[2];
    s.pad = 0;
    if (s.cy == s.ay) {
        s.mode = 0;
        // bubblesort3 by x; only the extremes are used
        s.ax = min(min(x[0], x[1]), x[2]);
        s.cx = max(max(x[0], x[1]), x[2]);
    } else if (s.by == s.ay) {
        s.mode = 1;
    } else {
        s.mode = 2;
    }
    return s;
}

// Rows the scanline iterator emits: [row0, row1] (row1 < row0 when none).
__device__ __forceinline__ void scan_rows(const ScanTri &s, int32_t &row0, int32_t &row1)
{
    row0 = s.ay;
    if (s.mode == 0) row1 = s.ay;
    else if (s.mode == 1) row1 = s.cy;
    else row1 = (s.cy == s.by) ? s.by - 1 : s.cy;   // flat bottom: row c.y is not emitted
}

__device__ __forceinline__ void scan_cols(const ScanTri &s, int32_t &col0, int32_t &col1)
{
    if (s.mode == 0) { col0 = s.ax; col1 = s.cx; }
    else { col0 = min(min(s.ax, s.bx), s.cx); col1 = max(max(s.ax, s.bx), s.cx); }
}

// Inclusive span [x1, x2] of row y (scanline.rs:107-166).  Returns false
// when the iterator emits no row y.  Rust's isize `/` truncates toward zero,
// as does C++'s; products fit in 32 bits for coordinates below 2^15, and use
// 64 bits above.
__device__ __forceinline__ bool scan_span(const ScanTri &s, int32_t y, int32_t &x1, int32_t &x2, bool wide)
{
    int32_t a, b;
    if (s.mode == 0) {
        if (y != s.ay) return false;
        x1 = s.ax; x2 = s.cx;
        return true;
    }
    int32_t acx = s.cx - s.ax, acy = s.cy - s.ay;
    int32_t i = y - s.ay;
    if (s.mode == 1) {
        if (y < s.ay || y > s.cy) return false;
        int32_t bcx = s.cx - s.bx, bcy = s.cy - s.by;
        if (wide) {
            a = s.ax + (int32_t)(((int64_t)acx * i) / acy);
            b = s.bx + (int32_t)(((int64_t)bcx * i) / bcy);
        } else {
            a = s.ax + (acx * i) / acy;
            b = s.bx + (bcx * i) / bcy;
        }
    } else {
        if (y < s.ay) return false;
        if (y < s.by) {
            int32_t abx = s.bx - s.ax, aby = s.by - s.ay;
            if (wide) {
                a = s.ax + (int32_t)(((int64_t)acx * i) / acy);
                b = s.ax + (int32_t)(((int64_t)abx * i) / aby);
            } else {
                a = s.ax + (acx * i) / acy;
                b = s.ax + (abx * i) / aby;
            }
        } else {
            int32_t bcy = s.cy - s.by;
            if (bcy == 0 || y > s.cy) return false;
            int32_t bcx = s.cx - s.bx;
            int32_t i2 = y - s.by;
            if (wide) {
                a = s.ax + (int32_t)(((int64_t)acx * i) / acy);
                b = s.bx + (int32_t)(((int64_t)bcx * i2) / bcy);
            } else {
                a = s.ax + (acx * i) / acy;
                b = s.bx + (bcx * i2) / bcy;
            }
        }
    }
    x1 = min(a, b);
    x2 = max(a, b);
    return true;
}

__device__ __forceinline__ bool scan_wide(const RasterTri &r)
{
    uint32_t m = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) m |= r.fbx[k] | r.fby[k];
    return m >= 32768u;
}

__device__ __forceinline__ bool coord_ok(const RasterTri &r)
{
    uint32_t m = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) m |= r.fbx[k] | r.fby[k];
    return m < (1u << 30);
}

// Tile footprint of the pixels the scanline rule emits inside the scissor.
// Returns false when it is empty.
__device__ __forceinline__ bool tile_bbox(const DrawParams &d, const ScanTri &s, uint32_t &tx0,
                                          uint32_t &ty0, uint32_t &tx1, uint32_t &ty1)
{
    int32_t r0, r1, c0, c1;
    scan_rows(s, r0, r1);
    scan_cols(s, c0, c1);
    int64_t sx0 = d.scissor[0], sy0 = d.scissor[1];
    int64_t sx1 = sx0 + d.scissor[2] - 1, sy1 = sy0 + d.scissor[3] - 1;
    int64_t x0 = max((int64_t)c0, sx0), x1 = min((int64_t)c1, sx1);
    int64_t y0 = max((int64_t)r0, sy0), y1 = min((int64_t)r1, sy1);
    if (x0 > x1 || y0 > y1) return false;
    tx0 = (uint32_t)(x0 >> d.tile_shift); tx1 = (uint32_t)(x1 >> d.tile_shift);
    ty0 = (uint32_t)(y0 >> d.tile_shift); ty1 = (uint32_t)(y1 >> d.tile_shift);
    return true;
}

// ------------------------------------------------------------------
// Fragment stage
// ------------------------------------------------------------------

// u8 / 255.0 (texture.rs:173-175), correctly rounded: the product with the
// rounded reciprocal, then one fused correction with the exact residual
// (Markstein's step).  For every byte this is the division's value (checked
// against exact rational arithmetic for all 256 inputs); no table, so no
// memory access.
__device__ __forceinline__ float u8f(uint32_t b)
{
    const float r = 1.0f / 255.0f;
    float x = (float)(b & 0xFFu);
    float q = x * r;
    float e = __builtin_fmaf(-q, 255.0f, x);
    return __builtin_fmaf(e, r, q);
}

// A bound texture with its sampler's addressing and filter.
struct TexView {
    const uint8_t *tex;
    uint32_t w, h, srgb, addr_u, addr_v, bilinear;
};

__device__ __forceinline__ TexView builtin_texture(const DrawParams &d)
{
    return TexView{d.tex, d.tex_w, d.tex_h, d.tex_srgb, d.addr_u, d.addr_v, d.bilinear};
}

// `srgb`: the draw's sRGB decode table, or a copy of it in LDS.
__device__ __forceinline__ void fetch_texel(const DrawParams &d, const TexView &tv, uint32_t u, uint32_t v,
                                            float out[4], const float *srgb)
{
    uint32_t t = *reinterpret_cast<const uint32_t *>(tv.tex + ((uint64_t)v * tv.w + u) * 4u);
    if (d.srgb && tv.srgb) {
        out[0] = srgb[t & 0xFFu];
        out[1] = srgb[(t >> 8) & 0xFFu];
        out[2] = srgb[(t >> 16) & 0xFFu];
        out[3] = u8f(t >> 24);
    } else {
        out[0] = u8f(t);   // texture.rs:173-175
        out[1] = u8f(t >> 8);
        out[2] = u8f(t >> 16);
        out[3] = u8f(t >> 24);
    }
}

__device__ __forceinline__ int32_t wrap_index(int32_t i, uint32_t mode, uint32_t size)
{
    int32_t n = (int32_t)size;
    if (mode == ADDR_REPEAT) { int32_t r = i % n; return r < 0 ? r + n : r; }
    if (mode == ADDR_MIRROR) {
        int32_t p = 2 * n;
        int32_t r = i % p;
        if (r < 0) r += p;
        return r < n ? r : p - 1 - r;
    }
    return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

// Extension: bilinear, WebGPU texel-centre convention.
__device__ __forceinline__ void sample_bilinear(const DrawParams &d, const TexView &tv, float u, float v,
                                                float out[4], const float *srgb)
{
    float fu = u * (float)tv.w - 0.5f;
    float fv = v * (float)tv.h - 0.5f;
    float u0f = floorf(fu), v0f = floorf(fv);
    float au = fu - u0f, av = fv - v0f;
    int32_t iu = (int32_t)fmaxf(fminf(u0f, 1073741824.0f), -1073741824.0f);
    int32_t iv = (int32_t)fmaxf(fminf(v0f, 1073741824.0f), -1073741824.0f);
    uint32_t x0 = (uint32_t)wrap_index(iu, tv.addr_u, tv.w);
    uint32_t y0 = (uint32_t)wrap_index(iv, tv.addr_v, tv.h);
    // Repeat: (i + 1) mod n is the next index after i mod n, wrapped
    uint32_t x1 = tv.addr_u == ADDR_REPEAT ? (x0 + 1u < tv.w ? x0 + 1u : 0u)
                                           : (uint32_t)wrap_index(iu + 1, tv.addr_u, tv.w);
    uint32_t y1 = tv.addr_v == ADDR_REPEAT ? (y0 + 1u < tv.h ? y0 + 1u : 0u)
                                           : (uint32_t)wrap_index(iv + 1, tv.addr_v, tv.h);
    float t00[4], t10[4], t01[4], t11[4];
    fetch_texel(d, tv, x0, y0, t00, srgb);
    fetch_texel(d, tv, x1, y0, t10, srgb);
    fetch_texel(d, tv, x0, y1, t01, srgb);
    fetch_texel(d, tv, x1, y1, t11, srgb);
    float su = 1.0f - au, sv = 1.0f - av;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        float top = t00[k] * su + t10[k] * au;
        float bot = t01[k] * su + t11[k] * au;
        out[k] = top * sv + bot * av;
    }
}

// textureSample (binding.rs:16-164; bilinear is the extension).
__device__ __forceinline__ void sample_texture(const DrawParams &d, const TexView &tv, float u, float v,
                                               float out[4], const float *srgb)
{
    if (tv.bilinear) {
        sample_bilinear(d, tv, u, v, out, srgb);
    } else {
        uint32_t x = texel_coordinate(u, tv.addr_u, tv.w);
        uint32_t y = texel_coordinate(v, tv.addr_v, tv.h);
        fetch_texel(d, tv, x, y, out, srgb);
    }
}

__device__ __forceinline__ void sample_texture(const DrawParams &d, const TexView &tv, float u, float v,
                                               float out[4])
{
    sample_texture(d, tv, u, v, out, d.srgb_lut);
}

// `srgb`: the sRGB decode table to read (the draw's, or a copy in LDS).
template <int PROG>
__device__ __forceinline__ void run_fs(const DrawParams &d, const float *var, float out[4], const float *srgb)
{
    if constexpr (PROG == PROG_MVP_TEXTURE) {
        sample_texture(d, builtin_texture(d), var[0], var[1], out, srgb);
    } else {
        out[0] = var[0];
        out[1] = var[1];
        out[2] = var[2];
        out[3] = (d.nvar >= 4u) ? var[3] : 1.0f;
    }
}

// TexelWriter::from_color packed in memory byte order.
__device__ __forceinline__ uint32_t pack_color(const float c[4], uint32_t bgra)
{
    uint32_t r = f32_to_unorm8(c[0]), g = f32_to_unorm8(c[1]);
    uint32_t b = f32_to_unorm8(c[2]), a = f32_to_unorm8(c[3]);
    return bgra ? (b | (g << 8) | (r << 16) | (a << 24)) : (r | (g << 8) | (b << 16) | (a << 24));
}

// Extension: src-over (BlendState::ALPHA_BLENDING) against the stored unorm8.
__device__ __forceinline__ uint32_t blend_over(const float s[4], uint32_t dst, uint32_t bgra)
{
    float dc[4];
    uint32_t c0 = dst & 0xFFu, c1 = (dst >> 8) & 0xFFu, c2 = (dst >> 16) & 0xFFu, c3 = dst >> 24;
    dc[0] = u8f(bgra ? c2 : c0);
    dc[1] = u8f(c1);
    dc[2] = u8f(bgra ? c0 : c2);
    dc[3] = u8f(c3);
    float ia = 1.0f - s[3];
    float o[4];
#pragma unroll
    for (int k = 0; k < 3; ++k) o[k] = s[k] * s[3] + dc[k] * ia;
    o[3] = s[3] + dc[3] * ia;
    return pack_color(o, bgra);
}

}  // namespace wgc
