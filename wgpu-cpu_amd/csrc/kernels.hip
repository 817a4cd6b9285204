// kernels.hip -- HIP kernels of the render-pass executor for gfx950.
//
// Per launch group: k_zero -> k_setup -> k_clip -> k_tile_scan -> k_bin_fill
// -> k_raster.  A group is either one draw (parameters passed by value) or
// a table of independent draws -- the batched form, where hundreds of render
// passes share every launch (MULTI).
//
//  k_setup     one thread per primitive: vertex shader for its three indices
//              (state.rs:519-535), winding and culling (state.rs:540-561),
//              the clipper's pass-through test, raster setup and the tile
//              footprint; wave-aggregated appends (ballot + prefix count)
//              to the record list, the big-triangle list and the clip list.
//  k_clip      primitives that cross a clip plane: the 6-plane clipper as an
//              in-order depth-first walk (same output order as the
//              reference's plane-by-plane queue, clipper.rs:756-775).
//  k_tile_scan exclusive scan of per-tile counts (all draws of the group).
//  k_bin_fill  scatters record keys into per-tile lists.
//  k_raster    one workgroup per TSxTS tile, PPT pixels per lane; colour and
//              depth live in registers from load (or clear) to one coalesced
//              store.  Triangles are staged 64 at a time in LDS, with per-row
//              spans of the reference's integer scanline rule.
#pragma clang fp contract(off)

#include "wgc_core.h"
#include "wgc_launch.h"

namespace wgc {

// ------------------------------------------------------------------
// helpers
// ------------------------------------------------------------------

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Compile-time loop: f(IC<I>{}) for I in [B, E), so per-pixel arrays are only
// ever indexed by constants and live in registers.
template <int I>
struct IC {
    static constexpr int value = I;
};
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (B < E) {
        f(IC<B>{});
        static_for<B + 1, E>(f);
    }
}

template <bool MULTI>
__device__ __forceinline__ const DrawParams &draw_at(const DrawParams &d0, const BatchArgs &b, uint32_t i)
{
    if constexpr (MULTI) return b.draws[i];
    else return d0;
}

// Largest i < n with base[i] <= g (empty ranges resolve to the last of them).
__device__ __forceinline__ uint32_t find_slot(const uint32_t *base, uint32_t n, uint32_t g)
{
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1u) {
        uint32_t mid = (lo + hi) >> 1;
        if (base[mid] <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Wave-aggregated append: one atomic per wave, slots in lane order.  Every
// lane of the wave must pass the same counter.
__device__ __forceinline__ uint32_t wave_append(uint32_t *counter, bool pred)
{
    uint64_t mask = __ballot(pred);
    if (mask == 0) return 0xFFFFFFFFu;
    uint32_t lane = lane_id();
    int leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0;
    if ((int)lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    uint64_t below = (lane == 0) ? 0ull : (mask & ((~0ull) >> (64 - lane)));
    return pred ? base + (uint32_t)__popcll(below) : 0xFFFFFFFFu;
}

__device__ __forceinline__ uint32_t xcd_tile(uint32_t i, uint32_t n)
{
    // Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
    // contiguous band of tiles so neighbouring tiles share its L2.
    uint32_t xcd = i & 7u, local = i >> 3;
    uint32_t per = n >> 3, rem = n & 7u;
    return xcd * per + min(xcd, rem) + local;
}

__device__ __forceinline__ void add_tile_counts(const DrawParams &d, uint32_t tx0, uint32_t ty0,
                                                uint32_t tx1, uint32_t ty1)
{
    for (uint32_t ty = ty0; ty <= ty1; ++ty)
        for (uint32_t tx = tx0; tx <= tx1; ++tx) atomicAdd(&d.s.tile_count[ty * d.tiles_x + tx], 1u);
}

// ------------------------------------------------------------------
// k_zero: tile counts of the group, per-draw counters, record counter
// ------------------------------------------------------------------

__global__ void k_zero(uint32_t *tile_count, uint32_t n_tiles, uint32_t *counters, uint32_t n_counters,
                       uint32_t *rec_count)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_tiles) tile_count[i] = 0u;
    if (i < n_counters) counters[i] = 0u;
    if (i == 0) *rec_count = 0u;
}

// ------------------------------------------------------------------
// k_setup
// ------------------------------------------------------------------

template <int PROG, bool MULTI>
__global__ __launch_bounds__(256) void k_setup(DrawParams d0, BatchArgs b)
{
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t di = 0, p = g;
    if constexpr (MULTI) {
        // draw ranges are padded to whole waves, so a wave never spans draws
        if (g >= b.total_prims) return;
        di = find_slot(b.prim_base, b.n, g);
        p = g - b.prim_base[di];
    }
    const DrawParams &d = draw_at<MULTI>(d0, b, di);
    bool active = p < d.total_prims;
    bool has_rec = false, needs_clip = false;
    uint4 rec = make_uint4(0, 0, 0, 0);
    uint32_t ntiles = 0;
    if (active) {
        VsOut v[3];
        if (shade_primitive<PROG>(d, p, v)) {
            bool ff, culled;
            if (!face_and_cull(d, v, ff, culled)) {
                atomicOr(d.s.error, (uint32_t)ERR_W_ZERO);
            } else if (!culled) {
                if (fully_inside(v)) {
                    float pos[3][4];
#pragma unroll
                    for (int k = 0; k < 3; ++k)
#pragma unroll
                        for (int c = 0; c < 4; ++c) pos[k][c] = v[k].clip[c];
                    RasterTri r;
                    raster_setup(d, pos, r);
                    if (!coord_ok(r)) {
                        atomicOr(d.s.error, (uint32_t)ERR_COORD_RANGE);
                    } else if (r.total != 0.0f) {
                        ScanTri s = scan_setup(r);
                        uint32_t tx0, ty0, tx1, ty1;
                        if (tile_bbox(d, s, tx0, ty0, tx1, ty1)) {
                            has_rec = true;
                            rec = make_uint4(p << 7, tx0 | (ty0 << 16), tx1 | (ty1 << 16), di);
                            ntiles = (tx1 - tx0 + 1) * (ty1 - ty0 + 1);
                        }
                    }
                } else {
                    needs_clip = true;
                }
            }
        }
    }
    bool small = has_rec && ntiles <= (uint32_t)kMaxBinTiles;
    bool big = has_rec && !small;
    uint32_t slot = wave_append(&d.s.counters[C_CLIP_LIST], needs_clip);
    if (needs_clip) {
        if (slot < d.s.clip_list_cap) d.s.clip_list[slot] = p;
        else atomicOr(d.s.error, (uint32_t)ERR_CLIP_OVERFLOW);
    }
    slot = wave_append(d.s.rec_count, small);
    if (small) {
        if (slot < d.s.rec_cap) {
            d.s.recs[slot] = rec;
            add_tile_counts(d, rec.y & 0xFFFFu, rec.y >> 16, rec.z & 0xFFFFu, rec.z >> 16);
        } else {
            atomicOr(d.s.error, (uint32_t)ERR_RECORD_OVERFLOW);
        }
    }
    slot = wave_append(&d.s.counters[C_BIG], big);
    if (big) {
        if (slot < d.s.big_cap) d.s.big[slot] = rec;
        else atomicOr(d.s.error, (uint32_t)ERR_RECORD_OVERFLOW);
    }
}

// ------------------------------------------------------------------
// k_clip
// ------------------------------------------------------------------

// Depth-first walk over the plane-by-plane splits; leaves come out in the
// order of the reference's queue (lexicographic in the split choices).  At
// most one pending sibling per level: 7 stack entries.
template <class F>
__device__ int clip_walk(const CTri &root, F &&emit)
{
    CTri stack[8];
    int lvl[8];
    int sp = 1;
    stack[0] = root;
    lvl[0] = 0;
    int leaf = 0;
    while (sp > 0) {
        --sp;
        CTri t = stack[sp];
        int L = lvl[sp];
        if (L == 6) {
            emit(t, leaf);
            ++leaf;
            continue;
        }
        CTri out[2];
        int n = clip_against_plane(t, L, out);
        if (n == 2) {
            stack[sp] = out[1]; lvl[sp] = L + 1; ++sp;
            stack[sp] = out[0]; lvl[sp] = L + 1; ++sp;
        } else if (n == 1) {
            stack[sp] = out[0]; lvl[sp] = L + 1; ++sp;
        }
    }
    return leaf;
}

template <int PROG>
__device__ void clip_primitive(const DrawParams &d, uint32_t di, uint32_t p)
{
    VsOut v[3];
    if (!shade_primitive<PROG>(d, p, v)) return;
    CTri root;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
#pragma unroll
        for (int c = 0; c < 4; ++c) root.p[k][c] = v[k].clip[c];
#pragma unroll
        for (int c = 0; c < 3; ++c) root.b[k][c] = (c == k) ? 1.0f : 0.0f;
    }
    int nleaf = clip_walk(root, [](const CTri &, int) {});
    if (nleaf == 0) return;
    uint32_t base = atomicAdd(&d.s.counters[C_CLIP_TRI], (uint32_t)nleaf);
    if (base + (uint32_t)nleaf > d.s.clip_tri_cap) {
        atomicOr(d.s.error, (uint32_t)ERR_CLIP_OVERFLOW);
        return;
    }
    d.s.clip_base[p] = base;
    clip_walk(root, [&](const CTri &t, int leaf) {
        float *o = d.s.clip_tris + (size_t)(base + (uint32_t)leaf) * kClipFloats;
        for (int k = 0; k < 3; ++k) {
            for (int c = 0; c < 4; ++c) o[k * 7 + c] = t.p[k][c];
            for (int c = 0; c < 3; ++c) o[k * 7 + 4 + c] = t.b[k][c];
        }
        RasterTri r;
        raster_setup(d, t.p, r);
        if (!coord_ok(r)) { atomicOr(d.s.error, (uint32_t)ERR_COORD_RANGE); return; }
        if (r.total == 0.0f) return;
        ScanTri s = scan_setup(r);
        uint32_t tx0, ty0, tx1, ty1;
        if (!tile_bbox(d, s, tx0, ty0, tx1, ty1)) return;
        uint4 rec = make_uint4((p << 7) | ((uint32_t)leaf << 1) | 1u, tx0 | (ty0 << 16),
                               tx1 | (ty1 << 16), di);
        uint32_t nt = (tx1 - tx0 + 1) * (ty1 - ty0 + 1);
        if (nt <= (uint32_t)kMaxBinTiles) {
            uint32_t slot = atomicAdd(d.s.rec_count, 1u);
            if (slot < d.s.rec_cap) {
                d.s.recs[slot] = rec;
                add_tile_counts(d, tx0, ty0, tx1, ty1);
            } else {
                atomicOr(d.s.error, (uint32_t)ERR_RECORD_OVERFLOW);
            }
        } else {
            uint32_t slot = atomicAdd(&d.s.counters[C_BIG], 1u);
            if (slot < d.s.big_cap) d.s.big[slot] = rec;
            else atomicOr(d.s.error, (uint32_t)ERR_RECORD_OVERFLOW);
        }
    });
}

template <int PROG, bool MULTI>
__global__ __launch_bounds__(256) void k_clip(DrawParams d0, BatchArgs b)
{
    uint32_t total = MULTI ? b.total_prims : d0.total_prims;
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < total; g += gridDim.x * blockDim.x) {
        uint32_t di = 0, i = g;
        if constexpr (MULTI) {
            di = find_slot(b.prim_base, b.n, g);
            i = g - b.prim_base[di];
        }
        const DrawParams &d = draw_at<MULTI>(d0, b, di);
        uint32_t n = min(d.s.counters[C_CLIP_LIST], d.s.clip_list_cap);
        if (i < n) clip_primitive<PROG>(d, di, d.s.clip_list[i]);
    }
}

// ------------------------------------------------------------------
// k_tile_scan: one workgroup of 1024 threads
// ------------------------------------------------------------------

__global__ __launch_bounds__(1024) void k_tile_scan(uint32_t *tile_count, uint32_t *tile_offset,
                                                    uint32_t *tile_fill, uint32_t n)
{
    __shared__ uint32_t sums[1024];
    uint32_t t = threadIdx.x;
    uint32_t chunk = (n + 1023u) / 1024u;
    uint32_t b = t * chunk, e = min(b + chunk, n);
    uint32_t s = 0;
    for (uint32_t i = b; i < e; ++i) s += tile_count[i];
    sums[t] = s;
    __syncthreads();
    for (uint32_t off = 1; off < 1024u; off <<= 1) {
        uint32_t v = (t >= off) ? sums[t - off] : 0u;
        __syncthreads();
        sums[t] += v;
        __syncthreads();
    }
    uint32_t run = sums[t] - s;
    for (uint32_t i = b; i < e; ++i) {
        tile_offset[i] = run;
        run += tile_count[i];
        tile_fill[i] = 0u;
    }
    if (t == 1023u) tile_offset[n] = sums[1023];
}

// ------------------------------------------------------------------
// k_bin_fill
// ------------------------------------------------------------------

template <bool MULTI>
__global__ __launch_bounds__(256) void k_bin_fill(DrawParams d0, BatchArgs b, const uint4 *recs,
                                                  const uint32_t *rec_count, uint32_t rec_cap, uint32_t *bins)
{
    uint32_t n = min(*rec_count, rec_cap);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint4 r = recs[i];
        const DrawParams &d = draw_at<MULTI>(d0, b, r.w);
        uint32_t tx0 = r.y & 0xFFFFu, ty0 = r.y >> 16, tx1 = r.z & 0xFFFFu, ty1 = r.z >> 16;
        for (uint32_t ty = ty0; ty <= ty1; ++ty)
            for (uint32_t tx = tx0; tx <= tx1; ++tx) {
                uint32_t t = ty * d.tiles_x + tx;
                uint32_t slot = atomicAdd(&d.s.tile_fill[t], 1u);
                bins[d.s.tile_offset[t] + slot] = r.x;
            }
    }
}

// ------------------------------------------------------------------
// k_raster
// ------------------------------------------------------------------

struct TriLds {
    float vpx[3], vpy[3];
    float k0, k1, k2, total;
    float z[3];
    float cb[9];
    float var[3][kMaxVar];
    float iw[3];
    uint32_t key;
    uint32_t valid;
};

template <int TS>
struct RasterShared {
    TriLds tri[kBatch];
    ScanTri scan[kBatch];
    uint32_t wide[kBatch];
    uint32_t span[kBatch][TS];
    uint32_t keys[kBatch];
    uint32_t nkeys;
};

// Stage one list entry (thread-local): the vertex shader for the unclipped
// vertices (varyings, 1/w), the clipped sub-triangle when the key says so,
// raster setup and the scanline case.
template <int PROG>
__device__ void stage_entry(const DrawParams &d, uint32_t key, TriLds &T, ScanTri &S, uint32_t &wide)
{
    uint32_t p = key >> 7, sub = (key >> 1) & 63u;
    bool clipped = key & 1u;
    VsOut v[3];
    T.key = key;
    T.valid = 0u;
    wide = 0u;
    if (!shade_primitive<PROG>(d, p, v)) return;
    float pos[3][4];
    float cb[3][3];
    if (clipped) {
        const float *ct = d.s.clip_tris + (size_t)(d.s.clip_base[p] + sub) * kClipFloats;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
#pragma unroll
            for (int c = 0; c < 4; ++c) pos[k][c] = ct[k * 7 + c];
#pragma unroll
            for (int c = 0; c < 3; ++c) cb[k][c] = ct[k * 7 + 4 + c];
        }
    } else {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
#pragma unroll
            for (int c = 0; c < 4; ++c) pos[k][c] = v[k].clip[c];
#pragma unroll
            for (int c = 0; c < 3; ++c) cb[k][c] = (c == k) ? 1.0f : 0.0f;
        }
    }
    RasterTri r;
    raster_setup(d, pos, r);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        T.vpx[k] = r.vpx[k];
        T.vpy[k] = r.vpy[k];
        T.z[k] = r.z[k];
        T.iw[k] = 1.0f / v[k].clip[3];
#pragma unroll
        for (int c = 0; c < kMaxVar; ++c) T.var[k][c] = v[k].var[c];
#pragma unroll
        for (int c = 0; c < 3; ++c) T.cb[k * 3 + c] = cb[k][c];
    }
    T.total = r.total;
    // constant terms of the three shoelace sums (the same products the
    // per-pixel evaluation would form)
    T.k0 = (r.vpy[2] - r.vpy[1]) * (r.vpx[2] + r.vpx[1]);
    T.k1 = (r.vpy[0] - r.vpy[2]) * (r.vpx[0] + r.vpx[2]);
    T.k2 = (r.vpy[1] - r.vpy[0]) * (r.vpx[1] + r.vpx[0]);
    S = scan_setup(r);
    wide = scan_wide(r) ? 1u : 0u;
    T.valid = (r.total != 0.0f && coord_ok(r)) ? 1u : 0u;
}

struct PixelState {
    uint32_t col;
    float dep;
    float init_dep;
    uint32_t key;
    bool has;
};

// One fragment of triangle T at pixel (pxf, pyf): barycentric at the pixel
// corner (raster.rs:313-346), the depth test or its order-independent
// restatement, varyings (fragment.rs:319-377), the fragment shader and the
// colour write (texture.rs:213-218).
template <int PROG, bool ORDERED>
__device__ __forceinline__ void shade_fragment(const DrawParams &d, const TriLds &T, float pxf, float pyf,
                                               PixelState &ps)
{
    float l0 = (((T.vpy[1] - pyf) * (T.vpx[1] + pxf) + T.k0) + (pyf - T.vpy[2]) * (pxf + T.vpx[2])) / T.total;
    float l1 = (((pyf - T.vpy[0]) * (pxf + T.vpx[0]) + (T.vpy[2] - pyf) * (T.vpx[2] + pxf)) + T.k1) / T.total;
    float l2 = ((T.k2 + (pyf - T.vpy[1]) * (pxf + T.vpx[1])) + (T.vpy[0] - pyf) * (T.vpx[0] + pxf)) / T.total;
    float lc[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float acc = T.cb[0 * 3 + c] * l0;
        acc = acc + T.cb[1 * 3 + c] * l1;
        acc = acc + T.cb[2 * 3 + c] * l2;
        lc[c] = acc;
    }
    float z = T.z[0] * lc[0];
    z = z + T.z[1] * lc[1];
    z = z + T.z[2] * lc[2];

    bool win;
    if (ORDERED) {
        win = d.depth_test ? compare(d.depth_compare, z, ps.dep) : true;
    } else {
        switch (d.resolve) {
        case RES_LAST: win = !ps.has || T.key > ps.key; break;
        case RES_MIN_FIRST:
            win = (z < ps.init_dep) && (!ps.has || z < ps.dep || (z == ps.dep && T.key < ps.key));
            break;
        case RES_MIN_LAST:
            win = (z <= ps.init_dep) && (!ps.has || z < ps.dep || (z == ps.dep && T.key > ps.key));
            break;
        case RES_MAX_FIRST:
            win = (z > ps.init_dep) && (!ps.has || z > ps.dep || (z == ps.dep && T.key < ps.key));
            break;
        case RES_MAX_LAST:
            win = (z >= ps.init_dep) && (!ps.has || z > ps.dep || (z == ps.dep && T.key > ps.key));
            break;
        default:  // RES_PASS_LAST
            win = compare(d.depth_compare, z, ps.init_dep) && (!ps.has || T.key > ps.key);
            break;
        }
    }
    if (!win) return;
    float var[kMaxVar];
    if (d.interp == INTERP_FLAT) {
#pragma unroll
        for (int c = 0; c < kMaxVar; ++c) var[c] = T.var[0][c];
    } else if (d.interp == INTERP_PERSPECTIVE) {
        float q0 = lc[0] * T.iw[0];
        float q1 = lc[1] * T.iw[1];
        float q2 = lc[2] * T.iw[2];
        float qs = (q0 + q1) + q2;
#pragma unroll
        for (int c = 0; c < kMaxVar; ++c) {
            float acc = T.var[0][c] * q0;
            acc = acc + T.var[1][c] * q1;
            acc = acc + T.var[2][c] * q2;
            var[c] = acc / qs;
        }
    } else {
#pragma unroll
        for (int c = 0; c < kMaxVar; ++c) {
            float acc = T.var[0][c] * lc[0];
            acc = acc + T.var[1][c] * lc[1];
            acc = acc + T.var[2][c] * lc[2];
            var[c] = acc;
        }
    }
    float color[4];
    run_fs<PROG>(d, var, color);
    if (ORDERED) {
        if (d.depth_test && d.depth_write) ps.dep = z;
        ps.col = d.blend ? blend_over(color, ps.col, d.color_bgra) : pack_color(color, d.color_bgra);
        ps.has = true;
    } else {
        ps.col = pack_color(color, d.color_bgra);
        ps.dep = z;
        ps.key = T.key;
        ps.has = true;
    }
}

// Stages sh.keys[0..n) into LDS and applies them to this lane's pixels.
template <int PROG, bool ORDERED, int TS, int NT>
__device__ __forceinline__ void raster_batch(const DrawParams &d, RasterShared<TS> &sh, uint32_t n, uint32_t x0, uint32_t y0,
                             PixelState (&ps)[TS * TS / NT], uint32_t (&nfrag)[TS * TS / NT])
{
    constexpr int PPT = TS * TS / NT;
    uint32_t tid = threadIdx.x;
    if (tid < n) stage_entry<PROG>(d, sh.keys[tid], sh.tri[tid], sh.scan[tid], sh.wide[tid]);
    __syncthreads();
    // per-row spans inside the scissor and the tile, tile-relative
    int64_t sx0 = d.scissor[0], sy0 = d.scissor[1];
    int64_t sx1 = sx0 + d.scissor[2], sy1 = sy0 + d.scissor[3];
    for (uint32_t idx = tid; idx < n * (uint32_t)TS; idx += NT) {
        uint32_t e = idx / TS, row = idx % TS;
        int32_t y = (int32_t)(y0 + row);
        uint32_t span = 1u;  // empty: lo = 1, hi = 0
        int32_t x1, x2;
        if (sh.tri[e].valid && (int64_t)y >= sy0 && (int64_t)y < sy1 &&
            scan_span(sh.scan[e], y, x1, x2, sh.wide[e] != 0u)) {
            int64_t lo = max((int64_t)x1, max(sx0, (int64_t)x0));
            int64_t hi = min((int64_t)x2, min(sx1 - 1, (int64_t)x0 + TS - 1));
            if (lo <= hi) span = (uint32_t)(lo - x0) | ((uint32_t)(hi - x0) << 16);
        }
        sh.span[e][row] = span;
    }
    __syncthreads();
    // pixels outer (unrolled, so the per-pixel state stays in registers),
    // triangles inner in list order
    static_for<0, PPT>([&](auto kc) __attribute__((always_inline)) {
        constexpr int k = decltype(kc)::value;
        uint32_t q = tid + (uint32_t)k * NT;
        uint32_t lx = q % TS, ly = q / TS;
        float pxf = (float)(x0 + lx), pyf = (float)(y0 + ly);
        for (uint32_t e = 0; e < n; ++e) {
            uint32_t sp = sh.span[e][ly];
            if (lx < (sp & 0xFFFFu) || lx > (sp >> 16)) continue;
            nfrag[k]++;
            shade_fragment<PROG, ORDERED>(d, sh.tri[e], pxf, pyf, ps[k]);
        }
    });
    __syncthreads();
}

template <int PROG, bool ORDERED, bool MULTI, int TS, int NT>
__global__ __launch_bounds__(NT) void k_raster(DrawParams d0, BatchArgs b)
{
    constexpr int PPT = TS * TS / NT;
    static_assert(NT >= kBatch && PPT >= 1 && TS * TS == PPT * NT, "tile shape");
    __shared__ RasterShared<TS> sh;
    __shared__ uint32_t olist[ORDERED ? kOrderedCap : 1];
    __shared__ uint32_t olist_n;

    uint32_t n_tiles = MULTI ? b.total_tiles : d0.tiles_x * d0.tiles_y;
    uint32_t gt = xcd_tile(blockIdx.x, n_tiles);
    uint32_t di = 0, tile = gt;
    if constexpr (MULTI) {
        di = find_slot(b.tile_base, b.n, gt);
        tile = gt - b.tile_base[di];
    }
    const DrawParams &d = draw_at<MULTI>(d0, b, di);
    uint32_t tx = tile % d.tiles_x, ty = tile / d.tiles_x;
    uint32_t x0 = tx * TS, y0 = ty * TS;
    uint32_t tid = threadIdx.x;

    uint32_t seg0 = d.s.tile_offset[tile];
    uint32_t seg_n = d.s.tile_offset[tile + 1] - seg0;
    uint32_t nbig = min(d.s.counters[C_BIG], d.s.big_cap);
    if (!d.init_clear && seg_n == 0 && nbig == 0) return;

    PixelState ps[PPT];
    uint32_t init_col[PPT];
    uint32_t nfrag[PPT];
    static_for<0, PPT>([&](auto kc) __attribute__((always_inline)) {
        constexpr int k = decltype(kc)::value;
        nfrag[k] = 0u;
        uint32_t q = tid + (uint32_t)k * NT;
        uint32_t px = x0 + q % TS, py = y0 + q / TS;
        bool inside = px < d.width && py < d.height;
        size_t pix = (size_t)py * d.width + px;
        ps[k].has = false;
        ps[k].key = 0u;
        ps[k].col = 0u;
        ps[k].dep = 0.0f;
        if (inside) {
            if (d.color)
                ps[k].col = (d.init_clear && d.clear_color_on) ? d.clear_rgba
                                                               : reinterpret_cast<const uint32_t *>(d.color)[pix];
            if (d.depth) ps[k].dep = (d.init_clear && d.clear_depth_on) ? d.clear_depth : d.depth[pix];
        }
        ps[k].init_dep = ps[k].dep;
        init_col[k] = ps[k].col;
    });

    uint32_t total = 0;  // ordered: length of the sorted list
    if (ORDERED) {
        // gather the tile's entries, sort them into API order
        if (tid == 0) olist_n = 0u;
        __syncthreads();
        bool overflow = seg_n > (uint32_t)kOrderedCap;
        if (!overflow) {
            for (uint32_t i = tid; i < seg_n; i += NT) olist[i] = d.s.bins[seg0 + i];
            if (tid == 0) olist_n = seg_n;
        }
        __syncthreads();
        for (uint32_t bb = 0; bb < nbig && !overflow; bb += kBatch) {
            uint32_t base = olist_n;
            __syncthreads();  // every thread has read olist_n before thread 0 moves it
            if (tid < kBatch) {
                bool hit = false;
                uint32_t key = 0u;
                if (bb + tid < nbig) {
                    uint4 r = d.s.big[bb + tid];
                    hit = tx >= (r.y & 0xFFFFu) && tx <= (r.z & 0xFFFFu) && ty >= (r.y >> 16) && ty <= (r.z >> 16);
                    key = r.x;
                }
                uint64_t mask = __ballot(hit);
                uint64_t below = (tid == 0) ? 0ull : (mask & ((~0ull) >> (64 - tid)));
                uint32_t pos = base + (uint32_t)__popcll(below);
                if (hit && pos < (uint32_t)kOrderedCap) olist[pos] = key;
                if (tid == 0) olist_n = base + (uint32_t)__popcll(mask);
            }
            __syncthreads();
            overflow = olist_n > (uint32_t)kOrderedCap;
        }
        total = olist_n;
        if (overflow || total > (uint32_t)kOrderedCap) {
            if (tid == 0) atomicOr(d.s.error, (uint32_t)ERR_ORDERED_OVERFLOW);
            total = 0u;
        }
        // bitonic sort of olist[0..P), padded with UINT32_MAX
        uint32_t P = 1u;
        while (P < total) P <<= 1;
        for (uint32_t i = total + tid; i < P; i += NT) olist[i] = 0xFFFFFFFFu;
        __syncthreads();
        for (uint32_t k = 2; k <= P; k <<= 1) {
            for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                for (uint32_t i = tid; i < P; i += NT) {
                    uint32_t ixj = i ^ j;
                    if (ixj > i) {
                        uint32_t a = olist[i], c = olist[ixj];
                        bool up = (i & k) == 0u;
                        if ((a > c) == up) { olist[i] = c; olist[ixj] = a; }
                    }
                }
                __syncthreads();
            }
        }
    }

    // One loop, one call site (so the per-pixel state stays in registers):
    // ordered -> the sorted list; otherwise the tile's bin segment, then the
    // big-triangle list filtered by tile overlap.
    uint32_t fpos = 0, bpos = 0;
    for (;;) {
        uint32_t n;
        if (ORDERED) {
            if (fpos >= total) break;
            n = min((uint32_t)kBatch, total - fpos);
            if (tid < n) sh.keys[tid] = olist[fpos + tid];
            fpos += n;
        } else if (fpos < seg_n) {
            n = min((uint32_t)kBatch, seg_n - fpos);
            if (tid < n) sh.keys[tid] = d.s.bins[seg0 + fpos + tid];
            fpos += n;
        } else if (bpos < nbig) {
            if (tid < kBatch) {  // wave 0 compacts the hits
                bool hit = false;
                uint32_t key = 0u;
                if (bpos + tid < nbig) {
                    uint4 r = d.s.big[bpos + tid];
                    hit = tx >= (r.y & 0xFFFFu) && tx <= (r.z & 0xFFFFu) && ty >= (r.y >> 16) && ty <= (r.z >> 16);
                    key = r.x;
                }
                uint64_t mask = __ballot(hit);
                uint64_t below = (tid == 0) ? 0ull : (mask & ((~0ull) >> (64 - tid)));
                if (hit) sh.keys[__popcll(below)] = key;
                if (tid == 0) sh.nkeys = (uint32_t)__popcll(mask);
            }
            __syncthreads();
            n = sh.nkeys;
            bpos += kBatch;
            if (n == 0) {
                __syncthreads();  // sh.nkeys is rewritten next round
                continue;
            }
        } else {
            break;
        }
        raster_batch<PROG, ORDERED, TS, NT>(d, sh, n, x0, y0, ps, nfrag);
    }

    uint32_t f = 0;
    static_for<0, PPT>([&](auto kc) __attribute__((always_inline)) {
        constexpr int k = decltype(kc)::value;
        f += nfrag[k];
        uint32_t q = tid + (uint32_t)k * NT;
        uint32_t px = x0 + q % TS, py = y0 + q / TS;
        if (px >= d.width || py >= d.height) return;
        size_t pix = (size_t)py * d.width + px;
        bool write_depth;
        if (ORDERED) {
            write_depth = d.depth_test && d.depth_write;
        } else {
            write_depth = ps[k].has && d.depth_test && d.depth_write &&
                          (d.resolve != RES_PASS_LAST || d.depth_compare == CMP_EQUAL ||
                           d.depth_compare == CMP_ALWAYS);
            if (!ps[k].has) ps[k].col = init_col[k];
        }
        if (d.color) reinterpret_cast<uint32_t *>(d.color)[pix] = ps[k].col;
        if (d.depth) d.depth[pix] = write_depth ? ps[k].dep : ps[k].init_dep;
        if (d.s.frag_px && nfrag[k]) d.s.frag_px[pix] += nfrag[k];
    });
    // fragment count: wave reduction, one atomic per wave
    for (int off = 32; off > 0; off >>= 1) f += __shfl_xor(f, off);
    if (lane_id() == 0 && f) atomicAdd(d.s.frags, (unsigned long long)f);
}

// ------------------------------------------------------------------
// k_clear / k_reset
// ------------------------------------------------------------------

__global__ void k_reset(unsigned long long *frags)
{
    if (threadIdx.x == 0) frags[0] = 0ull;
}

__global__ void k_clear(uint32_t *color, uint32_t rgba, float *depth, float dv, uint64_t n)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (color) color[i] = rgba;
    if (depth) depth[i] = dv;
}

// ------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------

#define WGC_LAUNCH(name, kern, grid, block, ...)                                        \
    do {                                                                                \
        if (kt) kt->begin(name, st);                                                    \
        hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, st, __VA_ARGS__);          \
        hipError_t e_ = hipGetLastError();                                              \
        if (kt) kt->end(st);                                                            \
        if (e_ != hipSuccess) return e_;                                                \
    } while (0)

template <int PROG, bool ORDERED, bool MULTI>
static hipError_t launch_raster(const DrawParams &d, const BatchArgs &b, uint32_t tiles, uint32_t cfg,
                                hipStream_t st, KernelTimer *kt)
{
    if (ORDERED) {
        WGC_LAUNCH("k_raster", (k_raster<PROG, true, MULTI, 16, 256>), tiles, 256, d, b);
        return hipSuccess;
    }
    switch (cfg) {
    case RASTER_8x8_64: WGC_LAUNCH("k_raster", (k_raster<PROG, false, MULTI, 8, 64>), tiles, 64, d, b); break;
    case RASTER_16x16_64: WGC_LAUNCH("k_raster", (k_raster<PROG, false, MULTI, 16, 64>), tiles, 64, d, b); break;
    case RASTER_32x32_256: WGC_LAUNCH("k_raster", (k_raster<PROG, false, MULTI, 32, 256>), tiles, 256, d, b); break;
    default: WGC_LAUNCH("k_raster", (k_raster<PROG, false, MULTI, 16, 256>), tiles, 256, d, b); break;
    }
    return hipSuccess;
}

template <int PROG, bool MULTI>
static hipError_t launch_group(const DrawParams &d, const BatchArgs &b, const GroupLaunch &gl, hipStream_t st,
                               KernelTimer *kt)
{
    uint32_t zero_n = max(max(gl.total_tiles, gl.n_counters), 1u);
    WGC_LAUNCH("k_zero", k_zero, (zero_n + 255) / 256, 256, gl.tile_count, gl.total_tiles, gl.counters,
               gl.n_counters, gl.rec_count);
    if (gl.total_prims > 0) {
        uint32_t g = (gl.total_prims + 255u) / 256u;
        WGC_LAUNCH("k_setup", (k_setup<PROG, MULTI>), g, 256, d, b);
        WGC_LAUNCH("k_clip", (k_clip<PROG, MULTI>), min(g, 1024u), 256, d, b);
    }
    WGC_LAUNCH("k_tile_scan", k_tile_scan, 1, 1024, gl.tile_count, gl.tile_offset, gl.tile_fill, gl.total_tiles);
    if (gl.total_prims > 0) {
        uint32_t g = min((gl.rec_cap + 255u) / 256u, 4096u);
        WGC_LAUNCH("k_bin_fill", k_bin_fill<MULTI>, g, 256, d, b, gl.recs, gl.rec_count, gl.rec_cap, gl.bins);
    }
    if (gl.ordered) return launch_raster<PROG, true, MULTI>(d, b, gl.total_tiles, gl.raster_cfg, st, kt);
    return launch_raster<PROG, false, MULTI>(d, b, gl.total_tiles, gl.raster_cfg, st, kt);
}

template <bool MULTI>
static hipError_t launch_group_prog(const DrawParams &d, const BatchArgs &b, const GroupLaunch &gl,
                                    hipStream_t st, KernelTimer *kt)
{
    switch (gl.program) {
    case PROG_MVP_COLOR: return launch_group<PROG_MVP_COLOR, MULTI>(d, b, gl, st, kt);
    case PROG_MVP_TEXTURE: return launch_group<PROG_MVP_TEXTURE, MULTI>(d, b, gl, st, kt);
    case PROG_COLORED_TRIANGLE: return launch_group<PROG_COLORED_TRIANGLE, MULTI>(d, b, gl, st, kt);
    case PROG_CLIP_COLOR: return launch_group<PROG_CLIP_COLOR, MULTI>(d, b, gl, st, kt);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_draw(const DrawParams &d, const GroupLaunch &gl, hipStream_t st, KernelTimer *kt)
{
    BatchArgs b{};
    return launch_group_prog<false>(d, b, gl, st, kt);
}

hipError_t launch_batch(const BatchArgs &b, const GroupLaunch &gl, hipStream_t st, KernelTimer *kt)
{
    DrawParams d{};
    return launch_group_prog<true>(d, b, gl, st, kt);
}

hipError_t launch_reset(unsigned long long *frags, hipStream_t st)
{
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, st, frags);
    return hipGetLastError();
}

hipError_t launch_clear(uint32_t *color, uint32_t rgba, float *depth, float dv, uint64_t n,
                        hipStream_t st, KernelTimer *kt)
{
    if (n == 0) return hipSuccess;
    WGC_LAUNCH("k_clear", k_clear, (unsigned)((n + 255) / 256), 256, color, rgba, depth, dv, n);
    return hipSuccess;
}

}  // namespace wgc
