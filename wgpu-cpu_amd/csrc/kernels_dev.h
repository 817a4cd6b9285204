// kernels_dev.h -- device code of the render-pass executor (gfx950).
//
// Per launch group: k_setup -> k_clip -> k_tiles (order-independent draws)
// or k_raster (ordered draws).  A group is one draw
// (parameters passed by value) or a table of independent draws -- the
// batched form, where hundreds of render passes share every launch (MULTI).
//
//  k_setup     one thread per primitive: vertex shader for its three indices
//              (state.rs:519-535), winding and culling (state.rs:540-561),
//              the clipper's pass-through test and raster setup; writes the
//              triangle record the raster reads and bins its key into the
//              tiles of its footprint (positions from one atomic add per tile
//              and group of lanes, the tile appended to the active list when
//              it is binned first), and appends (ballot + prefix count, one
//              atomic per wave) to the big-triangle list and the clip list.
//  k_clip      primitives that cross a clip plane: the 6-plane clipper as an
//              in-order depth-first walk (same output order as the
//              reference's plane-by-plane queue, clipper.rs:756-775).
//  k_tiles     persistent workgroups, statically scheduled over the active
//              tiles: the winners of each pixel (an order-independent minimum
//              of depth and API-order key in LDS), then the winners shaded,
//              one coalesced colour and depth store; then chunks of untouched
//              tiles (the LoadOp::Clear values, or only big triangles).
//  k_raster    the same schedule for ordered draws: each tile's keys sorted
//              into API order, colour and depth in registers from load or
//              clear to one store, triangles staged 64 at a time in LDS with
//              per-row spans of the reference's integer scanline rule.
//
// Compiled twice: into the library for the built-in programs, and at
// pipeline creation (hipRTC) together with the device functions emitted
// for a WGSL module, for the kernels that run shader code.
#pragma once
#pragma clang fp contract(off)

#include "wgc_core.h"
#include "wgc_batch.h"

namespace wgc {

// ------------------------------------------------------------------
// helpers
// ------------------------------------------------------------------

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Streaming stores (non-temporal) for target pixels that no later step of
// the same kernel reads back, so that they do not stay dirty in L2 until
// the end of the kernel (whose write-back then delays the next launch).
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void store_stream(uint32_t *p, uint32_t v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void store_stream4(uint32_t *p, uint32_t v)
{
    u32x4_t q = {v, v, v, v};
    __builtin_nontemporal_store(q, reinterpret_cast<u32x4_t *>(p));
}

// v from lane `lane`, the same lane for the whole wave (v_readlane: no LDS
// round trip, unlike a shuffle whose source lane differs per lane).
__device__ __forceinline__ uint32_t lane_value(uint32_t v, int lane)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)v, lane);
}

// Inclusive prefix sum over the 64 lanes of a full wave, with DPP: shifts
// by 1, 2, 4 and 8 lanes within each row of 16, then lane 15 of each row
// added to the next row and lane 31 to rows 2 and 3.
__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t v)
{
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);  // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);  // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);  // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);  // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);  // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return v;
}

template <bool MULTI>
__device__ __forceinline__ const DrawParams &draw_at(const DrawParams &d0, const BatchArgs &b, uint32_t i)
{
    if constexpr (MULTI) return b.draws[i];
    else return d0;
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
    base = lane_value(base, leader);
    uint64_t below = (lane == 0) ? 0ull : (mask & ((~0ull) >> (64 - lane)));
    return pred ? base + (uint32_t)__popcll(below) : 0xFFFFFFFFu;
}

// Lanes of the wave below this one.
__device__ __forceinline__ uint64_t lanes_below()
{
    uint32_t lane = lane_id();
    return lane == 0 ? 0ull : ((~0ull) >> (64 - lane));
}

// Wave-aggregated atomicAdd of v (below 256) to *addr: lanes (with act)
// that pass the same address share one atomic and receive consecutive
// ranges in lane order.  Every active lane of the wave must call it.  The
// prefix sums are formed bit by bit from ballots, so lanes that are not
// active (or have left the kernel) contribute nothing.
__device__ __forceinline__ uint32_t wave_add(uint32_t *addr, uint32_t v, bool act)
{
    uint32_t lane = lane_id();
    uint64_t lt = lanes_below();
    act = act && v > 0u;
    uint64_t pending = __ballot(act);
    uint32_t res = 0;
    unsigned long long a = (unsigned long long)addr;
    while (pending) {
        int leader = __ffsll((unsigned long long)pending) - 1;
        uint32_t lo = lane_value((uint32_t)a, leader);
        uint32_t hi = lane_value((uint32_t)(a >> 32), leader);
        bool mine = act && a == (((unsigned long long)hi << 32) | lo);
        uint64_t m = __ballot(mine);
        uint32_t below = 0, total = 0;
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            uint64_t mb = __ballot(mine && ((v >> bit) & 1u));
            below += (uint32_t)__popcll(mb & lt) << bit;
            total += (uint32_t)__popcll(mb) << bit;
        }
        uint32_t base = 0;
        if ((int)lane == leader) base = atomicAdd(addr, total);
        base = lane_value(base, leader);
        if (mine) res = base + below;
        pending &= ~m;
    }
    return res;
}

// The k-th tile (row-major) of a footprint w tiles wide.
__device__ __forceinline__ uint32_t footprint_tile(const DrawParams &d, uint32_t tx0, uint32_t ty0, uint32_t w,
                                                   uint32_t k)
{
    return (ty0 + k / w) * d.tiles_x + (tx0 + k % w);
}

// One bin entry that did not fit its tile's bin: (global tile, key) to the
// spill list.
__device__ __forceinline__ void spill_entry(const DrawParams &d, uint32_t at, uint32_t tile, uint32_t key)
{
    if (at < d.s.spill_cap) d.s.spill[at] = make_uint2(d.s.tile0 + tile, key);
    else atomicOr(d.s.error, (uint32_t)ERR_RECORD_OVERFLOW);
}

// Whether bin position `slot` starts a work item of its tile (position 0,
// and every `split` positions inside the bin, up to kMaxItems items), and
// which.  Each position is taken once, so each item is appended once.
__device__ __forceinline__ bool item_start(const DrawParams &d, uint32_t slot, uint32_t &item)
{
    item = 0u;
    if (slot == 0u) return true;
    uint32_t sp = d.s.split;
    if (sp == 0u || slot >= d.s.bin_cap || slot >= kMaxItems * sp) return false;
    // slot / sp for slot < 16 * sp < 2^24: the binary32 product is within one
    // of the quotient, one correction either way makes it exact
    uint32_t q = (uint32_t)((float)slot * d.s.split_inv);
    if (q * sp > slot) --q;
    else if ((q + 1u) * sp <= slot) ++q;
    item = q;
    return q * sp == slot;
}

// A work item of a tile to the active list.
__device__ __forceinline__ void activate_item(const DrawParams &d, uint32_t at, uint32_t tile, uint32_t item)
{
    if (at < d.s.active_cap) d.s.active_list[at] = (d.s.tile0 + tile) | (item << kItemShift);
    else atomicOr(d.s.error, (uint32_t)ERR_RECORD_OVERFLOW);
}

// Bins every lane's record (with `has`) into the tiles of its footprint
// (row-major, at most kMaxBinTiles).  For the k-th tile of every lane's
// footprint, runs of consecutive lanes on the same tile form a group
// (consecutive primitives of a mesh mostly share tiles, and one tile may
// collect thousands of them); every run's first lane takes the run's bin
// positions with one atomic add.  All the adds of the wave are issued before any result is used, so
// the wave waits for them once.  Then the keys go to their positions, the
// entries past a full bin to the spill list and the work items the taken
// positions start (position 0 of a tile, and every item's first position
// of a tile with more entries than one item holds) to the active list (one
// atomic per list per wave).  Every lane of the wave must call it.
// ANY_DRAW: lanes may bin into different draws of a batched group (the
// clipper's waves), so lanes group by the tile counter's address rather
// than by the draw's tile index.
template <bool ANY_DRAW = false>
__device__ __forceinline__ void bin_wave(const DrawParams &d, bool has, uint32_t key, uint32_t tx0, uint32_t ty0,
                                         uint32_t tx1, uint32_t ty1)
{
    uint32_t w = has ? tx1 - tx0 + 1u : 1u;
    uint32_t ntl = has ? w * (ty1 - ty0 + 1u) : 0u;
    uint32_t lane = lane_id();
    uint64_t lt = lanes_below();
    // got[k] is written only by the lanes that take an atomic and read only
    // from those lanes: it has no initial value, so that no merge of the
    // results with one waits for the atomics one by one
    uint32_t got[kMaxBinTiles], grp[kMaxBinTiles];
    uint32_t nk = 0;
#pragma unroll
    for (int k = 0; k < kMaxBinTiles; ++k) grp[k] = 0u;
#pragma unroll
    for (int k = 0; k < kMaxBinTiles; ++k) {
        bool act = (uint32_t)k < ntl;
        uint64_t pending = __ballot(act);
        if (!pending) break;
        nk = (uint32_t)k + 1u;
        uint32_t t = act ? footprint_tile(d, tx0, ty0, w, (uint32_t)k) : 0xFFFFFFFFu;
        // the global tile identifies the counter in a group (the draws'
        // tile ranges do not overlap)
        uint32_t gt = ANY_DRAW ? (act ? d.s.tile0 + t : 0xFFFFFFFFu) : t;
        // runs of consecutive lanes on the same counter (inactive lanes carry
        // an id no tile has, so they end every run): one shuffle and one
        // ballot, no loop
        uint32_t prev = (uint32_t)__shfl_up((int)gt, 1);
        uint64_t bounds = __ballot(lane == 0u || gt != prev);
        uint64_t le = lt | (1ull << lane);
        uint32_t start = 63u - (uint32_t)__clzll(bounds & le);
        uint64_t up = bounds & ~le;
        uint32_t end = up ? (uint32_t)(__ffsll((unsigned long long)up) - 1) : 64u;
        uint32_t lead = start, cnt = end - start, rank = lane - start;
        if (act && lead == lane) got[k] = atomicAdd(&d.s.tile_count[t], cnt);
        grp[k] = act ? ((1u << 16) | (rank << 8) | lead) : 0u;
    }
    uint32_t first = 0u, spilled = 0u;   // bit k: the k-th tile (starts an item; spilled)
    uint64_t items = 0u;                  // 4 bits per tile: the item it starts
#pragma unroll
    for (int k = 0; k < kMaxBinTiles; ++k) {
        if ((uint32_t)k >= nk) break;
        uint32_t g = grp[k];
        uint32_t lead = g & 0xFFu;
        uint32_t base = (uint32_t)__shfl((int)got[k], (int)lead);
        if (g >> 16) {
            uint32_t slot = base + ((g >> 8) & 0xFFu);
            uint32_t t = footprint_tile(d, tx0, ty0, w, (uint32_t)k);
            if (slot < d.s.bin_cap) d.s.bins[(size_t)t * d.s.bin_cap + slot] = key;
            else spilled |= 1u << k;
            uint32_t item;
            if (item_start(d, slot, item)) {
                first |= 1u << k;
                items |= (uint64_t)item << (4 * k);
            }
        }
    }
    uint32_t nf = (uint32_t)__popc(first), ns = (uint32_t)__popc(spilled);
    if (__ballot(nf != 0u)) {
        uint32_t at = wave_add(&d.s.header[H_ACTIVE], nf, nf != 0u);
        for (uint32_t m = first; m; m &= m - 1u) {
            uint32_t k = __ffs(m) - 1u;
            activate_item(d, at++, footprint_tile(d, tx0, ty0, w, k), (uint32_t)(items >> (4 * k)) & 15u);
        }
    }
    if (__ballot(ns != 0u)) {
        uint32_t at = wave_add(&d.s.header[H_SPILL], ns, ns != 0u);
        for (uint32_t m = spilled; m; m &= m - 1u)
            spill_entry(d, at++, footprint_tile(d, tx0, ty0, w, __ffs(m) - 1u), key);
    }
}

__device__ __forceinline__ uint32_t wave_max(uint32_t v)
{
    for (int off = 32; off > 0; off >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, off));
    return v;
}

// Compile-time loop: f(IC<I>{}) for I in [B, E).
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

// ------------------------------------------------------------------
// triangle records
// ------------------------------------------------------------------

__device__ __forceinline__ void write_record(uint32_t *rec, const RasterTri &r, const VsOut v[3], uint32_t key,
                                             bool front_facing)
{
    ScanTri s = scan_setup(r);
    uint32_t w[kRecWords];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        w[R_VP + 2 * k] = __float_as_uint(r.vpx[k]);
        w[R_VP + 2 * k + 1] = __float_as_uint(r.vpy[k]);
        w[R_Z + k] = __float_as_uint(r.z[k]);
        w[R_IW + k] = __float_as_uint(1.0f / v[k].clip[3]);
#pragma unroll
        for (int c = 0; c < kMaxVar; ++c) w[R_VAR + k * kMaxVar + c] = __float_as_uint(v[k].var[c]);
    }
    // constant terms of the three shoelace sums (the same products the
    // per-pixel evaluation would form)
    w[R_K + 0] = __float_as_uint((r.vpy[2] - r.vpy[1]) * (r.vpx[2] + r.vpx[1]));
    w[R_K + 1] = __float_as_uint((r.vpy[0] - r.vpy[2]) * (r.vpx[0] + r.vpx[2]));
    w[R_K + 2] = __float_as_uint((r.vpy[1] - r.vpy[0]) * (r.vpx[1] + r.vpx[0]));
    w[R_K + 3] = __float_as_uint(r.total);
    w[R_SCAN + 0] = (uint32_t)s.ax; w[R_SCAN + 1] = (uint32_t)s.ay;
    w[R_SCAN + 2] = (uint32_t)s.bx; w[R_SCAN + 3] = (uint32_t)s.by;
    w[R_SCAN + 4] = (uint32_t)s.cx; w[R_SCAN + 5] = (uint32_t)s.cy;
    w[R_MODE] = (uint32_t)s.mode | ((scan_wide(r) ? 1u : 0u) << 8) | ((front_facing ? 1u : 0u) << 9);
    w[R_KEY] = key;
    uint4 *o = reinterpret_cast<uint4 *>(rec);
#pragma unroll
    for (int q = 0; q < kRecWords / 4; ++q) o[q] = make_uint4(w[4 * q], w[4 * q + 1], w[4 * q + 2], w[4 * q + 3]);
}

// ------------------------------------------------------------------
// k_setup
// ------------------------------------------------------------------

// Grows the draw's big-triangle footprint union; kept as maxima of
// (65535 - min) and max.  Every lane of the wave calls it (same draw).
__device__ __forceinline__ void note_big_wave(const DrawParams &d, bool big, const uint4 &rec)
{
    if (!__ballot(big)) return;
    uint32_t nx0 = wave_max(big ? 65535u - (rec.y & 0xFFFFu) : 0u);
    uint32_t ny0 = wave_max(big ? 65535u - (rec.y >> 16) : 0u);
    uint32_t x1 = wave_max(big ? (rec.z & 0xFFFFu) : 0u);
    uint32_t y1 = wave_max(big ? (rec.z >> 16) : 0u);
    if (lane_id() == 0) {
        atomicOr(&d.s.header[H_BIG], 1u);
        atomicMax(&d.s.counters[C_BIG_NX0], nx0);
        atomicMax(&d.s.counters[C_BIG_NY0], ny0);
        atomicMax(&d.s.counters[C_BIG_X1], x1);
        atomicMax(&d.s.counters[C_BIG_Y1], y1);
    }
}

template <int PROG, bool ANY_DRAW, class DrawOf>
__device__ __forceinline__ void clip_wave(uint32_t di, uint32_t p, bool act, DrawOf &&draw_of);

// The LoadOp::Clear values into the whole target, spread over the setup
// grid (16-byte stores; issued before the setup work, they drain while it
// waits on its loads and atomics).
__device__ __forceinline__ void store_color(const DrawParams &d, size_t pix, uint32_t c)
{
    store_stream(reinterpret_cast<uint32_t *>(d.color) + pix, c);
}
__device__ __forceinline__ void store_depth(const DrawParams &d, size_t pix, float z)
{
    store_stream(reinterpret_cast<uint32_t *>(d.depth) + pix, __float_as_uint(z));
}

__device__ __forceinline__ void clear_target(const DrawParams &d)
{
    uint64_t n = (uint64_t)d.width * d.height;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (d.color && d.clear_color_on) {
        uint32_t *c = reinterpret_cast<uint32_t *>(d.color);
        for (uint64_t i = i0; i < n / 4; i += stride) store_stream4(c + 4 * i, d.clear_rgba);
        if (i0 < (n & 3u)) store_stream(c + (n & ~3ull) + i0, d.clear_rgba);
    }
    if (d.depth && d.clear_depth_on) {
        uint32_t z = __float_as_uint(d.clear_depth);
        uint32_t *p = reinterpret_cast<uint32_t *>(d.depth);
        for (uint64_t i = i0; i < n / 4; i += stride) store_stream4(p + 4 * i, z);
        if (i0 < (n & 3u)) store_stream(p + (n & ~3ull) + i0, z);
    }
}

// CLIP: the primitives that cross a clip plane are clipped here, by their
// wave (single draws whose waves all fit on the GPU at once: no k_clip
// launch); otherwise they go to the clip list.
template <int PROG, bool MULTI, bool CLIP = false>
__global__ __launch_bounds__(256) void k_setup(DrawParams d0, BatchArgs b)
{
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t di = 0, p = g;
    if constexpr (MULTI) {
        // draw ranges are padded to whole waves, so a wave never spans draws
        if (g >= b.total_prims) return;
        // every lane of the wave is in the same draw: look it up once, in
        // scalar registers, so that the draw's parameters are scalar loads
        di = __builtin_amdgcn_readfirstlane(b.wave_draw[__builtin_amdgcn_readfirstlane(g) >> 6]);
        p = g - b.prim_base[di];
    }
    const DrawParams &d = draw_at<MULTI>(d0, b, di);
    if constexpr (!MULTI) {
        if (b.preclear) clear_target(d);
        // whole waves past the end leave; the lanes of a partial wave stay
        // (inactive) for the wave-wide steps
        if (__builtin_amdgcn_readfirstlane(g) >= d.total_prims) return;
    }
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
                int cc = clip_class(v);
                if (cc == 0) {
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
                            write_record(d.s.tri + (size_t)p * kRecWords, r, v, p << 7, ff);
                        }
                    }
                } else if (cc == 1) {
                    needs_clip = true;
                }
            }
        }
    }
    bool small = has_rec && ntiles <= (uint32_t)kMaxBinTiles;
    bool big = has_rec && !small;
    uint32_t slot = 0;
    if constexpr (!CLIP) {
        slot = wave_append(&d.s.header[H_CLIP], needs_clip);
        if (needs_clip) {
            if (slot < d.s.clip_list_cap) d.s.clip_list[slot] = make_uint2(di, p);
            else atomicOr(d.s.error, (uint32_t)ERR_CLIP_OVERFLOW);
        }
    }
    bin_wave(d, small, rec.x, rec.y & 0xFFFFu, rec.y >> 16, rec.z & 0xFFFFu, rec.z >> 16);
    slot = wave_append(&d.s.counters[C_BIG], big);
    if (big) {
        if (slot < d.s.big_cap) d.s.big[slot] = rec;
        else atomicOr(d.s.error, (uint32_t)ERR_RECORD_OVERFLOW);
    }
    note_big_wave(d, big && slot < d.s.big_cap, rec);
    if constexpr (CLIP) {
        // every lane of the wave takes part (the clip slots are reserved per wave)
        if (__ballot(needs_clip))
            clip_wave<PROG, false>(di, p, needs_clip, [&](uint32_t) -> const DrawParams & { return d; });
    }
}

// ------------------------------------------------------------------
// k_clip
// ------------------------------------------------------------------

// Selects out[c] without indexing (keeps the triangles in registers).
__device__ __forceinline__ void pick(CTri &t, const CTri out[2], uint32_t c)
{
#pragma unroll
    for (int k = 0; k < 3; ++k) {
#pragma unroll
        for (int j = 0; j < 4; ++j) t.p[k][j] = c ? out[1].p[k][j] : out[0].p[k][j];
#pragma unroll
        for (int j = 0; j < 3; ++j) t.b[k][j] = c ? out[1].b[k][j] : out[0].b[k][j];
    }
}

// The clipper's output (clipper.rs:625-775) by split paths.  The reference
// clips a triangle against the six planes in turn, each plane turning every
// triangle of its queue into zero, one or two; its output order is
// lexicographic in the choices of piece at the levels where a triangle was
// split in two (plane 0 first).  Here a lane follows one path: it clips its
// triangle against the six planes in order and at each split takes the
// piece its next choice bit names (the first bit at the first split), with
// `bits` choices available.  The lanes of a group that take the choices 0
// .. 2^bits - 1 in rank order then hold every sub-triangle with at most
// `bits` splits on its path, in the clipper's order: two paths part at a
// level where both split after the same splits, so the order of the choice
// bits is the order of the paths.  A lane whose path dies (a plane leaves
// nothing) or does not exist (a choice of the second piece of a triangle
// that was not split, or a choice bit its path never used set) holds
// nothing; `over` marks a path that splits more often than `bits` allows.
// Each sub-triangle is computed by the same plane clips as the reference's.
struct ClipPath {
    CTri t;
    bool valid, over;
};

__device__ __forceinline__ ClipPath clip_path(const CTri &root, uint32_t rank, uint32_t bits, bool act)
{
    ClipPath r;
    r.t = root;
    r.valid = act;
    r.over = false;
    uint32_t used = 0;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        CTri out[2];
        uint32_t nn = (uint32_t)clip_against_plane(r.t, k, out);
        uint32_t ch = 0u;
        if (nn == 2u) {
            if (used < bits) ch = (rank >> (bits - 1u - used)) & 1u;
            else if (r.valid) r.over = true;
            ++used;
        }
        if (nn == 0u) r.valid = false;
        else pick(r.t, out, ch);
    }
    // choice bits past the path's splits must be 0 (one lane per path)
    uint32_t spare = used >= bits ? 0u : bits - used;
    r.valid = r.valid && !r.over && (rank & ((1u << spare) - 1u)) == 0u;
    return r;
}

// Primitives that cross a clip plane, clipped by one wave: the lanes with
// `act` hold them (draw di, primitive p).  Eight at a time, each by a group
// of eight lanes following its paths of up to three splits (clip_path); a
// primitive with a path of more splits is clipped again by the whole wave
// (64 lanes, six choice bits: every path).  A primitive's sub-triangles, in
// the clipper's order, become clip records (positions and clip
// barycentrics) and triangle records at consecutive clip slots, and are
// binned like k_setup's primitives, every lane of the wave together.  Every
// lane of the wave calls it.  ANY_DRAW: the primitives may belong to
// different draws of a batched group; `draw_of(di)` is draw di.
template <int PROG, bool ANY_DRAW, class DrawOf>
__device__ __forceinline__ void clip_wave(uint32_t di, uint32_t p, bool act, DrawOf &&draw_of)
{
    uint32_t lane = lane_id();
    uint64_t lt = lanes_below();
    uint64_t todo = __ballot(act), redo = 0u;
    while (todo | redo) {
        // the groups of this round and the lane holding each one's primitive
        uint32_t bits, gshift;
        int holder = -1;
        if (todo) {
            bits = 3u;
            gshift = 3u;
#pragma unroll
            for (uint32_t q = 0; q < 8u; ++q) {
                if (!todo) break;
                int h = __ffsll((unsigned long long)todo) - 1;
                todo &= todo - 1u;
                if ((lane >> 3) == q) holder = h;
            }
        } else {
            bits = 6u;
            gshift = 6u;
            holder = __ffsll((unsigned long long)redo) - 1;
            redo &= redo - 1u;
        }
        uint32_t rank = lane & ((1u << gshift) - 1u);
        uint32_t first = lane - rank;   // the group's first lane
        uint64_t gmask = gshift == 6u ? ~0ull : (0xFFull << first);
        bool has = holder >= 0;
        uint32_t gdi = (uint32_t)__shfl((int)di, has ? holder : 0);
        uint32_t gp = (uint32_t)__shfl((int)p, has ? holder : 0);
        const DrawParams &d = draw_of(gdi);

        VsOut v[3];
        CTri root{};
        bool ff = false;
        bool ok = has && shade_primitive<PROG>(d, gp, v);
        if (ok) {
            if constexpr (PROG == PROG_USER) {
                // @builtin(front_facing) of the sub-triangles (only a WGSL
                // fragment shader can read it)
                bool culled;
                (void)face_and_cull(d, v, ff, culled);
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
#pragma unroll
                for (int c = 0; c < 4; ++c) root.p[k][c] = v[k].clip[c];
#pragma unroll
                for (int c = 0; c < 3; ++c) root.b[k][c] = (c == k) ? 1.0f : 0.0f;
            }
        }
        ClipPath path = clip_path(root, rank, bits, ok);
        uint64_t overs = __ballot(path.over);
        bool group_over = (overs & gmask) != 0u;
        // primitives with a path of more splits go round again (by their holder lane)
        uint64_t again = __ballot(rank == 0u && group_over && has);
        for (uint64_t m = again; m; m &= m - 1u) {
            int l = __ffsll((unsigned long long)m) - 1;
            redo |= 1ull << lane_value((uint32_t)holder, l);
        }
        bool valid = path.valid && !group_over;
        uint64_t leaves = __ballot(valid);
        uint32_t nleaf = (uint32_t)__popcll(leaves & gmask);
        uint32_t leaf = (uint32_t)__popcll(leaves & gmask & lt);
        bool lead = rank == 0u && has && !group_over;
        if (lead && nleaf > 63u) {   // the record key holds 6 bits of sub-triangle index
            atomicOr(d.s.error, (uint32_t)ERR_CLIP_OVERFLOW);
        }
        bool take = lead && nleaf != 0u && nleaf <= 63u;
        uint32_t base = wave_add(&d.s.counters[C_CLIP_TRI], take ? nleaf : 0u, take);
        bool fits = take && base + nleaf <= d.s.clip_tri_cap;
        if (take && !fits) atomicOr(d.s.error, (uint32_t)ERR_CLIP_OVERFLOW);
        if (fits) d.s.clip_base[gp] = base;
        // the group's slot range, from its first lane
        base = (uint32_t)__shfl((int)base, (int)first);
        bool gfits = __shfl((int)(fits ? 1 : 0), (int)first) != 0;
        bool emit = valid && gfits;

        bool bin = false;
        uint32_t key = 0u, tx0 = 0u, ty0 = 0u, tx1 = 0u, ty1 = 0u;
        if (emit) {
            uint32_t slot = base + leaf;
            float *o = d.s.clip_tris + (size_t)slot * kClipFloats;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
#pragma unroll
                for (int c = 0; c < 4; ++c) o[k * 7 + c] = path.t.p[k][c];
#pragma unroll
                for (int c = 0; c < 3; ++c) o[k * 7 + 4 + c] = path.t.b[k][c];
            }
            RasterTri r;
            raster_setup(d, path.t.p, r);
            ScanTri s = scan_setup(r);
            if (!coord_ok(r)) {
                atomicOr(d.s.error, (uint32_t)ERR_COORD_RANGE);
            } else if (r.total != 0.0f && tile_bbox(d, s, tx0, ty0, tx1, ty1)) {
                key = (gp << 7) | (leaf << 1) | 1u;
                write_record(d.s.clip_tri + (size_t)slot * kRecWords, r, v, key, ff);
                bin = true;
            }
        }
        uint32_t nt = bin ? (tx1 - tx0 + 1u) * (ty1 - ty0 + 1u) : 0u;
        bool small = bin && nt <= (uint32_t)kMaxBinTiles;
        bool big = bin && !small;
        bin_wave<ANY_DRAW>(d, small, key, tx0, ty0, tx1, ty1);
        if (__ballot(big)) {
            uint32_t bs = wave_add(&d.s.counters[C_BIG], 1u, big);
            if (big) {
                if (bs < d.s.big_cap) {
                    d.s.big[bs] = make_uint4(key, tx0 | (ty0 << 16), tx1 | (ty1 << 16), gdi);
                    atomicOr(&d.s.header[H_BIG], 1u);
                    atomicMax(&d.s.counters[C_BIG_NX0], 65535u - tx0);
                    atomicMax(&d.s.counters[C_BIG_NY0], 65535u - ty0);
                    atomicMax(&d.s.counters[C_BIG_X1], tx1);
                    atomicMax(&d.s.counters[C_BIG_Y1], ty1);
                } else {
                    atomicOr(d.s.error, (uint32_t)ERR_RECORD_OVERFLOW);
                }
            }
        }
    }
}

// Primitives of the clip list per wave: one round of clip_wave's groups of
// eight lanes.
constexpr uint32_t kClipPerWave = 8;

// Every lane runs the same number of iterations, so that clip_wave's
// wave-wide steps see the whole wave.
template <int PROG, bool MULTI>
__global__ __launch_bounds__(256) void k_clip(DrawParams d0, BatchArgs b, const uint2 *clip_list, uint32_t cap)
{
    uint32_t n = min(b.header[H_CLIP], cap);
    uint32_t waves = gridDim.x * (blockDim.x / 64u);
    uint32_t wave = blockIdx.x * (blockDim.x / 64u) + threadIdx.x / 64u;
    for (uint32_t w0 = wave * kClipPerWave; w0 < n; w0 += waves * kClipPerWave) {
        uint32_t i = w0 + lane_id();
        bool act = lane_id() < kClipPerWave && i < n;
        uint2 e = act ? clip_list[i] : make_uint2(0u, 0u);
        clip_wave<PROG, MULTI>(e.x, e.y, act,
                               [&](uint32_t di) -> const DrawParams & { return draw_at<MULTI>(d0, b, di); });
    }
}

// ------------------------------------------------------------------
// what a tile holds
// ------------------------------------------------------------------

// Whether a tile lies in the footprint union of the draw's big triangles.
__device__ __forceinline__ bool in_big_union(const DrawParams &d, uint32_t tile)
{
    if (d.s.counters[C_BIG] == 0u) return false;
    uint32_t tx = tile % d.tiles_x, ty = tile / d.tiles_x;
    return tx >= 65535u - d.s.counters[C_BIG_NX0] && tx <= d.s.counters[C_BIG_X1] &&
           ty >= 65535u - d.s.counters[C_BIG_NY0] && ty <= d.s.counters[C_BIG_Y1];
}

// The sources of a work item's triangles, in the order the raster reads
// them: its range of the tile's bin; for the tile's last item the rest of
// the bin, the spill entries (only when the bin overflowed: the spill list
// is filtered by tile) and the big triangles whose footprint holds the tile.
struct TileFeed {
    uint32_t bin0, nbin, pos;     // bin: first slot, entries, next entry
    uint32_t nspill, spos;        // spill list: entries, next entry
    uint32_t nbig, bpos;          // big list: entries, next entry
    uint32_t gt, tx, ty;          // global tile, tile coordinates
    uint32_t items;               // the tile's work items
};

// From the tile's count of binned records and the group's spill and big
// counts.
__device__ __forceinline__ TileFeed tile_feed_from(const DrawParams &d, uint32_t tile, uint32_t item, uint32_t cnt,
                                                   uint32_t spill_n, uint32_t big_n)
{
    TileFeed f;
    uint32_t nbin = min(cnt, d.s.bin_cap), sp = d.s.split;
    f.items = 1u;
    if (sp) f.items = max(1u, min((nbin + sp - 1u) / sp, kMaxItems));
    item = min(item, f.items - 1u);
    bool last = item == f.items - 1u;
    uint32_t start = item * sp;
    f.bin0 = tile * d.s.bin_cap + start;
    f.nbin = last ? nbin - start : sp;
    f.pos = 0u;
    f.nspill = last && cnt > d.s.bin_cap ? min(spill_n, d.s.spill_cap) : 0u;
    f.spos = 0u;
    f.nbig = last ? min(big_n, d.s.big_cap) : 0u;
    f.bpos = 0u;
    f.gt = d.s.tile0 + tile;
    f.tx = tile % d.tiles_x;
    f.ty = tile / d.tiles_x;
    return f;
}

// The three counts are read together (none waits for another).
__device__ __forceinline__ TileFeed tile_feed(const DrawParams &d, uint32_t tile, uint32_t item = 0u)
{
    uint32_t cnt = d.s.tile_count[tile], spill_n = d.s.header[H_SPILL], big_n = d.s.counters[C_BIG];
    return tile_feed_from(d, tile, item, cnt, spill_n, big_n);
}

// Wave 0 compacts the next 64 entries of the spill or big list that belong
// to the tile into keys[0..); returns how many (in *nkeys for the workgroup).
__device__ __forceinline__ void feed_filtered(const DrawParams &d, TileFeed &f, bool spill, uint32_t *keys,
                                              uint32_t *nkeys)
{
    uint32_t tid = threadIdx.x;
    if (tid < 64u) {
        bool hit = false;
        uint32_t key = 0u;
        if (spill) {
            if (f.spos + tid < f.nspill) {
                uint2 e = d.s.spill[f.spos + tid];
                hit = e.x == f.gt;
                key = e.y;
            }
        } else if (f.bpos + tid < f.nbig) {
            uint4 r = d.s.big[f.bpos + tid];
            hit = f.tx >= (r.y & 0xFFFFu) && f.tx <= (r.z & 0xFFFFu) && f.ty >= (r.y >> 16) && f.ty <= (r.z >> 16);
            key = r.x;
        }
        uint64_t mask = __ballot(hit);
        if (hit) keys[__popcll(mask & lanes_below())] = key;
        if (tid == 0u) *nkeys = (uint32_t)__popcll(mask);
    }
    if (spill) f.spos += 64u;
    else f.bpos += 64u;
}

// The tile's next batch of at most nb keys (nb >= 64) into keys[]; 0 when
// the tile has no more.  Every thread of the workgroup calls it; the keys
// are visible to all of them on return, and the caller synchronises the
// workgroup before the next call.
__device__ __forceinline__ uint32_t next_batch(const DrawParams &d, TileFeed &f, uint32_t *keys, uint32_t *nkeys,
                                               uint32_t nb)
{
    if (f.pos < f.nbin) {
        uint32_t n = min(nb, f.nbin - f.pos);
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) keys[i] = d.s.bins[f.bin0 + f.pos + i];
        f.pos += n;
        __syncthreads();
        return n;
    }
    while (f.spos < f.nspill || f.bpos < f.nbig) {
        feed_filtered(d, f, f.spos < f.nspill, keys, nkeys);
        __syncthreads();
        uint32_t n = *nkeys;
        if (n) return n;
        __syncthreads();  // *nkeys is rewritten next round
    }
    return 0u;
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
    uint32_t ff;    // front facing
};

template <int TS>
struct RasterShared {
    TriLds tri[kBatch];
    ScanTri scan[kBatch];
    uint32_t wide[kBatch];
    uint32_t span[kBatch][TS];      // ordered path: per-row spans of the staged triangles
    uint32_t keys[kBatch];
    uint32_t nkeys;
};

// Loads one list entry's triangle record into LDS.
__device__ __forceinline__ void stage_record(const DrawParams &d, uint32_t key, TriLds &T, ScanTri &S, uint32_t &wide)
{
    uint32_t p = key >> 7, sub = (key >> 1) & 63u;
    bool clipped = key & 1u;
    uint32_t slot = clipped ? d.s.clip_base[p] + sub : 0u;
    const uint4 *r = reinterpret_cast<const uint4 *>(clipped ? d.s.clip_tri + (size_t)slot * kRecWords
                                                             : d.s.tri + (size_t)p * kRecWords);
    uint32_t w[kRecWords];
#pragma unroll
    for (int q = 0; q < kRecWords / 4; ++q) {
        uint4 v = r[q];
        w[4 * q] = v.x; w[4 * q + 1] = v.y; w[4 * q + 2] = v.z; w[4 * q + 3] = v.w;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        T.vpx[k] = __uint_as_float(w[R_VP + 2 * k]);
        T.vpy[k] = __uint_as_float(w[R_VP + 2 * k + 1]);
        T.z[k] = __uint_as_float(w[R_Z + k]);
        T.iw[k] = __uint_as_float(w[R_IW + k]);
#pragma unroll
        for (int c = 0; c < kMaxVar; ++c) T.var[k][c] = __uint_as_float(w[R_VAR + k * kMaxVar + c]);
    }
    T.k0 = __uint_as_float(w[R_K + 0]);
    T.k1 = __uint_as_float(w[R_K + 1]);
    T.k2 = __uint_as_float(w[R_K + 2]);
    T.total = __uint_as_float(w[R_K + 3]);
    T.key = key;
    if (clipped) {
        const float *ct = d.s.clip_tris + (size_t)slot * kClipFloats;
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int c = 0; c < 3; ++c) T.cb[k * 3 + c] = ct[k * 7 + 4 + c];
    } else {
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int c = 0; c < 3; ++c) T.cb[k * 3 + c] = (c == k) ? 1.0f : 0.0f;
    }
    S.ax = (int32_t)w[R_SCAN + 0]; S.ay = (int32_t)w[R_SCAN + 1];
    S.bx = (int32_t)w[R_SCAN + 2]; S.by = (int32_t)w[R_SCAN + 3];
    S.cx = (int32_t)w[R_SCAN + 4]; S.cy = (int32_t)w[R_SCAN + 5];
    S.mode = (int32_t)(w[R_MODE] & 0xFFu);
    S.pad = 0;
    wide = (w[R_MODE] >> 8) & 1u;
    T.ff = (w[R_MODE] >> 9) & 1u;
}

// Pins a value to a register: keeps the compiler from turning a select of
// two fields of the per-pixel state into a load through a selected address,
// which would move that state to scratch memory.
__device__ __forceinline__ float in_vgpr(float x)
{
    asm volatile("" : "+v"(x));
    return x;
}

struct PixelState {
    uint32_t col;
    float dep;
    float init_dep;
    uint32_t key;
    bool has;
};

// Barycentric at the pixel corner (raster.rs:313-336) composed with the
// clip barycentrics (raster.rs:337-343); returns the interpolated depth
// (raster.rs:345-346).
__device__ __forceinline__ float bary_depth(const TriLds &T, float pxf, float pyf, float lc[3])
{
    float l0 = (((T.vpy[1] - pyf) * (T.vpx[1] + pxf) + T.k0) + (pyf - T.vpy[2]) * (pxf + T.vpx[2])) / T.total;
    float l1 = (((pyf - T.vpy[0]) * (pxf + T.vpx[0]) + (T.vpy[2] - pyf) * (T.vpx[2] + pxf)) + T.k1) / T.total;
    float l2 = ((T.k2 + (pyf - T.vpy[1]) * (pxf + T.vpx[1])) + (T.vpy[0] - pyf) * (T.vpx[0] + pxf)) / T.total;
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
    return z;
}

// @builtin(position) of a fragment (raster.rs:148-155, 345-346): the
// composed barycentrics applied to the (clipped) triangle's viewport x, y,
// NDC z and perspective divisor 1/w, in the same order as the depth.
__device__ __forceinline__ void frag_position(const DrawParams &d, const TriLds &T, const float lc[3], float p[4])
{
    float iw[3];
    if (T.key & 1u) {
        uint32_t pr = T.key >> 7, sub = (T.key >> 1) & 63u;
        const float *ct = d.s.clip_tris + (size_t)(d.s.clip_base[pr] + sub) * kClipFloats;
#pragma unroll
        for (int k = 0; k < 3; ++k) iw[k] = 1.0f / ct[k * 7 + 3];
    } else {
#pragma unroll
        for (int k = 0; k < 3; ++k) iw[k] = T.iw[k];
    }
    float x = T.vpx[0] * lc[0], y = T.vpy[0] * lc[0], z = T.z[0] * lc[0], w = iw[0] * lc[0];
    x = x + T.vpx[1] * lc[1];
    y = y + T.vpy[1] * lc[1];
    z = z + T.z[1] * lc[1];
    w = w + iw[1] * lc[1];
    p[0] = x + T.vpx[2] * lc[2];
    p[1] = y + T.vpy[2] * lc[2];
    p[2] = z + T.z[2] * lc[2];
    p[3] = w + iw[2] * lc[2];
}

// Position provider of a triangle fragment for WGSL fragment shaders.
struct TriPos {
    const DrawParams &d;
    const TriLds &T;
    const float *lc;
    __device__ void operator()(float p[4]) const { frag_position(d, T, lc, p); }
};

// Varyings a built-in program's fragment shader reads (at most; the draw's
// nvar bounds them too).
template <int PROG>
struct VarUsed {
    static constexpr int value = PROG == PROG_MVP_TEXTURE ? 2 : kMaxVar;
};

// Varyings (fragment.rs:319-377) and the fragment shader.  Returns false
// when the fragment writes nothing (a WGSL shader discarded it or aborted).
// Only the varyings the shader can read are interpolated.
template <int PROG>
__device__ __forceinline__ bool fragment_color(const DrawParams &d, const TriLds &T, const float lc[3],
                                               float color[4], const float *srgb, float *frag_depth)
{
    constexpr int NV = VarUsed<PROG>::value;
    float var[kMaxVar];
#pragma unroll
    for (int c = 0; c < kMaxVar; ++c) var[c] = 0.0f;
#ifdef WGC_USER_PROGRAM
    if constexpr (PROG == PROG_USER) {
        // per word: flat (first vertex), perspective-corrected or linear
        float q0 = lc[0] * T.iw[0];
        float q1 = lc[1] * T.iw[1];
        float q2 = lc[2] * T.iw[2];
        float qs = (q0 + q1) + q2;
#pragma unroll
        for (int c = 0; c < kMaxVar; ++c) {
            if ((uint32_t)c >= d.nvar) break;
            if ((d.var_flat_mask >> c) & 1u) {
                var[c] = T.var[0][c];
            } else if ((d.var_persp_mask >> c) & 1u) {
                float acc = T.var[0][c] * q0;
                acc = acc + T.var[1][c] * q1;
                acc = acc + T.var[2][c] * q2;
                var[c] = acc / qs;
            } else {
                float acc = T.var[0][c] * lc[0];
                acc = acc + T.var[1][c] * lc[1];
                acc = acc + T.var[2][c] * lc[2];
                var[c] = acc;
            }
        }
        // primitive_index: the key's primitive within its instance
        uint32_t prim = (T.key >> 7) % d.nprim;
        uint32_t rc = user_fragment(d, TriPos{d, T, lc}, T.ff != 0u, prim, var, color, frag_depth);
        if (rc == 0u) return true;
        if (rc != 3u) atomicOr(d.s.error, (uint32_t)ERR_SHADER);   // 3: discard
        return false;
    }
#endif
    (void)frag_depth;
    if (d.interp == INTERP_FLAT) {
#pragma unroll
        for (int c = 0; c < NV; ++c) var[c] = T.var[0][c];
    } else if (d.interp == INTERP_PERSPECTIVE) {
        float q0 = lc[0] * T.iw[0];
        float q1 = lc[1] * T.iw[1];
        float q2 = lc[2] * T.iw[2];
        float qs = (q0 + q1) + q2;
#pragma unroll
        for (int c = 0; c < NV; ++c) {
            if ((uint32_t)c >= d.nvar) break;
            float acc = T.var[0][c] * q0;
            acc = acc + T.var[1][c] * q1;
            acc = acc + T.var[2][c] * q2;
            var[c] = acc / qs;
        }
    } else {
#pragma unroll
        for (int c = 0; c < NV; ++c) {
            if ((uint32_t)c >= d.nvar) break;
            float acc = T.var[0][c] * lc[0];
            acc = acc + T.var[1][c] * lc[1];
            acc = acc + T.var[2][c] * lc[2];
            var[c] = acc;
        }
    }
    run_fs<PROG>(d, var, color, srgb);
    return true;
}

// The depth tests around a fragment's shader, in API order against the
// running depth `dep` (fragment.rs:127-214, 427-488): with
// @early_depth_test the test (and write) before the shader -- force keeps
// its result, allow tests again after the shader; the late test runs when
// the colour output is written, on the frag_depth output if the shader wrote
// one before its colour, else the interpolated depth.  A discarded fragment
// keeps what an early test wrote.  `shade(&frag_depth)` runs the shader and
// returns false when it writes nothing.  Returns whether the colour is
// written.  Without early test or shader-written depth, the test comes
// first (it does not depend on the shader) so that failing fragments skip
// the shader.
template <class Shade>
__device__ __forceinline__ bool depth_and_shade(const DrawParams &d, float z, float &dep, Shade shade)
{
    if (!d.fs_early && !d.fs_depth_tested) {
        if (d.depth_test && !compare(d.depth_compare, z, dep)) return false;
        float unused = z;
        if (!shade(&unused)) return false;
        if (d.depth_test && d.depth_write) dep = z;
        return true;
    }
    if (d.fs_early && d.depth_test) {
        if (!compare(d.depth_compare, z, dep)) return false;
        if (d.depth_write) dep = z;
    }
    float frag_depth = z;
    if (!shade(&frag_depth)) return false;
    if (d.fs_early != 1u && d.depth_test) {
        if (!compare(d.depth_compare, frag_depth, dep)) return false;
        if (d.depth_write) dep = frag_depth;
    }
    return true;
}

// One fragment of triangle T at pixel (pxf, pyf).  Ordered: the depth test
// against the running depth, the fragment shader and the colour write or
// blend (fragment.rs:427-488, texture.rs:213-218), in API order.  Otherwise
// only visibility: the order-independent restatement of the depth test picks
// the winning (depth, key); its colour is shaded once after the tile's last
// triangle (the fragment shader has no side effects, so only the winner's
// output can reach the target).
template <int PROG, bool ORDERED>
__device__ __forceinline__ void shade_fragment(const DrawParams &d, const TriLds &T, float pxf, float pyf,
                                               PixelState &ps)
{
    float lc[3];
    float z = bary_depth(T, pxf, pyf, lc);
    if (ORDERED) {
        float color[4];
        bool ok = depth_and_shade(d, z, ps.dep, [&](float *frag_depth) {
            return fragment_color<PROG>(d, T, lc, color, d.srgb_lut, frag_depth);
        });
        if (!ok) return;
        ps.col = d.blend ? blend_over(color, ps.col, d.color_bgra) : pack_color(color, d.color_bgra);
        ps.has = true;
        return;
    }
    bool win;
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
    if (win) {
        ps.dep = z;
        ps.key = T.key;
        ps.has = true;
    }
}

// Colour and depth of a pixel's winning fragment: its triangle record is
// read back (from L2, shared by the neighbouring pixels it won) and the same
// barycentrics are evaluated again, bit for bit.
template <int PROG>
__device__ __forceinline__ uint32_t winner_color(const DrawParams &d, uint32_t key, float pxf, float pyf, float &z,
                                                 const float *srgb)
{
    TriLds T;
    ScanTri S;
    uint32_t wide;
    stage_record(d, key, T, S, wide);
    float lc[3];
    z = bary_depth(T, pxf, pyf, lc);
    float color[4];
    // pipelines whose fragment shader can discard use the ordered path, so
    // the winner always writes (an aborted invocation raised ERR_SHADER)
    float unused = z;
    (void)fragment_color<PROG>(d, T, lc, color, srgb, &unused);
    return pack_color(color, d.color_bgra);
}

// Order-independent resolve as a minimum over 64-bit values: the depth in
// an unsigned encoding that keeps IEEE order (with -0 equal to +0, as the
// comparison treats them) in the high word, the primitive key (or its
// complement, to prefer the later fragment) in the low word.  ~0 means no
// fragment; no candidate encodes to it (NaN depths never pass the test).
__device__ __forceinline__ uint32_t depth_order(float z)
{
    uint32_t u = __float_as_uint(z);
    if ((u << 1) == 0u) u = 0u;
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ unsigned long long oi_value(uint32_t resolve, float z, uint32_t key)
{
    switch (resolve) {
    case RES_MIN_FIRST: return ((unsigned long long)depth_order(z) << 32) | key;
    case RES_MIN_LAST: return ((unsigned long long)depth_order(z) << 32) | (~key);
    case RES_MAX_FIRST: return ((unsigned long long)(~depth_order(z)) << 32) | key;
    case RES_MAX_LAST: return ((unsigned long long)(~depth_order(z)) << 32) | (~key);
    default: return (unsigned long long)(~key);  // last in API order
    }
}

__device__ __forceinline__ uint32_t oi_key(uint32_t resolve, unsigned long long v)
{
    uint32_t lo = (uint32_t)v;
    return (resolve == RES_MIN_FIRST || resolve == RES_MAX_FIRST) ? lo : ~lo;
}

// The depth test of the order-independent modes, against the depth the
// pixel held before the draw.
__device__ __forceinline__ bool oi_pass(const DrawParams &d, float z, float init)
{
    switch (d.resolve) {
    case RES_LAST: return true;
    case RES_MIN_FIRST: return z < init;
    case RES_MIN_LAST: return z <= init;
    case RES_MAX_FIRST: return z > init;
    case RES_MAX_LAST: return z >= init;
    default: return compare(d.depth_compare, z, init);  // RES_PASS_LAST
    }
}

// Stages sh.keys[0..n) into LDS and applies them to this lane's pixels.
template <int PROG, bool ORDERED, int TS, int NT>
__device__ __forceinline__ void raster_batch(const DrawParams &d, RasterShared<TS> &sh, uint32_t n, uint32_t x0,
                                             uint32_t y0, PixelState (&ps)[TS * TS / NT],
                                             uint32_t (&nfrag)[TS * TS / NT])
{
    constexpr int PPT = TS * TS / NT;
    uint32_t tid = threadIdx.x;
    if (tid < n) stage_record(d, sh.keys[tid], sh.tri[tid], sh.scan[tid], sh.wide[tid]);
    __syncthreads();
    // per-row spans inside the scissor and the tile, tile-relative
    int64_t sx0 = d.scissor[0], sy0 = d.scissor[1];
    int64_t sx1 = sx0 + d.scissor[2], sy1 = sy0 + d.scissor[3];
    for (uint32_t idx = tid; idx < n * (uint32_t)TS; idx += NT) {
        uint32_t e = idx / TS, row = idx % TS;
        int32_t y = (int32_t)(y0 + row);
        uint32_t span = 1u;  // empty: lo = 1, hi = 0
        int32_t x1, x2;
        if ((int64_t)y >= sy0 && (int64_t)y < sy1 && scan_span(sh.scan[e], y, x1, x2, sh.wide[e] != 0u)) {
            int64_t lo = max((int64_t)x1, max(sx0, (int64_t)x0));
            int64_t hi = min((int64_t)x2, min(sx1 - 1, (int64_t)x0 + TS - 1));
            if (lo <= hi) span = (uint32_t)(lo - x0) | ((uint32_t)(hi - x0) << 16);
        }
        sh.span[e][row] = span;
    }
    __syncthreads();
    // pixels outer (compile-time, so the per-pixel state stays in registers),
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

// fn(key) for every key of the tile (its bin slots, its entries of the
// spill list, the big triangles whose footprint holds it), each key by one
// thread of the workgroup.
template <class F>
__device__ __forceinline__ void for_tile_keys(const DrawParams &d, const TileFeed &f, F &&fn)
{
    for (uint32_t i = threadIdx.x; i < f.nbin; i += blockDim.x) fn(d.s.bins[f.bin0 + i]);
    for (uint32_t i = threadIdx.x; i < f.nspill; i += blockDim.x) {
        uint2 e = d.s.spill[i];
        if (e.x == f.gt) fn(e.y);
    }
    for (uint32_t i = threadIdx.x; i < f.nbig; i += blockDim.x) {
        uint4 r = d.s.big[i];
        if (f.tx >= (r.y & 0xFFFFu) && f.tx <= (r.z & 0xFFFFu) && f.ty >= (r.y >> 16) && f.ty <= (r.z >> 16))
            fn(r.x);
    }
}

// The r-th smallest (from 1) of the tile's keys at or above lo, when at
// least r of them are: a radix select over four 8-bit digits, one pass over
// the keys per digit with a histogram in LDS (`hist`: 256 counters and two
// words for the choice).  Every thread of the workgroup calls it.
__device__ __forceinline__ uint32_t tile_key_select(const DrawParams &d, const TileFeed &f, uint64_t lo, uint32_t r,
                                                    uint32_t *hist)
{
    uint32_t prefix = 0u, pmask = 0u;
    for (int shift = 24; shift >= 0; shift -= 8) {
        for (uint32_t i = threadIdx.x; i < 256u; i += blockDim.x) hist[i] = 0u;
        __syncthreads();
        for_tile_keys(d, f, [&](uint32_t key) {
            if ((uint64_t)key >= lo && (key & pmask) == prefix) atomicAdd(&hist[(key >> shift) & 255u], 1u);
        });
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t acc = 0u, sel = 255u;
            for (uint32_t v = 0; v < 256u; ++v) {
                if (acc + hist[v] >= r) { sel = v; break; }
                acc += hist[v];
            }
            hist[256] = sel;
            hist[257] = acc;
        }
        __syncthreads();
        uint32_t sel = hist[256], below = hist[257];
        __syncthreads();  // the histogram is cleared next round
        prefix |= sel << shift;
        pmask |= 255u << shift;
        r -= below;
    }
    return prefix;
}

// Sorts olist[0..total) (total <= kOrderedCap) into ascending order: a
// bitonic sort padded with UINT32_MAX to a power of two.  Up to 64 keys
// (most tiles) are sorted by one wave in registers, the network running over
// its lanes with shuffles and no barrier between steps; longer lists by the
// workgroup in LDS.
template <int NT>
__device__ __forceinline__ void sort_olist(uint32_t *olist, uint32_t total)
{
    uint32_t tid = threadIdx.x;
    if (total <= 64u) {
        if (tid < 64u) {
            uint32_t v = tid < total ? olist[tid] : 0xFFFFFFFFu;
#pragma unroll
            for (uint32_t k = 2; k <= 64u; k <<= 1) {
#pragma unroll
                for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                    uint32_t o = (uint32_t)__shfl_xor((int)v, (int)j);
                    bool up = (tid & k) == 0u, lower = (tid & j) == 0u;
                    v = lower == up ? min(v, o) : max(v, o);
                }
            }
            if (tid < total) olist[tid] = v;
        }
        __syncthreads();
        return;
    }
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

// One tile of an ordered draw: load or clear, apply its triangles in API
// order, store.
template <int PROG, int TS, int NT>
__device__ __forceinline__ void raster_tile(const DrawParams &d, uint32_t tile, RasterShared<TS> &sh,
                                            uint32_t *olist, uint32_t &olist_n, uint32_t &frag_acc)
{
    constexpr bool ORDERED = true;
    constexpr int PPT = TS * TS / NT;
    uint32_t x0 = (tile % d.tiles_x) * TS, y0 = (tile / d.tiles_x) * TS;
    uint32_t tid = threadIdx.x;
    TileFeed feed = tile_feed(d, tile);

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

    // gather the tile's keys (bin, spill entries, overlapping big
    // triangles), then sort them into API order
    if (tid == 0) olist_n = 0u;
    __syncthreads();
    bool overflow = feed.nbin > (uint32_t)kOrderedCap;
    if (!overflow) {
        for (uint32_t i = tid; i < feed.nbin; i += NT) olist[i] = d.s.bins[feed.bin0 + i];
        if (tid == 0) olist_n = feed.nbin;
    }
    __syncthreads();
    while (!overflow && (feed.spos < feed.nspill || feed.bpos < feed.nbig)) {
        feed_filtered(d, feed, feed.spos < feed.nspill, sh.keys, &sh.nkeys);
        __syncthreads();
        uint32_t n = sh.nkeys, base = olist_n;
        if (base + n > (uint32_t)kOrderedCap) overflow = true;
        else if (tid < n) olist[base + tid] = sh.keys[tid];
        __syncthreads();  // every thread has read olist_n and sh.nkeys
        if (tid == 0 && !overflow) olist_n = base + n;
        __syncthreads();
    }
    // more keys than the list holds: the tile's keys in ascending chunks of
    // at most kOrderedCap, each chunk's end found by a radix select over the
    // keys (chunks in key order are API order; the pixels' state carries
    // over from one chunk to the next)
    uint64_t lo = 0u, hi = 1ull << 32;
    if (overflow) {
        __syncthreads();  // every thread has read olist_n
        hi = (uint64_t)tile_key_select(d, feed, 0u, (uint32_t)kOrderedCap, olist) + 1u;
    }
    for (;;) {
        if (overflow) {
            if (tid == 0) olist_n = 0u;
            __syncthreads();
            for_tile_keys(d, feed, [&](uint32_t key) {
                if ((uint64_t)key >= lo && (uint64_t)key < hi) {
                    uint32_t at = atomicAdd(&olist_n, 1u);
                    if (at < (uint32_t)kOrderedCap) olist[at] = key;   // (keys are distinct: at most kOrderedCap)
                }
            });
            __syncthreads();
        }
        uint32_t total = min(olist_n, (uint32_t)kOrderedCap);
        sort_olist<NT>(olist, total);
        for (uint32_t fpos = 0; fpos < total;) {
            uint32_t n = min((uint32_t)kBatch, total - fpos);
            if (tid < n) sh.keys[tid] = olist[fpos + tid];
            fpos += n;
            raster_batch<PROG, ORDERED, TS, NT>(d, sh, n, x0, y0, ps, nfrag);
        }
        if (!overflow || hi > 0xFFFFFFFFull) break;
        // the next chunk: the keys from hi on, at most kOrderedCap of them
        lo = hi;
        if (tid == 0) olist_n = 0u;
        __syncthreads();
        for_tile_keys(d, feed, [&](uint32_t key) {
            if ((uint64_t)key >= lo) atomicAdd(&olist_n, 1u);
        });
        __syncthreads();
        uint32_t rest = olist_n;
        __syncthreads();  // every thread has read olist_n
        if (rest == 0u) break;
        hi = rest <= (uint32_t)kOrderedCap ? (1ull << 32)
                                           : (uint64_t)tile_key_select(d, feed, lo, (uint32_t)kOrderedCap, olist) + 1u;
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
            float zw;
            ps[k].col = ps[k].has ? winner_color<PROG>(d, ps[k].key, (float)px, (float)py, zw, d.srgb_lut)
                                  : init_col[k];
        }
        if (d.color) store_color(d, pix, ps[k].col);
        if (d.depth) store_depth(d, pix, write_depth ? in_vgpr(ps[k].dep) : in_vgpr(ps[k].init_dep));
        if (d.s.frag_px && nfrag[k]) d.s.frag_px[pix] += nfrag[k];
    });
    frag_acc += f;
}

// What the visibility pass needs of a triangle: its record without the
// varyings (those are read again when the winner is shaded), kept in LDS as
// structure-of-arrays (lanes that read different triangles hit different
// banks).
enum VisField {
    V_VPX = 0, V_VPY = 3, V_K = 6, V_Z = 10, V_CB = 13, V_NUM = 22,
};
enum VisScan { S_AX = 0, S_AY, S_BX, S_BY, S_CX, S_CY, S_MODE, S_NUM };

template <int TS, int NT>
struct VisBatch {
    static constexpr int value = (int)vis_batch_size(TS, NT);
};

// Waves per SIMD the tile kernel's registers are budgeted for (the second
// __launch_bounds__ argument): six, i.e. six workgroups per CU for 16x16
// tiles of 256 threads and three for 32x32 tiles of 512; other shapes as the
// compiler sees fit.
template <int TS, int NT>
struct TilesOccupancy {
    static constexpr int value = ((NT == 256 && TS == 16) || NT == 512) ? 6 : 1;
};

template <int TS, int NT>
struct VisShared {
    static constexpr int NB = VisBatch<TS, NT>::value;
    float f[V_NUM][NB];
    int32_t sc[S_NUM][NB];
    uint32_t tkey[NB];
    uint32_t rows[NB];          // first covered tile row | count << 8 | wide << 16
    uint32_t owner[NB * TS];    // per covered row: triangle | row << 8 | span start << 16 | length << 24
    uint32_t flist[2 * NT];     // a window of the batch's fragments: owner entry | pixel in span << 16
    uint32_t keys[NB];
    unsigned long long best[TS * TS];  // encoded winner (see oi_value), ~0 = none
    float idep[TS * TS];               // depth before the draw
    uint32_t wsum[NT / 64];
    uint32_t nkeys;
    uint32_t last;                     // split tiles: this workgroup finished the tile's last item
};

// Spreads consecutive items over the waves of the workgroup (item i goes to
// wave i mod (NT/64)), so that a short batch keeps every SIMD busy.
template <int NT>
__device__ __forceinline__ uint32_t spread(uint32_t tid)
{
    constexpr uint32_t W = NT / 64;
    return (tid & 63u) * W + (tid >> 6);
}

template <int TS, int NT>
__device__ __forceinline__ void stage_vis(const DrawParams &d, uint32_t key, VisShared<TS, NT> &sh, uint32_t e,
                                          ScanTri &S, uint32_t &wide)
{
    uint32_t p = key >> 7, sub = (key >> 1) & 63u;
    bool clipped = key & 1u;
    uint32_t slot = clipped ? d.s.clip_base[p] + sub : 0u;
    const uint4 *r = reinterpret_cast<const uint4 *>(clipped ? d.s.clip_tri + (size_t)slot * kRecWords
                                                             : d.s.tri + (size_t)p * kRecWords);
    static_assert(R_VAR == 24, "the visibility fields are the record's first 24 words");
    uint32_t w[24];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        uint4 v = r[q];
        w[4 * q] = v.x; w[4 * q + 1] = v.y; w[4 * q + 2] = v.z; w[4 * q + 3] = v.w;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        sh.f[V_VPX + k][e] = __uint_as_float(w[R_VP + 2 * k]);
        sh.f[V_VPY + k][e] = __uint_as_float(w[R_VP + 2 * k + 1]);
        sh.f[V_Z + k][e] = __uint_as_float(w[R_Z + k]);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) sh.f[V_K + k][e] = __uint_as_float(w[R_K + k]);
    sh.tkey[e] = key;
    if (clipped) {
        const float *ct = d.s.clip_tris + (size_t)slot * kClipFloats;
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int c = 0; c < 3; ++c) sh.f[V_CB + k * 3 + c][e] = ct[k * 7 + 4 + c];
    } else {
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int c = 0; c < 3; ++c) sh.f[V_CB + k * 3 + c][e] = (c == k) ? 1.0f : 0.0f;
    }
    S.ax = (int32_t)w[R_SCAN + 0]; S.ay = (int32_t)w[R_SCAN + 1];
    S.bx = (int32_t)w[R_SCAN + 2]; S.by = (int32_t)w[R_SCAN + 3];
    S.cx = (int32_t)w[R_SCAN + 4]; S.cy = (int32_t)w[R_SCAN + 5];
    S.mode = (int32_t)(w[R_MODE] & 0xFFu);
    S.pad = 0;
    sh.sc[S_AX][e] = S.ax; sh.sc[S_AY][e] = S.ay;
    sh.sc[S_BX][e] = S.bx; sh.sc[S_BY][e] = S.by;
    sh.sc[S_CX][e] = S.cx; sh.sc[S_CY][e] = S.cy;
    sh.sc[S_MODE][e] = S.mode;
    wide = (w[R_MODE] >> 8) & 1u;
}

// The same arithmetic as bary_depth, on a staged triangle.
template <int TS, int NT>
__device__ __forceinline__ float vis_depth(const VisShared<TS, NT> &sh, uint32_t e, float pxf, float pyf)
{
    float vx0 = sh.f[V_VPX][e], vx1 = sh.f[V_VPX + 1][e], vx2 = sh.f[V_VPX + 2][e];
    float vy0 = sh.f[V_VPY][e], vy1 = sh.f[V_VPY + 1][e], vy2 = sh.f[V_VPY + 2][e];
    float k0 = sh.f[V_K][e], k1 = sh.f[V_K + 1][e], k2 = sh.f[V_K + 2][e], total = sh.f[V_K + 3][e];
    float l0 = (((vy1 - pyf) * (vx1 + pxf) + k0) + (pyf - vy2) * (pxf + vx2)) / total;
    float l1 = (((pyf - vy0) * (pxf + vx0) + (vy2 - pyf) * (vx2 + pxf)) + k1) / total;
    float l2 = ((k2 + (pyf - vy1) * (pxf + vx1)) + (vy0 - pyf) * (vx0 + pxf)) / total;
    float lc[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float acc = sh.f[V_CB + 0 * 3 + c][e] * l0;
        acc = acc + sh.f[V_CB + 1 * 3 + c][e] * l1;
        acc = acc + sh.f[V_CB + 2 * 3 + c][e] * l2;
        lc[c] = acc;
    }
    float z = sh.f[V_Z][e] * lc[0];
    z = z + sh.f[V_Z + 1][e] * lc[1];
    z = z + sh.f[V_Z + 2][e] * lc[2];
    return z;
}

// Stages sh.keys[0..n) and merges their fragments into the tile's winners.
// Each staged triangle's covered rows (the reference's scanline rows, within
// the tile and the scissor) are compacted into one list; one lane per
// (triangle, row) computes the row's integer span; the spans' fragments are
// then listed and spread evenly over the lanes.  Returns the fragments of
// this thread's entries (the workgroup's sum is the batch's fragments).
template <int TS, int NT>
__device__ __forceinline__ uint32_t vis_batch(const DrawParams &d, VisShared<TS, NT> &sh, uint32_t n, uint32_t x0,
                                              uint32_t y0)
{
    uint32_t tid = threadIdx.x;
    uint32_t e = tid;  // the triangle this thread stages (a batch takes as few waves as it needs)
    int64_t sx0 = d.scissor[0], sy0 = d.scissor[1];
    int64_t sx1 = sx0 + d.scissor[2], sy1 = sy0 + d.scissor[3];
    uint32_t cnt_rows = 0, first = 0;
    if (e < n) {
        uint32_t wide;
        ScanTri S;
        stage_vis<TS, NT>(d, sh.keys[e], sh, e, S, wide);
        int32_t r0, r1;
        scan_rows(S, r0, r1);
        int64_t lo = max((int64_t)r0, max((int64_t)y0, sy0));
        int64_t hi = min((int64_t)r1, min((int64_t)y0 + TS - 1, sy1 - 1));
        if (lo <= hi) {
            cnt_rows = (uint32_t)(hi - lo + 1);
            first = (uint32_t)(lo - y0);
        }
        sh.rows[e] = first | (cnt_rows << 8) | (wide << 16);
    }
    // exclusive scan of the row counts in thread order: within each wave,
    // then across waves
    uint32_t lane = lane_id(), wave = tid >> 6;
    uint32_t v = wave_inclusive_sum(cnt_rows);
    if (lane == 63u) sh.wsum[wave] = v;
    __syncthreads();
    uint32_t before = 0, total = 0;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) {
        uint32_t x = sh.wsum[w];
        if ((uint32_t)w < wave) before += x;
        total += x;
    }
    if (e < n) {
        uint32_t b = before + v - cnt_rows;
        for (uint32_t j = 0; j < cnt_rows; ++j) sh.owner[b + j] = e | ((first + j) << 8);
    }
    __syncthreads();

    // the span of every (triangle, row) entry, tile-relative; this thread's
    // entries are tid, tid + NT, ... and its fragments follow each other.
    // Debug fragment counts (rare) go straight to the per-pixel buffer.
    uint32_t *counts = d.s.frag_px;
    uint32_t nfr = 0;
    for (uint32_t idx = tid; idx < total; idx += NT) {
        uint32_t o = sh.owner[idx];
        uint32_t te = o & 0xFFu, row = (o >> 8) & 0xFFu;
        ScanTri S;
        S.ax = sh.sc[S_AX][te]; S.ay = sh.sc[S_AY][te];
        S.bx = sh.sc[S_BX][te]; S.by = sh.sc[S_BY][te];
        S.cx = sh.sc[S_CX][te]; S.cy = sh.sc[S_CY][te];
        S.mode = sh.sc[S_MODE][te];
        S.pad = 0;
        int32_t x1, x2;
        uint32_t start = 0, len = 0;
        if (scan_span(S, (int32_t)(y0 + row), x1, x2, ((sh.rows[te] >> 16) & 1u) != 0u)) {
            int64_t lo = max((int64_t)x1, max(sx0, (int64_t)x0));
            int64_t hi = min((int64_t)x2, min(sx1 - 1, (int64_t)x0 + TS - 1));
            if (lo <= hi) {
                start = (uint32_t)(lo - x0);
                len = (uint32_t)(hi - lo) + 1u;
            }
        }
        nfr += len;
        if (d.resolve == RES_LAST) {
            // no depth test: the last fragment in API order wins, depth
            // unused; one atomic per fragment, applied right here
            uint32_t q0 = row * TS + start;
            unsigned long long val = (unsigned long long)(~sh.tkey[te]);
            for (uint32_t i = 0; i < len; ++i) {
                if (counts) atomicAdd(&counts[(size_t)(y0 + row) * d.width + x0 + start + i], 1u);
                atomicMin(&sh.best[q0 + i], val);
            }
            continue;
        }
        sh.owner[idx] = te | (row << 8) | (start << 16) | (len << 24);
    }
    if (d.resolve == RES_LAST) {
        __syncthreads();
        return nfr;
    }
    // exclusive scan of the fragment counts in thread order
    uint32_t c = wave_inclusive_sum(nfr);
    if (lane == 63u) sh.wsum[wave] = c;  // every thread read the row-count sums before the last barrier
    __syncthreads();
    uint32_t fbase = c - nfr, nfrag = 0;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) {
        uint32_t x = sh.wsum[w];
        if ((uint32_t)w < wave) fbase += x;
        nfrag += x;
    }

    // the fragments, a window at a time: each entry lists its fragments in
    // the window, then every lane takes fragments of the list (consecutive
    // ones spread over the waves), so that long and short spans cost alike
    constexpr uint32_t W = 2 * NT;
    for (uint32_t w0 = 0; w0 < nfrag; w0 += W) {
        uint32_t pos = fbase;
        for (uint32_t idx = tid; idx < total; idx += NT) {
            uint32_t len = sh.owner[idx] >> 24;
            uint32_t a = max(pos, w0), bnd = min(pos + len, w0 + W);
            for (uint32_t q = a; q < bnd; ++q) sh.flist[q - w0] = idx | ((q - pos) << 16);
            pos += len;
        }
        __syncthreads();
        uint32_t nw = min(W, nfrag - w0);
        for (uint32_t k = spread<NT>(tid); k < nw; k += NT) {
            uint32_t fl = sh.flist[k];
            uint32_t o = sh.owner[fl & 0xFFFFu];
            uint32_t te = o & 0xFFu, row = (o >> 8) & 0xFFu;
            uint32_t px = ((o >> 16) & 0xFFu) + (fl >> 16);
            uint32_t q = row * TS + px;
            if (counts) atomicAdd(&counts[(size_t)(y0 + row) * d.width + x0 + px], 1u);
            uint32_t key = sh.tkey[te];
            float z = vis_depth<TS, NT>(sh, te, (float)(x0 + px), (float)(y0 + row));
            if (oi_pass(d, z, sh.idep[q])) atomicMin(&sh.best[q], oi_value(d.resolve, z, key));
        }
        __syncthreads();
    }
    return nfr;
}

// One tile of an order-independent draw: the depth the pixels hold before
// the draw (cleared or loaded), the winners of the tile's triangles (its
// bin, spill entries and overlapping big triangles) as a minimum per pixel
// in LDS (encoded as in oi_value), then each winner shaded once: its
// record is read back (from L2, shared by the neighbouring pixels it won)
// and the same barycentrics are evaluated again, bit for bit.  Colour and
// depth are stored where the pass changes them (a winner, or the
// LoadOp::Clear values).
template <int PROG, int TS, int NT>
__device__ __forceinline__ void raster_tile_oi(const DrawParams &d, uint32_t tile, uint32_t item,
                                               VisShared<TS, NT> &sh, const float *srgb, uint32_t &frag_acc)
{
    constexpr int PPT = TS * TS / NT;
    constexpr uint32_t NB = VisShared<TS, NT>::NB;
    uint32_t x0 = (tile % d.tiles_x) * TS, y0 = (tile / d.tiles_x) * TS;
    uint32_t tid = threadIdx.x;
    // the item's first keys, the tile's count and the group's spill and big
    // counts, all in flight together (key slots are clamped to the bin's
    // capacity, at least 1; those at or past the count are not used)
    uint32_t slot0 = min(item * d.s.split + tid, d.s.bin_cap - 1u);
    uint32_t k0 = d.s.bins[tile * d.s.bin_cap + slot0];
    uint32_t cnt = d.s.tile_count[tile], spill_n = d.s.header[H_SPILL], big_n = d.s.counters[C_BIG];
    asm volatile("" : : "v"(k0), "v"(cnt), "v"(spill_n), "v"(big_n));
    if (tid < NB) sh.keys[tid] = k0;
    TileFeed feed = tile_feed_from(d, tile, item, cnt, spill_n, big_n);

#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        uint32_t q = tid + (uint32_t)k * NT;
        uint32_t px = x0 + q % TS, py = y0 + q / TS;
        float z = 0.0f;
        if (d.depth && px < d.width && py < d.height)
            z = (d.init_clear && d.clear_depth_on) ? d.clear_depth : d.depth[(size_t)py * d.width + px];
        sh.idep[q] = z;
        sh.best[q] = ~0ull;
    }
    // the first batch: the slots read above (the same as next_batch's)
    bool pre = item < feed.items && feed.nbin != 0u;
    uint32_t n = pre ? min(NB, feed.nbin) : 0u;
    feed.pos = n;
    __syncthreads();

    uint32_t f = 0;
    for (;;) {
        if (!pre) n = next_batch(d, feed, sh.keys, &sh.nkeys, NB);
        pre = false;
        if (n == 0u) break;
        f += vis_batch<TS, NT>(d, sh, n, x0, y0);
    }
    frag_acc += f;

    bool split = feed.items > 1u;
    if (split) {
        // a split tile: the item's winners merge into the tile's (one 64-bit
        // maximum of the complemented value per pixel, so that zero means
        // none); the item that finishes last shades the tile
        unsigned long long *gv = d.s.gvis + (size_t)tile * (TS * TS);
        unsigned long long done = 0ull;
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            uint32_t q = tid + (uint32_t)k * NT;
            unsigned long long v = sh.best[q];
            if (v != ~0ull) done |= atomicMax(&gv[q], ~v);
        }
        // the maxima have returned (their results are used here), so they
        // were performed before this item's ticket below
        asm volatile("" : : "v"(done));
        __syncthreads();
        if (tid == 0) {
            uint32_t last = atomicAdd(&d.s.tile_done[tile], 1u) == feed.items - 1u ? 1u : 0u;
            if (last) __threadfence();  // the tile's merged winners are read fresh
            sh.last = last;
        }
        __syncthreads();
        if (sh.last == 0u) return;
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            uint32_t q = tid + (uint32_t)k * NT;
            unsigned long long g = gv[q];
            sh.best[q] = g ? ~g : ~0ull;
            gv[q] = 0ull;  // the buffer is kept zero between uses
        }
        // each thread shades the pixels it read: no barrier needed
    }

    bool write_depth_mode = d.depth_test && d.depth_write &&
                            (d.resolve != RES_PASS_LAST || d.depth_compare == CMP_EQUAL ||
                             d.depth_compare == CMP_ALWAYS);
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        uint32_t q = tid + (uint32_t)k * NT;
        uint32_t px = x0 + q % TS, py = y0 + q / TS;
        if (px >= d.width || py >= d.height) continue;
        size_t pix = (size_t)py * d.width + px;
        unsigned long long v = sh.best[q];
        if (v != ~0ull) {
            float z;
            uint32_t col = winner_color<PROG>(d, oi_key(d.resolve, v), (float)px, (float)py, z, srgb);
            if (d.color) store_color(d, pix, col);
            if (d.depth) {
                if (write_depth_mode) store_depth(d, pix, z);
                else if (d.init_clear && d.clear_depth_on && !d.preclear) store_depth(d, pix, d.clear_depth);
            }
        } else if (d.init_clear && !d.preclear) {
            if (d.color && d.clear_color_on) store_color(d, pix, d.clear_rgba);
            if (d.depth && d.clear_depth_on) store_depth(d, pix, d.clear_depth);
        }
    }
    __syncthreads();  // the next tile resets the per-pixel state
}

// Stores the clear values into one tile.
template <int TS, int NT>
__device__ __forceinline__ void clear_tile(const DrawParams &d, uint32_t tile)
{
    constexpr int PPT = TS * TS / NT;
    uint32_t x0 = (tile % d.tiles_x) * TS, y0 = (tile / d.tiles_x) * TS;
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        uint32_t q = threadIdx.x + (uint32_t)k * NT;
        uint32_t px = x0 + q % TS, py = y0 + q / TS;
        if (px >= d.width || py >= d.height) continue;
        size_t pix = (size_t)py * d.width + px;
        if (d.color && d.clear_color_on) store_color(d, pix, d.clear_rgba);
        if (d.depth && d.clear_depth_on) store_depth(d, pix, d.clear_depth);
    }
}

// The tiles without binned records, kClearChunk per work item: those in
// the footprint union of their draw's big triangles are rastered (`tile_fn`),
// the others get the LoadOp::Clear values (when the draw clears).  The
// chunk's bin counts are read together, one per lane.
template <bool MULTI, int TS, int NT, class F>
__device__ __forceinline__ void untouched_chunk(const DrawParams &d0, const BatchArgs &b, uint32_t chunk,
                                                uint32_t *kinds, F &&tile_fn)
{
    uint32_t g0 = chunk * kClearChunk;
    if (threadIdx.x < kClearChunk) {
        uint32_t gt = g0 + threadIdx.x, kind = 0u, di = 0u;
        if (gt < b.total_tiles && b.tile_count[gt] == 0u) {
            uint32_t tile = gt;
            if constexpr (MULTI) {
                di = b.tile_draw[gt];
                tile = gt - b.tile_base[di];
            }
            const DrawParams &d = draw_at<MULTI>(d0, b, di);
            kind = in_big_union(d, tile) ? 2u : (d.init_clear && b.any_clear ? 1u : 0u);
        }
        kinds[threadIdx.x] = kind | (di << 2);
    }
    __syncthreads();
    for (uint32_t i = 0; i < kClearChunk; ++i) {
        uint32_t kd = __builtin_amdgcn_readfirstlane(kinds[i]);
        if ((kd & 3u) == 0u) continue;
        uint32_t di = kd >> 2, gt = g0 + i;
        const DrawParams &d = draw_at<MULTI>(d0, b, di);
        uint32_t tile = gt - d.s.tile0;
        if ((kd & 3u) == 1u) clear_tile<TS, NT>(d, tile);
        else tile_fn(d, tile);
    }
    __syncthreads();  // kinds is rewritten by the next chunk
}

// Whether the untouched tiles need a pass: some draw clears, or has big
// triangles whose footprint may hold tiles without binned records.
__device__ __forceinline__ bool untouched_pass(const BatchArgs &b)
{
    return b.any_clear != 0u || b.header[H_BIG] != 0u;
}

// The workgroups' share of the active list: round r gives XCD x the
// entries [r*G + x*G/8, r*G + (x+1)*G/8) (G = the grid, a multiple of 8;
// workgroup i runs on XCD i mod 8), so that tiles binned close in time
// (neighbouring triangles of a mesh, the texels they sample) share an L2.
// Round 0 is static: workgroup j of XCD x takes entry x*G/8 + j, read
// together with the active count.  When there are more entries than
// workgroups, the rest of each XCD's share goes out one entry per ticket
// of that XCD's counter (index_of maps the t-th ticket to its entry), so
// workgroups that drew heavy tiles take fewer: one same-address atomic
// per item, eight addresses, spread over the kernel.
struct ActiveSchedule {
    uint32_t first, stride, per, xcd, n, e0;
    __device__ ActiveSchedule(const BatchArgs &b)
    {
        stride = gridDim.x;
        per = stride >> 3;
        xcd = blockIdx.x & 7u;
        first = xcd * per + (blockIdx.x >> 3);
        // the list holds kMaxItems entries per tile (a group has at least
        // one tile); an entry at or beyond the count is not used
        uint32_t cap = b.total_tiles * kMaxItems;
        e0 = b.active_list[min(first, cap - 1u)];
        uint32_t count = b.header[H_ACTIVE];
        asm volatile("" : : "s"(count), "s"(e0));  // both loads in flight together
        n = min(count, cap);
    }
    __device__ bool dynamic() const { return n > stride; }
    // The entry of this XCD's share that ticket t (counted after round 0) names.
    __device__ uint32_t index_of(uint32_t t) const
    {
        uint32_t u = per + t;
        return (u / per) * stride + xcd * per + u % per;
    }
    __device__ uint32_t *queue(const BatchArgs &b) const { return b.header + H_QUEUE + kQueueStride * xcd; }
    __device__ uint32_t load(const BatchArgs &b, uint32_t item) const
    {
        return item < n ? b.active_list[item] : 0u;
    }
};

// Active-list entry e: its draw, tile and the tile's work item.  The
// draw's first global tile comes with its parameters (scalar loads the
// tile needs anyway), not from a table of its own.
template <bool MULTI>
__device__ __forceinline__ void locate_tile(const DrawParams &d0, const BatchArgs &b, uint32_t e, uint32_t &di,
                                            uint32_t &tile, uint32_t &it)
{
    uint32_t gt = e & ((1u << kItemShift) - 1u);
    it = e >> kItemShift;
    di = 0;
    tile = gt;
    if constexpr (MULTI) {
        di = __builtin_amdgcn_readfirstlane(b.tile_draw[gt]);
        tile = gt - draw_at<MULTI>(d0, b, di).s.tile0;
    }
}

// Reduces the workgroup's fragment count into one of kFragSlots counters.
template <bool MULTI, int NT>
__device__ __forceinline__ void count_fragments(const DrawParams &d0, const BatchArgs &b, uint32_t frag_acc,
                                                uint32_t *wave_frags)
{
    for (int off = 32; off > 0; off >>= 1) frag_acc += __shfl_xor(frag_acc, off);
    if (lane_id() == 0) wave_frags[threadIdx.x >> 6] = frag_acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t f = 0;
        for (int w = 0; w < NT / 64; ++w) f += wave_frags[w];
        if (f) atomicAdd(draw_at<MULTI>(d0, b, 0).s.frags + (blockIdx.x % kFragSlots), (unsigned long long)f);
    }
}

// Ordered draws (blending, NotEqual + write): persistent and statically
// scheduled (no shared work counter); each tile's fragments are applied in
// API order.  First the active tiles, then the untouched ones in chunks,
// strided over the whole grid.
template <int PROG, bool MULTI, int TS, int NT>
__global__ __launch_bounds__(NT) void k_raster(DrawParams d0, BatchArgs b)
{
    constexpr int PPT = TS * TS / NT;
    static_assert(NT >= kBatch && PPT >= 1 && TS * TS == PPT * NT, "tile shape");
    __shared__ RasterShared<TS> sh;
    __shared__ uint32_t olist[kOrderedCap];
    __shared__ uint32_t olist_n;
    __shared__ uint32_t ticket;
    __shared__ uint32_t wave_frags[NT / 64];
    __shared__ uint32_t kinds[kClearChunk];

    ActiveSchedule w(b);
    uint32_t frag_acc = 0;
    bool dyn = w.dynamic();
    uint32_t raw = w.e0;
    for (uint32_t item = w.first; item < w.n;) {
        uint32_t e = __builtin_amdgcn_readfirstlane(raw);
        // the ticket of this workgroup's next item, taken while this one is
        // processed (thread 0 takes it, the workgroup reads it after)
        uint32_t tk = 0u;
        if (dyn && threadIdx.x == 0) tk = atomicAdd(w.queue(b), 1u);
        uint32_t di, tile, it;
        locate_tile<MULTI>(d0, b, e, di, tile, it);
        raster_tile<PROG, TS, NT>(draw_at<MULTI>(d0, b, di), tile, sh, olist, olist_n, frag_acc);
        if (!dyn) break;
        if (threadIdx.x == 0) ticket = tk;
        __syncthreads();
        item = w.index_of(ticket);
        raw = w.load(b, item);
    }
    // workgroups late in the schedule have the fewest active tiles: chunks from there
    if (untouched_pass(b))
        for (uint32_t c = gridDim.x - 1u - w.first; c < b.n_chunks; c += gridDim.x)
            untouched_chunk<MULTI, TS, NT>(d0, b, c, kinds, [&](const DrawParams &d, uint32_t tile) {
                raster_tile<PROG, TS, NT>(d, tile, sh, olist, olist_n, frag_acc);
            });
    count_fragments<MULTI, NT>(d0, b, frag_acc, wave_frags);
}

// Order-independent draws: visibility and shading of every active tile
// (raster_tile_oi), then the untouched tiles in chunks, strided over the
// whole grid.
template <int PROG, bool MULTI, int TS, int NT>
__global__ __launch_bounds__(NT, (TilesOccupancy<TS, NT>::value)) void k_tiles(DrawParams d0, BatchArgs b)
{
    constexpr int PPT = TS * TS / NT;
    static_assert(NT >= kBatch && PPT >= 1 && TS * TS == PPT * NT, "tile shape");
    static_assert(VisShared<TS, NT>::NB <= 256 && TS <= 32, "owner entries hold 8-bit triangle and row indices");
    __shared__ VisShared<TS, NT> sh;
    __shared__ uint32_t ticket;
    __shared__ uint32_t wave_frags[NT / 64];
    __shared__ uint32_t kinds[kClearChunk];
    __shared__ float srgb[256];   // the sRGB decode table (one device table for every draw)

    // read only when shading, after the first tile's barriers; every draw
    // carries its device's table (allocated with the device, never null)
    const float *srgb_g = draw_at<MULTI>(d0, b, 0).srgb_lut;
    if constexpr (NT == 256) {
        srgb[threadIdx.x] = srgb_g[threadIdx.x];
    } else {
        for (uint32_t i = threadIdx.x; i < 256u; i += NT) srgb[i] = srgb_g[i];
    }
    ActiveSchedule w(b);
    uint32_t frag_acc = 0;
    bool dyn = w.dynamic();
    uint32_t raw = w.e0;
    for (uint32_t item = w.first; item < w.n;) {
        uint32_t e = __builtin_amdgcn_readfirstlane(raw);
        // the ticket of this workgroup's next item, taken while this one is
        // processed (thread 0 takes it, the workgroup reads it after)
        uint32_t tk = 0u;
        if (dyn && threadIdx.x == 0) tk = atomicAdd(w.queue(b), 1u);
        uint32_t di, tile, it;
        locate_tile<MULTI>(d0, b, e, di, tile, it);
        raster_tile_oi<PROG, TS, NT>(draw_at<MULTI>(d0, b, di), tile, it, sh, srgb, frag_acc);
        if (!dyn) break;
        if (threadIdx.x == 0) ticket = tk;
        __syncthreads();
        item = w.index_of(ticket);
        raw = w.load(b, item);
    }
    // workgroups late in the schedule have the fewest active tiles: chunks from there
    if (untouched_pass(b))
        for (uint32_t c = gridDim.x - 1u - w.first; c < b.n_chunks; c += gridDim.x)
            untouched_chunk<MULTI, TS, NT>(d0, b, c, kinds, [&](const DrawParams &d, uint32_t tile) {
                raster_tile_oi<PROG, TS, NT>(d, tile, 0u, sh, srgb, frag_acc);
            });
    count_fragments<MULTI, NT>(d0, b, frag_acc, wave_frags);
}

// ------------------------------------------------------------------
// Points and lines (raster.rs:184-265, clipper.rs:207-407,
// util/bresenham.rs:78-153)
// ------------------------------------------------------------------

// Cohen-Sutherland (clipper.rs:266-366); false when nothing remains.  The
// second end point is updated from the already updated first one, as the
// reference does.
__device__ __forceinline__ bool clip_line(const float *a, const float *b, float out[2][4], float al[2])
{
    float da[6], db[6];
    uint32_t oa = 0, ob = 0;
#pragma unroll
    for (int pl = 0; pl < 6; ++pl) {
        da[pl] = clip_distance(pl, a);
        db[pl] = clip_distance(pl, b);
        oa |= (da[pl] < 0.0f) ? (1u << pl) : 0u;
        ob |= (db[pl] < 0.0f) ? (1u << pl) : 0u;
    }
    al[0] = 0.0f;
    al[1] = 1.0f;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        out[0][c] = a[c];
        out[1][c] = b[c];
    }
    if ((oa | ob) == 0u) return true;
    if (oa & ob) return false;
    for (int pl = 0; pl < 6; ++pl) {
        uint32_t bit = 1u << pl;
        if (!((oa | ob) & bit)) continue;
        float alpha = da[pl] / (da[pl] - db[pl]);
        if (oa & bit) {
            if (alpha > al[0]) al[0] = alpha;
        } else {
            if (alpha < al[1]) al[1] = alpha;
        }
        if (al[0] > al[1]) return false;
    }
    if (oa) {
        float s = 1.0f - al[0];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            float x = out[0][c] * s;
            float y = out[1][c] * al[0];
            out[0][c] = x + y;
        }
    }
    if (ob) {
        float s = 1.0f - al[1];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            float x = out[0][c] * s;
            float y = out[1][c] * al[1];
            out[1][c] = x + y;
        }
    }
    return true;
}

__device__ __forceinline__ bool in_scissor(const DrawParams &d, uint32_t x, uint32_t y)
{
    return x >= d.scissor[0] && (uint64_t)x < (uint64_t)d.scissor[0] + d.scissor[2] && y >= d.scissor[1] &&
           (uint64_t)y < (uint64_t)d.scissor[1] + d.scissor[3];
}

// One fragment of a point or line: pixel, step (key bits), depth, the lerp
// coefficient of the second vertex (lines) and the fragment position.
struct LpFrag {
    uint32_t x, y, k;
    float z, L;
    float pos[4];
};

// End points of a line (or the point) after clipping and to_raster.
struct LpSetup {
    uint32_t fb[2][2];
    float fr[2][4];    // viewport x, y, NDC z, 1/w of the (clipped) end points
    float al[2];
    bool visible;
};

// Clip + to_raster of primitive v[0..N).  False raises an error bit (w == 0).
template <int N>
__device__ __forceinline__ bool lp_setup(const DrawParams &d, const VsOut *v, LpSetup &s)
{
    float cp[2][4];
    s.visible = true;
    if (N == 1) {
#pragma unroll
        for (int c = 0; c < 4; ++c) cp[0][c] = v[0].clip[c];
        s.al[0] = 0.0f;
        s.al[1] = 1.0f;
    } else {
        s.visible = clip_line(v[0].clip, v[1].clip, cp, s.al);
        if (!s.visible) return true;
    }
#pragma unroll
    for (int k = 0; k < N; ++k) {
        if (cp[k][3] == 0.0f) {
            atomicOr(d.s.error, (uint32_t)ERR_W_ZERO);
            return false;
        }
        to_raster(d.viewport, cp[k], s.fb[k][0], s.fb[k][1], s.fr[k][0], s.fr[k][1], s.fr[k][2]);
        s.fr[k][3] = 1.0f / cp[k][3];
        if ((s.fb[k][0] | s.fb[k][1]) >= (1u << 30)) {
            atomicOr(d.s.error, (uint32_t)ERR_COORD_RANGE);
            return false;
        }
    }
    return true;
}

// Bresenham iterator (bresenham.rs:78-153): start inclusive, end exclusive;
// t is the running binary32 sum of 1/major.
struct Bres {
    int64_t x, y, dx, dy, x1, diff;
    int oct;
    float t, dt;
    __device__ static void to0(int o, int64_t &x, int64_t &y)
    {
        int64_t a = x, b = y;
        switch (o) {
        case 0: x = a; y = b; break;
        case 1: x = b; y = a; break;
        case 2: x = b; y = -a; break;
        case 3: x = -a; y = b; break;
        case 4: x = -a; y = -b; break;
        case 5: x = -b; y = -a; break;
        case 6: x = -b; y = a; break;
        default: x = a; y = -b; break;
        }
    }
    __device__ static void from0(int o, int64_t &x, int64_t &y)
    {
        int64_t a = x, b = y;
        switch (o) {
        case 0: x = a; y = b; break;
        case 1: x = b; y = a; break;
        case 2: x = -b; y = a; break;
        case 3: x = -a; y = b; break;
        case 4: x = -a; y = -b; break;
        case 5: x = -b; y = -a; break;
        case 6: x = b; y = -a; break;
        default: x = a; y = -b; break;
        }
    }
    __device__ void init(const uint32_t s[2], const uint32_t e[2])
    {
        int64_t sx = s[0], sy = s[1], ex = e[0], ey = e[1];
        int64_t ddx = ex - sx, ddy = ey - sy;
        oct = 0;
        if (ddy < 0) {
            ddx = -ddx;
            ddy = -ddy;
            oct += 4;
        }
        if (ddx < 0) {
            int64_t t0 = ddx;
            ddx = ddy;
            ddy = -t0;
            oct += 2;
        }
        if (ddx < ddy) oct += 1;
        to0(oct, sx, sy);
        to0(oct, ex, ey);
        dx = ex - sx;
        dy = ey - sy;
        dt = 1.0f / (float)dx;
        x = sx;
        y = sy;
        x1 = ex;
        diff = dy - dx;
        t = 0.0f;
    }
    __device__ bool next(uint32_t &px, uint32_t &py, float &tt)
    {
        if (x >= x1) return false;
        int64_t qx = x, qy = y;
        if (diff >= 0) {
            y += 1;
            diff -= dx;
        }
        diff += dy;
        tt = t;
        t = t + dt;
        x += 1;
        from0(oct, qx, qy);
        px = (uint32_t)qx;
        py = (uint32_t)qy;
        return true;
    }
};

// Calls f(frag) for every in-scissor fragment of a set-up point or line, in
// rasterizer order; false (error bit raised) when a line is too long for
// the fragment key.
template <int N, class F>
__device__ __forceinline__ bool lp_walk(const DrawParams &d, const LpSetup &s, F &&f)
{
    if (!s.visible) return true;
    LpFrag fr;
    if (N == 1) {
        if (!in_scissor(d, s.fb[0][0], s.fb[0][1])) return true;
        fr.x = s.fb[0][0];
        fr.y = s.fb[0][1];
        fr.k = 0u;
        fr.z = s.fr[0][2];
        fr.L = 0.0f;
#pragma unroll
        for (int c = 0; c < 4; ++c) fr.pos[c] = s.fr[0][c];
        f(fr);
        return true;
    }
    Bres b;
    b.init(s.fb[0], s.fb[1]);
    uint32_t k = 0;
    float t;
    while (b.next(fr.x, fr.y, t)) {
        uint32_t step = k++;
        if (step >= kLineSteps) {   // also bounds the walk of lines far outside the scissor
            atomicOr(d.s.error, (uint32_t)ERR_COORD_RANGE);
            return false;
        }
        if (!in_scissor(d, fr.x, fr.y)) continue;
        // raster.rs:251-254: L = lerp of the clip alphas; the fragment is L
        // applied to the clipped end points
        float u = 1.0f - t;
        float l0 = s.al[0] * u;
        float l1 = s.al[1] * t;
        float L = l0 + l1;
        float m = 1.0f - L;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            float a = s.fr[0][c] * m;
            float bb = s.fr[1][c] * L;
            fr.pos[c] = a + bb;
        }
        fr.z = fr.pos[2];
        fr.L = L;
        fr.k = step;
        f(fr);
    }
    return true;
}

// Varyings of a point (the vertex's own) or line fragment (lerp by L of the
// unclipped vertices; flat: the first; perspective: the extension).
template <int PROG, int N>
__device__ __forceinline__ void lp_varyings(const DrawParams &d, const VsOut *v, float L, float var[kMaxVar])
{
    if (N == 1) {
#pragma unroll
        for (int c = 0; c < kMaxVar; ++c) var[c] = v[0].var[c];
        return;
    }
    float m = 1.0f - L;
    float q0 = m * (1.0f / v[0].clip[3]);
    float q1 = L * (1.0f / v[1].clip[3]);
    float qs = q0 + q1;
#pragma unroll
    for (int c = 0; c < kMaxVar; ++c) {
        bool flat, persp;
        if constexpr (PROG == PROG_USER) {
            flat = (d.var_flat_mask >> c) & 1u;
            persp = (d.var_persp_mask >> c) & 1u;
        } else {
            flat = d.interp == INTERP_FLAT;
            persp = d.interp == INTERP_PERSPECTIVE;
        }
        if (flat) {
            var[c] = v[0].var[c];
        } else if (persp) {
            float acc = v[0].var[c] * q0;
            acc = acc + v[1].var[c] * q1;
            var[c] = acc / qs;
        } else {
            float a = v[0].var[c] * m;
            float b = v[1].var[c] * L;
            var[c] = a + b;
        }
    }
}

struct VecPos {
    const float *p;
    __device__ void operator()(float o[4]) const
    {
        for (int c = 0; c < 4; ++c) o[c] = p[c];
    }
};

// Fragment shader of a point / line fragment (front_facing is true for
// them, state.rs:560-561).  False: nothing is written.
template <int PROG, int N>
__device__ __forceinline__ bool lp_color(const DrawParams &d, const VsOut *v, const LpFrag &f, uint32_t p,
                                         float color[4], float *frag_depth)
{
    float var[kMaxVar];
    lp_varyings<PROG, N>(d, v, f.L, var);
#ifdef WGC_USER_PROGRAM
    if constexpr (PROG == PROG_USER) {
        uint32_t rc = user_fragment(d, VecPos{f.pos}, true, p % d.nprim, var, color, frag_depth);
        if (rc == 0u) return true;
        if (rc != 3u) atomicOr(d.s.error, (uint32_t)ERR_SHADER);
        return false;
    }
#endif
    (void)p;
    (void)frag_depth;
    run_fs<PROG>(d, var, color, d.srgb_lut);
    return true;
}

__device__ __forceinline__ uint32_t lp_key(uint32_t p, uint32_t k, int n) { return n == 1 ? p : (p << 13) | k; }

__device__ __forceinline__ void lp_count_fragment(const DrawParams &d, uint32_t pix)
{
    if (d.s.frag_px) atomicAdd(&d.s.frag_px[pix], 1u);
}

// Pass 1 of an order-independent point / line draw: every fragment that
// passes the depth test against the depth before the draw proposes
// (depth, key) for its pixel (the same encoding as the triangle path).
template <int PROG, int N>
__global__ __launch_bounds__(256) void k_lp_visibility(DrawParams d, unsigned long long *vis)
{
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long frags = 0;
    if (p < d.total_prims) {
        VsOut v[N];
        LpSetup s;
        if (shade_primitive<PROG, N>(d, p, v) && lp_setup<N>(d, v, s)) {
            lp_walk<N>(d, s, [&](const LpFrag &f) {
                uint32_t pix = f.y * d.width + f.x;
                ++frags;
                lp_count_fragment(d, pix);
                uint32_t key = lp_key(p, f.k, N);
                if (d.resolve == RES_LAST) {
                    atomicMin(&vis[pix], (unsigned long long)(~key));
                } else if (oi_pass(d, f.z, d.depth[pix])) {
                    atomicMin(&vis[pix], oi_value(d.resolve, f.z, key));
                }
            });
        }
    }
    // fragment statistics: one atomic per wave
    unsigned long long tot = frags;
    for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off);
    if (lane_id() == 0 && tot) atomicAdd(&d.s.frags[(blockIdx.x * 4 + threadIdx.x / 64) % kFragSlots], tot);
}

// Pass 2: the winning fragment of each pixel runs the fragment shader and
// writes colour (and depth).
template <int PROG, int N>
__global__ __launch_bounds__(256) void k_lp_shade(DrawParams d, const unsigned long long *vis)
{
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.total_prims) return;
    VsOut v[N];
    LpSetup s;
    if (!shade_primitive<PROG, N>(d, p, v) || !lp_setup<N>(d, v, s)) return;
    lp_walk<N>(d, s, [&](const LpFrag &f) {
        uint32_t pix = f.y * d.width + f.x;
        uint32_t key = lp_key(p, f.k, N);
        unsigned long long want = d.resolve == RES_LAST ? (unsigned long long)(~key) : oi_value(d.resolve, f.z, key);
        if (vis[pix] != want) return;
        float color[4];
        float unused = f.z;
        (void)lp_color<PROG, N>(d, v, f, p, color, &unused);
        if (d.color) reinterpret_cast<uint32_t *>(d.color)[pix] = pack_color(color, d.color_bgra);
        if (d.depth_test && d.depth_write) d.depth[pix] = f.z;
    });
}

// Ordered points / lines (blending, NotEqual + write, or a fragment shader
// that may discard).  k_lp_prep runs every primitive's vertex stage and
// set-up once; then each workgroup of k_lp_ordered owns one kLpTile-square
// screen tile and takes the primitives in API order, 256 at a time: those
// whose pixel bounds meet the tile, in order, have thread 0 walk them and
// list their fragments inside the tile, all threads shade those (the
// fragments of one primitive are distinct pixels), in batches.  A pixel
// belongs to one tile, so its fragments are applied in API order as by one
// walk over every primitive.
constexpr int kLpBatch = 512;
constexpr uint32_t kLpTile = 32;

// A set-up point or line: its vertices, clipped end points and the pixel
// bounds of its walk (x0 > x1: nothing to draw).
template <int N>
struct LpPrim {
    VsOut v[N];
    LpSetup s;
    uint32_t x0, y0, x1, y1;
};
static_assert(sizeof(LpPrim<2>) <= kLpPrimBytes && sizeof(LpPrim<1>) <= kLpPrimBytes, "ordered point / line scratch");

template <int PROG, int N>
__global__ __launch_bounds__(256) void k_lp_prep(DrawParams d, LpPrim<N> *prims)
{
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.total_prims) return;
    LpPrim<N> r;
    bool ok = shade_primitive<PROG, N>(d, p, r.v) && lp_setup<N>(d, r.v, r.s) && r.s.visible;
    r.x0 = 1u;
    r.x1 = 0u;
    r.y0 = r.y1 = 0u;
    if (ok) {
        uint32_t e = N == 1 ? 0u : 1u;
        r.x0 = min(r.s.fb[0][0], r.s.fb[e][0]);
        r.x1 = max(r.s.fb[0][0], r.s.fb[e][0]);
        r.y0 = min(r.s.fb[0][1], r.s.fb[e][1]);
        r.y1 = max(r.s.fb[0][1], r.s.fb[e][1]);
        // a line of more than kLineSteps steps: its walk stops there with an
        // error (lp_walk), wherever its later steps fall
        if (N == 2) {
            uint32_t dx = r.x1 - r.x0, dy = r.y1 - r.y0;
            if (max(dx, dy) > kLineSteps) atomicOr(d.s.error, (uint32_t)ERR_COORD_RANGE);
        }
    }
    prims[p] = r;
}

template <int PROG, int N>
__global__ __launch_bounds__(256) void k_lp_ordered(DrawParams d, const LpPrim<N> *prims)
{
    __shared__ LpPrim<N> cur;
    __shared__ LpFrag fr[kLpBatch];
    __shared__ int nf, done;
    __shared__ Bres b;
    __shared__ uint32_t step, entered;
    __shared__ unsigned long long touch[4];
    unsigned long long frags = 0;
    uint32_t tiles_x = (d.width + kLpTile - 1u) / kLpTile, tiles_y = (d.height + kLpTile - 1u) / kLpTile;
    for (uint32_t tile = blockIdx.x; tile < tiles_x * tiles_y; tile += gridDim.x) {
        // the tile's pixels, inclusive
        uint32_t tx0 = (tile % tiles_x) * kLpTile, ty0 = (tile / tiles_x) * kLpTile;
        uint32_t tx1 = min(tx0 + kLpTile, d.width) - 1u, ty1 = min(ty0 + kLpTile, d.height) - 1u;
        for (uint32_t base = 0; base < d.total_prims; base += 256u) {
            uint32_t p = base + threadIdx.x;
            bool t = false;
            if (p < d.total_prims) {
                const LpPrim<N> &q = prims[p];
                t = q.x0 <= q.x1 && q.x0 <= tx1 && q.x1 >= tx0 && q.y0 <= ty1 && q.y1 >= ty0;
            }
            uint64_t m = __ballot(t);
            if (lane_id() == 0) touch[threadIdx.x >> 6] = m;
            __syncthreads();
#pragma unroll 1
            for (uint32_t w = 0; w < 4u; ++w) {
                for (uint64_t mm = touch[w]; mm; mm &= mm - 1u) {
                    uint32_t pp = base + w * 64u + (uint32_t)(__ffsll((unsigned long long)mm) - 1);
                    if (threadIdx.x == 0) {
                        cur = prims[pp];
                        done = 0;
                        step = 0;
                        entered = 0;
                        if (N == 2) b.init(cur.s.fb[0], cur.s.fb[1]);
                    }
                    __syncthreads();
                    while (!done) {
                        if (threadIdx.x == 0) {
                            nf = 0;
                            if (N == 1) {
                                lp_walk<N>(d, cur.s, [&](const LpFrag &f) {
                                    if (f.x >= tx0 && f.x <= tx1 && f.y >= ty0 && f.y <= ty1) fr[nf++] = f;
                                });
                                done = 1;
                            } else {
                                // the steps inside the tile are consecutive (both
                                // coordinates move monotonically): the walk ends
                                // at the first step past them
                                uint32_t x, y;
                                float t;
                                while (nf < kLpBatch) {
                                    if (!b.next(x, y, t)) {
                                        done = 1;
                                        break;
                                    }
                                    uint32_t k = step++;
                                    if (k >= kLineSteps) {   // the error is raised by k_lp_prep
                                        done = 1;
                                        break;
                                    }
                                    bool in = x >= tx0 && x <= tx1 && y >= ty0 && y <= ty1;
                                    if (!in) {
                                        if (entered) {
                                            done = 1;
                                            break;
                                        }
                                        continue;
                                    }
                                    entered = 1;
                                    if (!in_scissor(d, x, y)) continue;
                                    // raster.rs:251-254, as lp_walk
                                    float u = 1.0f - t;
                                    float l0 = cur.s.al[0] * u;
                                    float l1 = cur.s.al[1] * t;
                                    float L = l0 + l1;
                                    float mv = 1.0f - L;
                                    LpFrag f;
                                    f.x = x;
                                    f.y = y;
                                    f.k = k;
                                    for (int c = 0; c < 4; ++c) {
                                        float a = cur.s.fr[0][c] * mv;
                                        float bb = cur.s.fr[1][c] * L;
                                        f.pos[c] = a + bb;
                                    }
                                    f.z = f.pos[2];
                                    f.L = L;
                                    fr[nf++] = f;
                                }
                            }
                        }
                        __syncthreads();
                        for (int i = threadIdx.x; i < nf; i += blockDim.x) {
                            const LpFrag &f = fr[i];
                            uint32_t pix = f.y * d.width + f.x;
                            ++frags;
                            lp_count_fragment(d, pix);
                            float color[4];
                            float dep = d.depth_test ? d.depth[pix] : 0.0f;
                            bool ok = depth_and_shade(d, f.z, dep, [&](float *frag_depth) {
                                return lp_color<PROG, N>(d, cur.v, f, pp, color, frag_depth);
                            });
                            if (d.depth_test && d.depth_write) d.depth[pix] = dep;
                            if (!ok) continue;
                            if (d.color) {
                                uint32_t *cp = reinterpret_cast<uint32_t *>(d.color) + pix;
                                *cp = d.blend ? blend_over(color, *cp, d.color_bgra) : pack_color(color, d.color_bgra);
                            }
                        }
                        __syncthreads();
                    }
                }
            }
            __syncthreads();   // touch[] is written again by the next chunk
        }
    }
    unsigned long long tot = frags;
    for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off);
    if (lane_id() == 0 && tot) atomicAdd(&d.s.frags[(threadIdx.x / 64) % kFragSlots], tot);
}

}  // namespace wgc
