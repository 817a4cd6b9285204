// wgc_launch.h -- host-side interface between the runtime and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace wgc {

struct DrawParams;

// Group-level arguments of every kernel: the table of independent draws
// that share the launches (batched form; empty for a single draw passed by
// value) and the group's tile bookkeeping.
struct BatchArgs {
    const DrawParams *draws;     // device array of n draws
    const uint32_t *prim_base;   // n + 1 prefix of primitives, padded to multiples of 64
    const uint32_t *tile_base;   // n + 1 prefix of tiles
    uint32_t n, total_prims, total_tiles, n_chunks;
    uint32_t *header;            // H_* words
    uint32_t *tile_count, *tile_offset;
    uint32_t *tile_active;       // 1 when the raster must visit the tile
    uint32_t *active_list;       // compacted active tiles
    uint32_t *block_sums;        // 2 words per scan block
    uint32_t n_blocks, pad;
};

constexpr uint32_t kScanBlock = 1024;   // tiles per scan block
constexpr uint32_t kClearChunk = 8;     // tiles per clear work item

// Raster tile shapes: tile side x threads per workgroup.
enum RasterCfg : uint32_t {
    RASTER_16x16_256 = 0,
    RASTER_8x8_64 = 1,
    RASTER_16x16_64 = 2,
    RASTER_32x32_256 = 3,
    RASTER_NUM_CFG = 4,
};

inline uint32_t raster_tile_shift(uint32_t cfg)
{
    return cfg == RASTER_8x8_64 ? 3u : (cfg == RASTER_32x32_256 ? 5u : 4u);
}

// Group-wide scratch and launch shape.
struct GroupLaunch {
    uint32_t program, ordered, raster_cfg;
    uint32_t total_prims, total_tiles, n_counters, raster_grid;
    uint32_t *tile_count, *bins, *counters, *header;
    uint4 *recs;
    uint32_t *rec_slots;
    uint32_t rec_prims, rec_clip_cap;   // record regions: per primitive, appended
    uint2 *clip_list;            // (draw, primitive) pairs that cross a clip plane
    uint32_t clip_cap;
    unsigned long long *vis;     // visibility buffer: tile side^2 values per active tile
    uint32_t shade_grid;
};

// Brackets launches with HIP events on the launching stream; the runtime
// resolves the pairs once the stream has passed them.
struct KernelTimer {
    struct Pending {
        std::string name;
        hipEvent_t begin, end;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;

    hipEvent_t get()
    {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        (void)hipEventCreate(&e);
        return e;
    }
    void begin(const char *name, hipStream_t st)
    {
        Pending p{name, get(), nullptr};
        (void)hipEventRecord(p.begin, st);
        pending.push_back(p);
    }
    void end(hipStream_t st)
    {
        pending.back().end = get();
        (void)hipEventRecord(pending.back().end, st);
    }
};

// A single draw passes its parameters by value (b.draws unused, b.n == 1);
// a batch passes the device table in b.draws (d0 unused).
hipError_t launch_draw(const DrawParams &d, const BatchArgs &b, const GroupLaunch &gl, hipStream_t st,
                       KernelTimer *kt);
hipError_t launch_batch(const DrawParams &d0, const BatchArgs &b, const GroupLaunch &gl, hipStream_t st,
                        KernelTimer *kt);
// Workgroups of the raster kernels for (program, ordered, batched, tile
// shape) that fit on one CU at once: the visibility (or ordered raster)
// kernel and the shading kernel.
void raster_blocks_per_cu(uint32_t program, bool ordered, bool multi, uint32_t cfg, uint32_t &raster,
                          uint32_t &shade);
// Zeroes the submission counters (fragment slots and error bits).
hipError_t launch_reset(unsigned long long *frags, uint32_t *error, hipStream_t st);
hipError_t launch_clear(uint32_t *color, uint32_t rgba, float *depth, float dv, uint64_t n,
                        hipStream_t st, KernelTimer *kt);

}  // namespace wgc
