// wgc_launch.h -- host-side interface between the runtime and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace wgc {

struct DrawParams;

// A table of independent draws that share every launch (batched form).
struct BatchArgs {
    const DrawParams *draws;     // device array of n draws
    const uint32_t *prim_base;   // n + 1 prefix of primitives, padded to multiples of 64
    const uint32_t *tile_base;   // n + 1 prefix of tiles
    uint32_t n, total_prims, total_tiles, pad;
};

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
    uint32_t total_prims, total_tiles, n_counters;
    uint32_t *tile_count, *tile_offset, *tile_fill, *bins, *counters, *rec_count;
    uint4 *recs;
    uint32_t rec_cap;
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

hipError_t launch_draw(const DrawParams &d, const GroupLaunch &gl, hipStream_t st, KernelTimer *kt);
hipError_t launch_batch(const BatchArgs &b, const GroupLaunch &gl, hipStream_t st, KernelTimer *kt);
hipError_t launch_reset(unsigned long long *frags, hipStream_t st);
hipError_t launch_clear(uint32_t *color, uint32_t rgba, float *depth, float dv, uint64_t n,
                        hipStream_t st, KernelTimer *kt);

}  // namespace wgc
