// wgc_batch.h -- launch-group arguments shared by the kernels and the
// host-side launchers (device-compilable: no host library headers).
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif
#include "wgc_core.h"

namespace wgc {

struct DrawParams;

// Group-level arguments of every kernel: the table of independent draws
// that share the launches (batched form; empty for a single draw passed by
// value) and the group's tile bookkeeping.
struct BatchArgs {
    const DrawParams *draws;     // device array of n draws
    const uint32_t *prim_base;   // n + 1 prefix of primitives, padded to multiples of 64
    const uint32_t *tile_base;   // n + 1 prefix of tiles
    const uint32_t *wave_draw;   // the draw of every 64 primitives (the padded prefix)
    const uint32_t *tile_draw;   // the draw of every tile
    uint32_t n, total_prims, total_tiles, n_chunks;
    uint32_t *header;            // H_* words
    uint32_t *tile_count;        // all tiles of the group
    uint32_t *active_list;       // tiles with binned records, in the order they were first binned
    uint32_t *counters;          // C_NUM per draw
    uint32_t n_counters;
    uint32_t any_clear;          // the tile kernel stores LoadOp::Clear values into untouched tiles
    uint32_t preclear;           // single draw: setup stores the LoadOp::Clear values into the whole target
    uint32_t pad;
};

constexpr uint32_t kClearChunk = 8;     // tiles per work item of the untouched-tile pass

// Raster tile shapes: tile side x threads per workgroup.
enum RasterCfg : uint32_t {
    RASTER_16x16_256 = 0,
    RASTER_8x8_64 = 1,
    RASTER_16x16_64 = 2,
    RASTER_32x32_256 = 3,
    RASTER_16x16_128 = 4,
    RASTER_32x32_128 = 5,
    RASTER_32x32_512 = 6,
    RASTER_32x32_64 = 7,
    RASTER_NUM_CFG = 8,
};

// Threads per workgroup of a tile shape.
__host__ __device__ inline uint32_t raster_threads(uint32_t cfg)
{
    switch (cfg) {
    case RASTER_8x8_64: case RASTER_16x16_64: case RASTER_32x32_64: return 64u;
    case RASTER_16x16_128: case RASTER_32x32_128: return 128u;
    case RASTER_32x32_512: return 512u;
    default: return 256u;
    }
}

// Triangles staged per visibility batch of the order-independent tile
// kernel: 96 for 16x16 tiles of 256 threads (so that six workgroups fit a
// CU's LDS), 64 for 16x16 tiles of at most 128 threads, else 128 (64 for
// one wave).
__host__ __device__ constexpr uint32_t vis_batch_size(uint32_t ts, uint32_t nt)
{
    return (ts == 16u && nt == 256u) ? 96u : (ts == 16u || nt < 128u) ? 64u : 128u;
}

__host__ __device__ inline uint32_t raster_tile_shift(uint32_t cfg)
{
    return cfg == RASTER_8x8_64 ? 3u
         : (cfg == RASTER_32x32_256 || cfg == RASTER_32x32_128 || cfg == RASTER_32x32_512 || cfg == RASTER_32x32_64)
               ? 5u
               : 4u;
}

}  // namespace wgc
