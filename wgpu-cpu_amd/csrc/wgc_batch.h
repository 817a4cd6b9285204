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
    RASTER_NUM_CFG = 4,
};

__host__ __device__ inline uint32_t raster_tile_shift(uint32_t cfg)
{
    return cfg == RASTER_8x8_64 ? 3u : (cfg == RASTER_32x32_256 ? 5u : 4u);
}

}  // namespace wgc
