// wgc_launch.h -- host-side interface between the runtime and the kernels.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "wgc_batch.h"

namespace wgc {

struct UserKernels;

// Group-wide scratch and launch shape.
struct GroupLaunch {
    uint32_t program, ordered, raster_cfg;
    const UserKernels *user;     // PROG_USER: the pipeline's kernels
    uint32_t total_prims, total_tiles, n_counters, raster_grid;
    uint32_t *tile_count, *counters, *header;
    uint2 *clip_list;            // (draw, primitive) pairs that cross a clip plane
    uint32_t clip_cap;
    uint32_t inline_clip;        // single draw whose setup waves all fit at once: setup clips
    unsigned long long *vis;     // points and lines: one visibility value per pixel
    void *lp;                    // ordered points and lines: kLpPrimBytes per primitive (LpPrim)
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
    // Two events for one launch that records them itself (the start and end
    // of the kernel's execution, from its dispatch packet).
    void pair(const char *name, hipEvent_t &b, hipEvent_t &e)
    {
        Pending p{name, get(), get()};
        pending.push_back(p);
        b = p.begin;
        e = p.end;
    }
};

// Launches a kernel on stream `st`; with a kernel timer `kt`, the launch
// records the kernel's start and end in its own events (from the dispatch
// packet, like a profiler's trace).  Returns the launch error from the
// enclosing function.
#define WGC_LAUNCH(name, kern, grid, block, ...)                                                      \
    do {                                                                                              \
        if (kt) {                                                                                     \
            hipEvent_t b_ = nullptr, f_ = nullptr;                                                    \
            kt->pair(name, b_, f_);                                                                   \
            hipExtLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, st, b_, f_, 0u, __VA_ARGS__);     \
        } else {                                                                                      \
            hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, st, __VA_ARGS__);                    \
        }                                                                                             \
        hipError_t e_ = hipGetLastError();                                                            \
        if (e_ != hipSuccess) return e_;                                                              \
    } while (0)

// Kernels of a pipeline whose shaders come from WGSL: the instantiations
// for PROG_USER of the kernels that run shader code, compiled with hipRTC
// (index: 0 single draw, 1 batched), and their resident workgroups per CU.
struct UserKernels {
    hipModule_t module = nullptr;
    hipFunction_t setup[2]{}, clip[2]{}, tiles[2]{}, raster[2]{};
    hipFunction_t lp_vis[2]{}, lp_shade[2]{}, lp_ordered[2]{}, lp_prep[2]{};   // index: points 0, lines 1
    hipFunction_t setup_clip{};   // single draws: setup that clips in place
    uint32_t occ_tiles[2]{}, occ_raster[2]{}, occ_setup_clip{};
};

// Launch group of a WGSL pipeline (16x16 tiles): the shader kernels from
// `k`, the rest from the library.
hipError_t launch_user(const UserKernels &k, const DrawParams &d, const BatchArgs &b, const GroupLaunch &gl,
                       bool multi, hipStream_t st, KernelTimer *kt);
// A point or line draw: visibility + shading passes over its primitives
// (or one ordered workgroup); gl.vis holds one 64-bit value per pixel.
hipError_t launch_lines(const DrawParams &d, const GroupLaunch &gl, hipStream_t st, KernelTimer *kt);

// A single draw passes its parameters by value (b.draws unused, b.n == 1);
// a batch passes the device table in b.draws (d0 unused).
hipError_t launch_draw(const DrawParams &d, const BatchArgs &b, const GroupLaunch &gl, hipStream_t st,
                       KernelTimer *kt);
hipError_t launch_batch(const DrawParams &d0, const BatchArgs &b, const GroupLaunch &gl, hipStream_t st,
                        KernelTimer *kt);
// Workgroups of the raster kernel for (program, ordered, batched, tile
// shape) that fit on one CU at once.
uint32_t raster_blocks_per_cu(uint32_t program, bool ordered, bool multi, uint32_t cfg);
// Workgroups per CU of the setup kernel that also clips (single draws of a
// built-in program); 0 for other programs.
uint32_t setup_clip_blocks_per_cu(uint32_t program);
// Zeroes the submission counters (fragment slots and error bits).
hipError_t launch_reset(unsigned long long *frags, uint32_t *error, hipStream_t st);
// Copies the submission counters to host memory (16 bytes, aligned: the
// fragment count summed over the kFragSlots counters as two 32-bit words,
// then the error word and 0) and zeroes them, and clears the first
// zone_bytes (a multiple of 16) of the scratch arena's zero region; `done`
// is recorded when the kernel completes (the submission's completion).
hipError_t launch_end_submission(unsigned long long *frags, uint32_t *error, void *host_out, void *zone,
                                 size_t zone_bytes, hipStream_t st, hipEvent_t done);
// Row copy (4-byte multiples; 16-byte accesses where aligned) and word fill,
// ordered on the stream like the render launches.
hipError_t launch_copy_rows(const void *src, uint64_t src_pitch, void *dst, uint64_t dst_pitch, uint64_t row_bytes,
                            uint32_t rows, hipStream_t st, KernelTimer *kt);
hipError_t launch_fill(void *dst, uint64_t bytes, uint32_t value, hipStream_t st, KernelTimer *kt);
hipError_t launch_clear(uint32_t *color, uint32_t rgba, float *depth, float dv, uint64_t n,
                        hipStream_t st, KernelTimer *kt);

}  // namespace wgc
