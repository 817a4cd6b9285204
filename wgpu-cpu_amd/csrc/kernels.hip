// kernels.hip -- host-side launchers of the render-pass kernels for the
// built-in programs (the kernels themselves are in kernels_dev.h).
#include <algorithm>

#include <hip/hip_ext.h>

#include "kernels_dev.h"
#include "wgc_launch.h"

namespace wgc {

// ------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------

// With a kernel timer, the launch records the kernel's start and end in
// its own events (from the dispatch packet, like a profiler's trace).
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

template <int PROG, bool MULTI>
static hipError_t launch_raster(const DrawParams &d, const BatchArgs &b, const GroupLaunch &gl, hipStream_t st,
                                KernelTimer *kt)
{
    if (gl.ordered) {
        WGC_LAUNCH("k_raster", (k_raster<PROG, MULTI, 16, 256>), gl.raster_grid, 256, d, b);
        return hipSuccess;
    }
    switch (gl.raster_cfg) {
    case RASTER_8x8_64: WGC_LAUNCH("k_tiles", (k_tiles<PROG, MULTI, 8, 64>), gl.raster_grid, 64, d, b); break;
    case RASTER_16x16_64: WGC_LAUNCH("k_tiles", (k_tiles<PROG, MULTI, 16, 64>), gl.raster_grid, 64, d, b); break;
    case RASTER_32x32_256: WGC_LAUNCH("k_tiles", (k_tiles<PROG, MULTI, 32, 256>), gl.raster_grid, 256, d, b); break;
    case RASTER_16x16_128: WGC_LAUNCH("k_tiles", (k_tiles<PROG, MULTI, 16, 128>), gl.raster_grid, 128, d, b); break;
    case RASTER_32x32_128: WGC_LAUNCH("k_tiles", (k_tiles<PROG, MULTI, 32, 128>), gl.raster_grid, 128, d, b); break;
    case RASTER_32x32_512: WGC_LAUNCH("k_tiles", (k_tiles<PROG, MULTI, 32, 512>), gl.raster_grid, 512, d, b); break;
    default: WGC_LAUNCH("k_tiles", (k_tiles<PROG, MULTI, 16, 256>), gl.raster_grid, 256, d, b); break;
    }
    return hipSuccess;
}

// Workgroups of a kernel that fit on one CU at once (registers, LDS, waves);
// the persistent grids are sized to what can be resident, so that the static
// schedule never waits for a second round of workgroups.
template <class K>
static uint32_t occupancy_of(K kernel, int threads)
{
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, threads, 0) != hipSuccess || n < 1) n = 1;
    return (uint32_t)n;
}

template <int PROG, bool MULTI, int TS, int NT>
static uint32_t occupancy_tiles()
{
    static uint32_t r = 0;
    if (!r) r = occupancy_of(k_tiles<PROG, MULTI, TS, NT>, NT);
    return r;
}

template <int PROG, bool MULTI>
static uint32_t occupancy_cfg(bool ordered, uint32_t cfg)
{
    if (ordered) {
        static uint32_t r = 0;
        if (!r) r = occupancy_of(k_raster<PROG, MULTI, 16, 256>, 256);
        return r;
    }
    switch (cfg) {
    case RASTER_8x8_64: return occupancy_tiles<PROG, MULTI, 8, 64>();
    case RASTER_16x16_64: return occupancy_tiles<PROG, MULTI, 16, 64>();
    case RASTER_32x32_256: return occupancy_tiles<PROG, MULTI, 32, 256>();
    case RASTER_16x16_128: return occupancy_tiles<PROG, MULTI, 16, 128>();
    case RASTER_32x32_128: return occupancy_tiles<PROG, MULTI, 32, 128>();
    case RASTER_32x32_512: return occupancy_tiles<PROG, MULTI, 32, 512>();
    default: return occupancy_tiles<PROG, MULTI, 16, 256>();
    }
}

uint32_t raster_blocks_per_cu(uint32_t program, bool ordered, bool multi, uint32_t cfg)
{
    switch (program) {
    case PROG_MVP_COLOR:
        return multi ? occupancy_cfg<PROG_MVP_COLOR, true>(ordered, cfg) : occupancy_cfg<PROG_MVP_COLOR, false>(ordered, cfg);
    case PROG_MVP_TEXTURE:
        return multi ? occupancy_cfg<PROG_MVP_TEXTURE, true>(ordered, cfg)
                     : occupancy_cfg<PROG_MVP_TEXTURE, false>(ordered, cfg);
    case PROG_COLORED_TRIANGLE:
        return multi ? occupancy_cfg<PROG_COLORED_TRIANGLE, true>(ordered, cfg)
                     : occupancy_cfg<PROG_COLORED_TRIANGLE, false>(ordered, cfg);
    case PROG_CLIP_COLOR:
        return multi ? occupancy_cfg<PROG_CLIP_COLOR, true>(ordered, cfg)
                     : occupancy_cfg<PROG_CLIP_COLOR, false>(ordered, cfg);
    default: return 1;
    }
}

uint32_t setup_clip_blocks_per_cu(uint32_t program)
{
    static uint32_t r[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (program >= 8) return 0;
    uint32_t &v = r[program];
    if (!v) {
        switch (program) {
        case PROG_MVP_COLOR: v = occupancy_of(k_setup<PROG_MVP_COLOR, false, true>, 256); break;
        case PROG_MVP_TEXTURE: v = occupancy_of(k_setup<PROG_MVP_TEXTURE, false, true>, 256); break;
        case PROG_COLORED_TRIANGLE: v = occupancy_of(k_setup<PROG_COLORED_TRIANGLE, false, true>, 256); break;
        case PROG_CLIP_COLOR: v = occupancy_of(k_setup<PROG_CLIP_COLOR, false, true>, 256); break;
        default: return 0;
        }
    }
    return v;
}

template <int PROG, bool MULTI>
static hipError_t launch_group(const DrawParams &d, const BatchArgs &b, const GroupLaunch &gl, hipStream_t st,
                               KernelTimer *kt)
{
    if (!MULTI && d.strip_restart && d.indexed && d.topology == TOPO_TRIANGLE_STRIP && d.count > 0) {
        uint32_t chunks = (d.count + kStripChunk - 1) / kStripChunk;
        WGC_LAUNCH("k_strip_chunks", k_strip_chunks, (chunks + 255) / 256, 256, d, d.s.strip_sep);
        WGC_LAUNCH("k_strip_prefix", k_strip_prefix, 1, 64, d, d.s.strip_sep);
    }
    if (gl.total_prims > 0) {
        uint32_t g = (gl.total_prims + 255u) / 256u;
        if constexpr (!MULTI) {
            if (gl.inline_clip) {
                WGC_LAUNCH("k_setup", (k_setup<PROG, false, true>), g, 256, d, b);
                return launch_raster<PROG, MULTI>(d, b, gl, st, kt);
            }
        }
        WGC_LAUNCH("k_setup", (k_setup<PROG, MULTI>), g, 256, d, b);
        // kClipPerWave primitives per wave; the count is only known on the device
        uint32_t gc = min(max(g / 2u, 1u), 2048u);
        WGC_LAUNCH("k_clip", (k_clip<PROG, MULTI>), gc, 256, d, b, gl.clip_list, gl.clip_cap);
    }
    return launch_raster<PROG, MULTI>(d, b, gl, st, kt);
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

hipError_t launch_draw(const DrawParams &d, const BatchArgs &b, const GroupLaunch &gl, hipStream_t st,
                       KernelTimer *kt)
{
    return launch_group_prog<false>(d, b, gl, st, kt);
}

hipError_t launch_batch(const DrawParams &d0, const BatchArgs &b, const GroupLaunch &gl, hipStream_t st,
                        KernelTimer *kt)
{
    return launch_group_prog<true>(d0, b, gl, st, kt);
}

// ---- WGSL pipelines: the shader kernels are module functions ----

#define WGC_MLAUNCH(name, fn, grid, block, ...)                                                   \
    do {                                                                                          \
        void *args_[] = {__VA_ARGS__};                                                            \
        hipError_t e_;                                                                            \
        if (kt) {                                                                                 \
            hipEvent_t b_ = nullptr, f_ = nullptr;                                                \
            kt->pair(name, b_, f_);                                                               \
            uint32_t g_ = (uint32_t)(grid), t_ = (uint32_t)(block);                               \
            e_ = hipExtModuleLaunchKernel(fn, g_ * t_, 1, 1, t_, 1, 1, 0, st, args_, nullptr, b_, f_, 0u); \
        } else {                                                                                  \
            e_ = hipModuleLaunchKernel(fn, (grid), 1, 1, (block), 1, 1, 0, st, args_, nullptr);   \
        }                                                                                         \
        if (e_ != hipSuccess) return e_;                                                          \
    } while (0)

template <bool MULTI>
static hipError_t launch_user_t(const UserKernels &k, const DrawParams &d0, const BatchArgs &b0,
                                const GroupLaunch &gl, hipStream_t st, KernelTimer *kt)
{
    DrawParams d = d0;
    BatchArgs b = b0;
    const uint2 *clip_list = gl.clip_list;
    uint32_t clip_cap = gl.clip_cap;
    const int m = MULTI ? 1 : 0;
    if (!MULTI && d.strip_restart && d.indexed && d.topology == TOPO_TRIANGLE_STRIP && d.count > 0) {
        uint32_t chunks = (d.count + kStripChunk - 1) / kStripChunk;
        WGC_LAUNCH("k_strip_chunks", k_strip_chunks, (chunks + 255) / 256, 256, d, d.s.strip_sep);
        WGC_LAUNCH("k_strip_prefix", k_strip_prefix, 1, 64, d, d.s.strip_sep);
    }
    if (gl.total_prims > 0) {
        uint32_t g = (gl.total_prims + 255u) / 256u;
        WGC_MLAUNCH("k_setup", k.setup[m], g, 256, &d, &b);
        uint32_t gc = min(max(g / 2u, 1u), 2048u);
        WGC_MLAUNCH("k_clip", k.clip[m], gc, 256, &d, &b, &clip_list, &clip_cap);
    }
    if (gl.ordered) WGC_MLAUNCH("k_raster", k.raster[m], gl.raster_grid, 256, &d, &b);
    else WGC_MLAUNCH("k_tiles", k.tiles[m], gl.raster_grid, 256, &d, &b);
    return hipSuccess;
}

hipError_t launch_user(const UserKernels &k, const DrawParams &d, const BatchArgs &b, const GroupLaunch &gl,
                       bool multi, hipStream_t st, KernelTimer *kt)
{
    return multi ? launch_user_t<true>(k, d, b, gl, st, kt) : launch_user_t<false>(k, d, b, gl, st, kt);
}

// ---- copies (command encoder copy_* / clear_*) ----

// `rows` rows of `row_words` 32-bit words from src (pitch src_pitch words)
// to dst (pitch dst_pitch words); 16-byte accesses when every address and
// pitch allows them.  Grid-stride over (row, chunk).
template <bool VEC>
__global__ __launch_bounds__(256) void k_copy_rows(const uint32_t *src, uint64_t src_pitch, uint32_t *dst,
                                                   uint64_t dst_pitch, uint32_t row_words, uint32_t rows)
{
    constexpr uint32_t W = VEC ? 4u : 1u;
    uint32_t chunks = row_words / W;
    uint64_t total = (uint64_t)chunks * rows;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t r = (uint32_t)(i / chunks), c = (uint32_t)(i - (uint64_t)r * chunks);
        const uint32_t *s = src + r * src_pitch + (uint64_t)c * W;
        uint32_t *d = dst + r * dst_pitch + (uint64_t)c * W;
        if constexpr (VEC) *reinterpret_cast<uint4 *>(d) = *reinterpret_cast<const uint4 *>(s);
        else *d = *s;
    }
}

__global__ __launch_bounds__(256) void k_fill_words(uint32_t *dst, uint64_t n, uint32_t value)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = value;
}

hipError_t launch_copy_rows(const void *src, uint64_t src_pitch, void *dst, uint64_t dst_pitch, uint64_t row_bytes,
                            uint32_t rows, hipStream_t st, KernelTimer *kt)
{
    if (!rows || !row_bytes) return hipSuccess;
    bool vec = ((uintptr_t)src % 16u) == 0 && ((uintptr_t)dst % 16u) == 0 && src_pitch % 16u == 0 &&
               dst_pitch % 16u == 0 && row_bytes % 16u == 0;
    uint64_t units = (row_bytes / (vec ? 16u : 4u)) * rows;
    uint32_t grid = (uint32_t)std::min<uint64_t>((units + 255) / 256, 65536);
    const uint32_t *s = (const uint32_t *)src;
    uint32_t *d = (uint32_t *)dst;
    if (vec) WGC_LAUNCH("k_copy", k_copy_rows<true>, grid, 256, s, src_pitch / 4, d, dst_pitch / 4,
                        (uint32_t)(row_bytes / 4), rows);
    else WGC_LAUNCH("k_copy", k_copy_rows<false>, grid, 256, s, src_pitch / 4, d, dst_pitch / 4,
                    (uint32_t)(row_bytes / 4), rows);
    return hipSuccess;
}

hipError_t launch_fill(void *dst, uint64_t bytes, uint32_t value, hipStream_t st, KernelTimer *kt)
{
    uint64_t n = bytes / 4;
    if (!n) return hipSuccess;
    uint32_t grid = (uint32_t)std::min<uint64_t>((n + 255) / 256, 65536);
    WGC_LAUNCH("k_fill", k_fill_words, grid, 256, (uint32_t *)dst, n, value);
    return hipSuccess;
}

// ---- points and lines ----

template <int PROG, int N>
static hipError_t launch_lines_t(const DrawParams &d, const GroupLaunch &gl, hipStream_t st, KernelTimer *kt)
{
    uint32_t g = (d.total_prims + 255u) / 256u;
    if (gl.ordered) {
        WGC_LAUNCH("k_lp_ordered", (k_lp_ordered<PROG, N>), 1, 256, d);
        return hipSuccess;
    }
    if (!g) return hipSuccess;
    hipError_t e = launch_fill(gl.vis, 8ull * d.width * d.height, 0xFFFFFFFFu, st, kt);
    if (e != hipSuccess) return e;
    WGC_LAUNCH("k_lp_visibility", (k_lp_visibility<PROG, N>), g, 256, d, gl.vis);
    WGC_LAUNCH("k_lp_shade", (k_lp_shade<PROG, N>), g, 256, d, (const unsigned long long *)gl.vis);
    return hipSuccess;
}

template <int PROG>
static hipError_t launch_lines_p(const DrawParams &d, const GroupLaunch &gl, hipStream_t st, KernelTimer *kt)
{
    return d.topology == TOPO_POINT_LIST ? launch_lines_t<PROG, 1>(d, gl, st, kt) : launch_lines_t<PROG, 2>(d, gl, st, kt);
}

static hipError_t launch_lines_user(const UserKernels &k, const DrawParams &d0, const GroupLaunch &gl, hipStream_t st,
                                    KernelTimer *kt)
{
    DrawParams d = d0;
    unsigned long long *vis = gl.vis;
    const int n = d.topology == TOPO_POINT_LIST ? 0 : 1;
    uint32_t g = (d.total_prims + 255u) / 256u;
    if (gl.ordered) {
        WGC_MLAUNCH("k_lp_ordered", k.lp_ordered[n], 1, 256, &d);
        return hipSuccess;
    }
    if (!g) return hipSuccess;
    hipError_t e = launch_fill(gl.vis, 8ull * d.width * d.height, 0xFFFFFFFFu, st, kt);
    if (e != hipSuccess) return e;
    WGC_MLAUNCH("k_lp_visibility", k.lp_vis[n], g, 256, &d, &vis);
    WGC_MLAUNCH("k_lp_shade", k.lp_shade[n], g, 256, &d, &vis);
    return hipSuccess;
}

hipError_t launch_lines(const DrawParams &d, const GroupLaunch &gl, hipStream_t st, KernelTimer *kt)
{
    if (gl.user) return launch_lines_user(*gl.user, d, gl, st, kt);
    switch (d.program) {
    case PROG_MVP_COLOR: return launch_lines_p<PROG_MVP_COLOR>(d, gl, st, kt);
    case PROG_MVP_TEXTURE: return launch_lines_p<PROG_MVP_TEXTURE>(d, gl, st, kt);
    case PROG_COLORED_TRIANGLE: return launch_lines_p<PROG_COLORED_TRIANGLE>(d, gl, st, kt);
    case PROG_CLIP_COLOR: return launch_lines_p<PROG_CLIP_COLOR>(d, gl, st, kt);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_reset(unsigned long long *frags, uint32_t *error, hipStream_t st)
{
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, st, frags, error);
    return hipGetLastError();
}

hipError_t launch_end_submission(unsigned long long *frags, uint32_t *error, void *host_out, void *zone,
                                 size_t zone_bytes, hipStream_t st)
{
    uint32_t n16 = (uint32_t)(zone_bytes / 16);
    uint32_t grid = std::max(1u, std::min((n16 + 1023u) / 1024u, 1024u));
    hipLaunchKernelGGL(k_end_submission, dim3(grid), dim3(256), 0, st, frags, error, (uint4 *)host_out,
                       (uint4 *)zone, n16);
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
