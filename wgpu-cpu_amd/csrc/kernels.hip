// kernels.hip -- host-side launchers: dispatch of the built-in programs'
// launch groups to their per-program units (kernels_prog.hip), the launch
// groups of WGSL pipelines, and the program-independent kernels (counters,
// clears, copies, the restart-strip pre-pass).
#include <algorithm>

#include "kernels_dev.h"
#include "kernels_prog.h"

namespace wgc {

// ------------------------------------------------------------------
// program-independent kernels
// ------------------------------------------------------------------

__global__ void k_reset(unsigned long long *frags, uint32_t *error)
{
    if (threadIdx.x < kFragSlots) frags[threadIdx.x] = 0ull;
    if (threadIdx.x == 0) *error = 0u;
}

// End of a submission: its counters to host memory (out: the fragment
// count summed over the kFragSlots counters, its high word, the error word,
// 0 -- one 16-byte store) and zero for the next one; the zero region of the
// scratch arena that its launch groups used (tile counts, group headers,
// per-draw counters) cleared, n16 16-byte units.
__global__ __launch_bounds__(256) void k_end_submission(unsigned long long *frags, uint32_t *error, uint4 *out,
                                                        uint4 *zone, uint32_t n16)
{
    static_assert(kFragSlots == 64, "one fragment counter per lane of a wave");
    if (blockIdx.x == 0 && threadIdx.x < 64u) {
        unsigned long long f = frags[threadIdx.x];
        frags[threadIdx.x] = 0ull;
        for (int off = 32; off > 0; off >>= 1) f += __shfl_xor(f, off);
        if (threadIdx.x == 0) {
            uint32_t e = *error;
            *error = 0u;
            *out = make_uint4((uint32_t)f, (uint32_t)(f >> 32), e, 0u);
        }
    }
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += gridDim.x * blockDim.x)
        zone[i] = make_uint4(0u, 0u, 0u, 0u);
}

__global__ void k_clear(uint32_t *color, uint32_t rgba, float *depth, float dv, uint64_t n)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (color) color[i] = rgba;
    if (depth) depth[i] = dv;
}



// ------------------------------------------------------------------
// Restart triangle strips: separator summary per chunk of positions
// ------------------------------------------------------------------

// Last separator position inside each chunk (-1: none); one thread per chunk.
__global__ __launch_bounds__(256) void k_strip_chunks(DrawParams d, int32_t *last)
{
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t n_chunks = (d.count + kStripChunk - 1) / kStripChunk;
    if (c >= n_chunks) return;
    uint32_t sep = d.index_format == 2u ? 0xFFFFu : 0xFFFFFFFFu;
    int32_t found = -1;
    uint32_t end = min(d.count, (c + 1) * kStripChunk);
    for (uint32_t i = c * kStripChunk; i < end; ++i) {
        uint32_t idx;
        if (raw_index(d, d.first + i, idx) && idx == sep) found = (int32_t)i;
    }
    last[c] = found;
}

// Inclusive prefix maximum over the chunks (one thread: a few thousand
// chunks at most per draw).
__global__ void k_strip_prefix(DrawParams d, int32_t *last)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint32_t n_chunks = (d.count + kStripChunk - 1) / kStripChunk;
    int32_t m = -1;
    for (uint32_t c = 0; c < n_chunks; ++c) {
        m = max(m, last[c]);
        last[c] = m;
    }
}

hipError_t launch_strip_prepass(const DrawParams &d, hipStream_t st, KernelTimer *kt)
{
    uint32_t chunks = (d.count + kStripChunk - 1) / kStripChunk;
    WGC_LAUNCH("k_strip_chunks", k_strip_chunks, (chunks + 255) / 256, 256, d, d.s.strip_sep);
    WGC_LAUNCH("k_strip_prefix", k_strip_prefix, 1, 64, d, d.s.strip_sep);
    return hipSuccess;
}

// ------------------------------------------------------------------
// built-in programs: one unit per (program, form)
// ------------------------------------------------------------------

#define WGC_EACH_PROGRAM(X) X(PROG_MVP_COLOR) X(PROG_MVP_TEXTURE) X(PROG_COLORED_TRIANGLE) X(PROG_CLIP_COLOR)

hipError_t launch_draw(const DrawParams &d, const BatchArgs &b, const GroupLaunch &gl, hipStream_t st,
                       KernelTimer *kt)
{
    switch (gl.program) {
#define X(P) case P: return launch_program<P, false>(d, b, gl, st, kt);
    WGC_EACH_PROGRAM(X)
#undef X
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_batch(const DrawParams &d0, const BatchArgs &b, const GroupLaunch &gl, hipStream_t st,
                        KernelTimer *kt)
{
    switch (gl.program) {
#define X(P) case P: return launch_program<P, true>(d0, b, gl, st, kt);
    WGC_EACH_PROGRAM(X)
#undef X
    default: return hipErrorInvalidValue;
    }
}

uint32_t raster_blocks_per_cu(uint32_t program, bool ordered, bool multi, uint32_t cfg)
{
    switch (program) {
#define X(P)                                                                                           \
    case P: return multi ? raster_blocks_per_cu_of<P, true>(ordered, cfg) : raster_blocks_per_cu_of<P, false>(ordered, cfg);
    WGC_EACH_PROGRAM(X)
#undef X
    default: return 1;
    }
}

uint32_t setup_clip_blocks_per_cu(uint32_t program)
{
    switch (program) {
#define X(P) case P: return setup_clip_blocks_per_cu_of<P>();
    WGC_EACH_PROGRAM(X)
#undef X
    default: return 0;
    }
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
        hipError_t e = launch_strip_prepass(d, st, kt);
        if (e != hipSuccess) return e;
    }
    if (gl.total_prims > 0) {
        uint32_t g = (gl.total_prims + 255u) / 256u;
        if (!MULTI && gl.inline_clip) {
            WGC_MLAUNCH("k_setup", k.setup_clip, g, 256, &d, &b);
        } else {
            WGC_MLAUNCH("k_setup", k.setup[m], g, 256, &d, &b);
            uint32_t gc = min(max(g / 2u, 1u), 2048u);
            WGC_MLAUNCH("k_clip", k.clip[m], gc, 256, &d, &b, &clip_list, &clip_cap);
        }
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

static hipError_t launch_lines_user(const UserKernels &k, const DrawParams &d0, const GroupLaunch &gl, hipStream_t st,
                                    KernelTimer *kt)
{
    DrawParams d = d0;
    unsigned long long *vis = gl.vis;
    void *lp = gl.lp;
    const int n = d.topology == TOPO_POINT_LIST ? 0 : 1;
    uint32_t g = (d.total_prims + 255u) / 256u;
    if (gl.ordered) {
        if (!g) return hipSuccess;
        WGC_MLAUNCH("k_lp_prep", k.lp_prep[n], g, 256, &d, &lp);
        uint32_t tiles = ((d.width + kLpTile - 1u) / kLpTile) * ((d.height + kLpTile - 1u) / kLpTile);
        WGC_MLAUNCH("k_lp_ordered", k.lp_ordered[n], std::min(tiles, 65535u), 256, &d, &lp);
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
#define X(P) case P: return launch_lines_of<P>(d, gl, st, kt);
    WGC_EACH_PROGRAM(X)
#undef X
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_reset(unsigned long long *frags, uint32_t *error, hipStream_t st)
{
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, st, frags, error);
    return hipGetLastError();
}

hipError_t launch_end_submission(unsigned long long *frags, uint32_t *error, void *host_out, void *zone,
                                 size_t zone_bytes, hipStream_t st, hipEvent_t done)
{
    uint32_t n16 = (uint32_t)(zone_bytes / 16);
    uint32_t grid = std::max(1u, std::min((n16 + 1023u) / 1024u, 1024u));
    // the completion event rides on the kernel's own dispatch (no marker
    // packet of its own after it)
    hipExtLaunchKernelGGL(k_end_submission, dim3(grid), dim3(256), 0, st, nullptr, done, 0u, frags, error,
                          (uint4 *)host_out, (uint4 *)zone, n16);
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
