// kernels_prog.hip -- host-side launchers of one built-in program's
// render-pass kernels (the kernels themselves are in kernels_dev.h).
//
// Compiled once per (program, form): WGC_PROG is the program number,
// WGC_MULTI 0 for single draws (and the point/line kernels), 1 for batches.
#include <algorithm>

#include "kernels_dev.h"
#include "kernels_prog.h"

#if !defined(WGC_PROG) || !defined(WGC_MULTI)
#error "compile with -DWGC_PROG=<program> -DWGC_MULTI=<0|1>"
#endif

namespace wgc {

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
    case RASTER_32x32_64: WGC_LAUNCH("k_tiles", (k_tiles<PROG, MULTI, 32, 64>), gl.raster_grid, 64, d, b); break;
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
uint32_t raster_blocks_per_cu_of(bool ordered, uint32_t cfg)
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
    case RASTER_32x32_64: return occupancy_tiles<PROG, MULTI, 32, 64>();
    default: return occupancy_tiles<PROG, MULTI, 16, 256>();
    }
}

template <int PROG, bool MULTI>
hipError_t launch_program(const DrawParams &d, const BatchArgs &b, const GroupLaunch &gl, hipStream_t st,
                          KernelTimer *kt)
{
    if (!MULTI && d.strip_restart && d.indexed && d.topology == TOPO_TRIANGLE_STRIP && d.count > 0) {
        hipError_t e = launch_strip_prepass(d, st, kt);
        if (e != hipSuccess) return e;
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

template hipError_t launch_program<WGC_PROG, WGC_MULTI != 0>(const DrawParams &, const BatchArgs &,
                                                             const GroupLaunch &, hipStream_t, KernelTimer *);
template uint32_t raster_blocks_per_cu_of<WGC_PROG, WGC_MULTI != 0>(bool, uint32_t);

#if !WGC_MULTI

template <int PROG>
uint32_t setup_clip_blocks_per_cu_of()
{
    static uint32_t r = 0;
    if (!r) r = occupancy_of(k_setup<PROG, false, true>, 256);
    return r;
}

template <int PROG, int N>
static hipError_t launch_lines_t(const DrawParams &d, const GroupLaunch &gl, hipStream_t st, KernelTimer *kt)
{
    uint32_t g = (d.total_prims + 255u) / 256u;
    if (gl.ordered) {
        if (!g) return hipSuccess;
        WGC_LAUNCH("k_lp_prep", (k_lp_prep<PROG, N>), g, 256, d, (LpPrim<N> *)gl.lp);
        uint32_t tiles = ((d.width + kLpTile - 1u) / kLpTile) * ((d.height + kLpTile - 1u) / kLpTile);
        WGC_LAUNCH("k_lp_ordered", (k_lp_ordered<PROG, N>), min(tiles, 65535u), 256, d, (const LpPrim<N> *)gl.lp);
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
hipError_t launch_lines_of(const DrawParams &d, const GroupLaunch &gl, hipStream_t st, KernelTimer *kt)
{
    return d.topology == TOPO_POINT_LIST ? launch_lines_t<PROG, 1>(d, gl, st, kt) : launch_lines_t<PROG, 2>(d, gl, st, kt);
}

template uint32_t setup_clip_blocks_per_cu_of<WGC_PROG>();
template hipError_t launch_lines_of<WGC_PROG>(const DrawParams &, const GroupLaunch &, hipStream_t, KernelTimer *);

#endif

}  // namespace wgc
