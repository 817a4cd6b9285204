// kernels_prog.h -- launchers of one built-in program's render-pass kernels.
//
// kernels_prog.hip defines these templates and is compiled once per
// (program, form) pair with WGC_PROG and WGC_MULTI set, each object holding
// the explicit instantiations for its pair; kernels.hip dispatches to them
// by the program number.  The pairs compile in parallel.
#pragma once

#include "wgc_launch.h"

namespace wgc {

// The launch group of a triangle draw (MULTI false) or batch (MULTI true):
// setup, clip, raster.
template <int PROG, bool MULTI>
hipError_t launch_program(const DrawParams &d, const BatchArgs &b, const GroupLaunch &gl, hipStream_t st,
                          KernelTimer *kt);
// Workgroups of the program's raster kernel (ordered or not, tile shape
// cfg) resident on one CU at once.
template <int PROG, bool MULTI>
uint32_t raster_blocks_per_cu_of(bool ordered, uint32_t cfg);
// Workgroups per CU of the single-draw setup kernel that also clips.
template <int PROG>
uint32_t setup_clip_blocks_per_cu_of();
// A point or line draw of the program.
template <int PROG>
hipError_t launch_lines_of(const DrawParams &d, const GroupLaunch &gl, hipStream_t st, KernelTimer *kt);

// The restart-strip pre-pass of an indexed triangle strip with primitive
// restart (separator summary per kStripChunk positions), any program.
hipError_t launch_strip_prepass(const DrawParams &d, hipStream_t st, KernelTimer *kt);

}  // namespace wgc
