// user_program.h -- render pipelines and entry-point runs whose shaders come
// from WGSL modules: the glue between the translator's emitted functions
// and the render kernels, and their hipRTC compilation.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "wgc_launch.h"
#include "wgsl.h"

namespace wgc {

// A compiled code object (unloaded when the last holder lets go).
struct RtcModule {
    hipModule_t module = nullptr;
    ~RtcModule();
};

// The shader program of one pipeline.
struct UserProgram {
    std::shared_ptr<RtcModule> code;
    UserKernels kernels;
    struct Slot {
        uint32_t group = 0, binding = 0;
        wgsl::ResKind kind = wgsl::ResKind::Uniform;
        uint64_t min_size = 0;
    };
    std::vector<Slot> slots;              // DrawParams::res, in order
    std::vector<wgsl::IoVar> vertex_inputs;
    uint32_t nvar = 0, flat_mask = 0, persp_mask = 0;
    bool discards = false;
    // the fragment shader's frag_depth output is the depth of the late test
    // (a member before its colour), and its @early_depth_test mode
    bool frag_depth_tested = false;
    uint32_t early_depth = 0;
};

// Pairs a vertex entry point with an optional fragment entry point (entry
// names may be empty: the module's only entry point of that stage), checks
// the stage interface and compiles the kernels (hip_device < 0: compile
// only, without loading).  `extended` allows the perspective interpolation
// the reference leaves as todo!().
bool build_user_program(int hip_device, const wgsl::Module &vm, const std::string &vs_entry,
                        const wgsl::Module *fm, const std::string &fs_entry, bool extended, UserProgram &out,
                        std::string &err);

// Per invocation of an entry-point run: abort code, the position (vertex)
// or frag depth (fragment), then 16 locations x 4 words of outputs.
constexpr uint32_t kRunWords = 1 + 4 + 16 * 4;

struct EntryRun {
    std::shared_ptr<RtcModule> code;
    hipFunction_t fn = nullptr;
    std::vector<wgsl::Resource> slots;    // the module's resource slots
    wgsl::Stage stage = wgsl::Stage::Vertex;
};

bool build_entry_run(int hip_device, const wgsl::Module &m, const std::string &entry, EntryRun &out,
                     std::string &err);

// Finds an entry point by name (empty: the only one of `stage`); -1 if none.
int find_entry(const wgsl::Module &m, const std::string &name, wgsl::Stage stage, std::string &err);

// The module's entry points and resources as JSON.
std::string reflection_json(const wgsl::Module &m);

}  // namespace wgc
