// wgsl.h -- WGSL front end: parses and type-checks a WGSL module of the
// subset the reference's shader compiler implements, and emits C++ device
// functions (against wgsl_prelude.h) that the runtime compiles with hipRTC
// into the render kernels at pipeline creation.
//
// The reference runs WGSL through naga's front end and its own Cranelift
// JIT (naga-cranelift/src/lib.rs:83-113, compiler.rs); this translator
// replaces both for the constructs that JIT implements
// (naga-cranelift/src/expression/mod.rs:114-149 minus its todo!() list).
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

namespace wgc {
namespace wgsl {

enum class Stage : uint32_t { Vertex = 0, Fragment = 1 };
enum class ResKind : uint32_t { Uniform = 0, Storage = 1, Texture = 2, Sampler = 3 };
enum Interp : uint32_t { IO_LINEAR = 0, IO_PERSPECTIVE = 1, IO_FLAT = 2 };

// A resource variable (@group @binding) some entry point uses; its index in
// Module::resources is the slot the emitted code names.
struct Resource {
    uint32_t group = 0, binding = 0;
    ResKind kind = ResKind::Uniform;
    uint64_t min_size = 0;      // bytes a bound buffer must hold
    std::string name;
};

// A user-defined input or output (@location): `ncomp` 32-bit words of
// scalar kind 'f', 'i', 'u' (bool is not allowed at the stage interface).
struct IoVar {
    uint32_t location = 0, ncomp = 0;
    char scalar = 'f';
    uint32_t interp = IO_PERSPECTIVE;
};

struct EntryPoint {
    std::string name;
    Stage stage = Stage::Vertex;
    std::string symbol;                 // emitted wrapper: template <class Env> uint32_t symbol(Env &)
    std::vector<IoVar> inputs, outputs; // sorted by location
    bool vertex_index = false, instance_index = false;
    bool front_facing = false, position_in = false;
    bool position_out = false;          // vertex: writes @builtin(position)
    bool frag_depth = false;            // fragment: writes @builtin(frag_depth)
    // fragment: the frag_depth output comes before the first @location
    // output in the result's member order, so the late depth test (run when
    // the first location is written, fragment.rs:465-478) uses it
    bool frag_depth_tested = false;
    bool primitive_index = false;       // fragment: reads @builtin(primitive_index)
    // fragment: @early_depth_test -- 0 none (late test only), 1 force (early
    // test only), 2 allow with a conservative depth (early and late test),
    // fragment.rs:166-194
    uint32_t early_depth = 0;
    bool may_discard = false;           // a discard is reachable
    std::vector<uint32_t> resources;    // slots reachable from the entry point
};

struct Module {
    std::string code;                   // C++ source; the caller wraps it in a namespace
    std::vector<Resource> resources;
    std::vector<EntryPoint> entries;
};

// Translates `src`.  Returns false with a message (line:column: text) in
// `err` when the module is not valid WGSL or uses a construct outside the
// subset.
bool translate(const std::string &src, Module &out, std::string &err);

}  // namespace wgsl
}  // namespace wgc
