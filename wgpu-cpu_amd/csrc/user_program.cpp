// user_program.cpp -- WGSL pipelines: stage-interface checks, the generated
// glue between the translator's entry points and the render kernels, and
// hipRTC compilation of the kernels that run shader code.
//
// The reference compiles a module once with Cranelift and calls its entry
// points from the CPU executor (naga-cranelift/src/product.rs:200-300,
// wgpu-cpu/src/render_pass/{vertex,fragment}.rs).  Here the entry points
// become device functions that the setup, clip, raster and shade kernels
// call in place of the built-in programs; everything the kernels do around
// them is the same code the built-in programs run.
#include "user_program.h"

#include <hip/hiprtc.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <set>

#include "wgc_core.h"

namespace wgc {

namespace {

// Device sources the programs are compiled from, embedded at build time.
#include "wgc_embedded.inc"

const char *const kRtcOptions[] = {
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
};

struct Compiled {
    std::shared_ptr<RtcModule> code;
    std::vector<hipFunction_t> fns;
};

std::mutex g_cache_mu;
std::map<std::pair<int, std::string>, Compiled> g_cache;   // (device, source) -> program

// Compiles `src` (with the embedded headers and `extra` headers) and loads
// it on device `hip_device` (< 0: compile only); `names` are name
// expressions of the kernels.
bool rtc_build(int hip_device, const std::string &src, const std::vector<std::pair<std::string, std::string>> &extra,
               const std::vector<std::string> &names, Compiled &out, std::string &err)
{
    std::string key = src;
    for (auto &h : extra) key += "\n//--" + h.first + "\n" + h.second;
    {
        std::lock_guard<std::mutex> g(g_cache_mu);
        auto it = g_cache.find({hip_device, key});
        if (it != g_cache.end()) {
            out = it->second;
            return true;
        }
    }
    std::vector<const char *> hdr, hdr_names;
    for (size_t i = 0; i < kEmbeddedCount; ++i) {
        hdr.push_back(kEmbeddedSources[i]);
        hdr_names.push_back(kEmbeddedNames[i]);
    }
    for (auto &h : extra) {
        hdr.push_back(h.second.c_str());
        hdr_names.push_back(h.first.c_str());
    }
    hiprtcProgram prog = nullptr;
    if (hiprtcCreateProgram(&prog, src.c_str(), "wgc_program.hip", (int)hdr.size(), hdr.data(), hdr_names.data()) !=
        HIPRTC_SUCCESS) {
        err = "hiprtcCreateProgram failed";
        return false;
    }
    for (auto &n : names) hiprtcAddNameExpression(prog, n.c_str());
    hiprtcResult r = hiprtcCompileProgram(prog, (int)(sizeof kRtcOptions / sizeof kRtcOptions[0]), kRtcOptions);
    if (r != HIPRTC_SUCCESS) {
        size_t ls = 0;
        hiprtcGetProgramLogSize(prog, &ls);
        std::string log(ls, '\0');
        if (ls) hiprtcGetProgramLog(prog, &log[0]);
        err = std::string("hipRTC compilation failed: ") + hiprtcGetErrorString(r) + "\n" + log.substr(0, 4000);
        hiprtcDestroyProgram(&prog);
        return false;
    }
    std::vector<std::string> lowered;
    for (auto &n : names) {
        const char *lw = nullptr;
        if (hiprtcGetLoweredName(prog, n.c_str(), &lw) != HIPRTC_SUCCESS || !lw) {
            err = "hipRTC: no kernel " + n;
            hiprtcDestroyProgram(&prog);
            return false;
        }
        lowered.push_back(lw);
    }
    size_t cs = 0;
    hiprtcGetCodeSize(prog, &cs);
    std::vector<char> code(cs);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    if (hip_device < 0) return true;
    auto mod = std::make_shared<RtcModule>();
    hipError_t e = hipSetDevice(hip_device);
    if (e == hipSuccess) e = hipModuleLoadData(&mod->module, code.data());
    if (e != hipSuccess) {
        mod->module = nullptr;
        err = std::string("hipModuleLoadData: ") + hipGetErrorString(e);
        return false;
    }
    Compiled c;
    c.code = mod;
    for (auto &lw : lowered) {
        hipFunction_t f = nullptr;
        e = hipModuleGetFunction(&f, mod->module, lw.c_str());
        if (e != hipSuccess) {
            err = std::string("hipModuleGetFunction: ") + hipGetErrorString(e);
            return false;
        }
        c.fns.push_back(f);
    }
    std::lock_guard<std::mutex> g(g_cache_mu);
    g_cache[{hip_device, key}] = c;
    out = c;
    return true;
}

uint32_t occupancy(hipFunction_t f, int threads)
{
    int n = 0;
    if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&n, f, threads, 0) != hipSuccess || n < 1) n = 1;
    return (uint32_t)n;
}

std::string u(uint32_t v) { return std::to_string(v) + "u"; }

// switch (s) { case k: return map[k]; } as a device function
std::string slot_map(const std::string &name, const std::vector<uint32_t> &map)
{
    std::string s = "__device__ __forceinline__ uint32_t " + name + "(uint32_t s)\n{\n    switch (s) {\n";
    for (size_t k = 0; k < map.size(); ++k) s += "    case " + u((uint32_t)k) + ": return " + u(map[k]) + ";\n";
    return s + "    default: return 0u;\n    }\n}\n";
}

// Resource accessors shared by the environments: `MAP` turns the module's
// resource slots into DrawParams::res slots.
std::string env_resources(const std::string &map)
{
    return "    __device__ const uint8_t *buffer(uint32_t s) const { return d.res[" + map + "(s)].ptr; }\n"
           "    __device__ uint64_t buffer_size(uint32_t s) const { return d.res[" + map + "(s)].size; }\n"
           "    __device__ void sample(uint32_t t, uint32_t s, float x, float y, float *out) const\n"
           "    {\n"
           "        const ResBind &rt = d.res[" + map + "(t)];\n"
           "        const ResBind &rs = d.res[" + map + "(s)];\n"
           "        sample_texture(d, TexView{rt.ptr, rt.w, rt.h, rt.srgb, rs.addr_u, rs.addr_v, rs.bilinear}, x, y, out);\n"
           "    }\n";
}

}  // namespace

RtcModule::~RtcModule()
{
    if (module) (void)hipModuleUnload(module);
}

int find_entry(const wgsl::Module &m, const std::string &name, wgsl::Stage stage, std::string &err)
{
    int found = -1;
    for (size_t i = 0; i < m.entries.size(); ++i) {
        const wgsl::EntryPoint &e = m.entries[i];
        if (!name.empty()) {
            if (e.name == name) return e.stage == stage ? (int)i : (err = "entry point '" + name + "' has another stage", -1);
            continue;
        }
        if (e.stage != stage) continue;
        if (found >= 0) {
            err = "the module has several entry points of this stage; name one";
            return -1;
        }
        found = (int)i;
    }
    if (found < 0) err = name.empty() ? "the module has no entry point of this stage" : "no entry point '" + name + "'";
    return found;
}

bool build_user_program(int hip_device, const wgsl::Module &vm, const std::string &vs_entry,
                        const wgsl::Module *fm, const std::string &fs_entry, bool extended, UserProgram &out,
                        std::string &err)
{
    int vi = find_entry(vm, vs_entry, wgsl::Stage::Vertex, err);
    if (vi < 0) return false;
    const wgsl::EntryPoint &ve = vm.entries[vi];
    const wgsl::EntryPoint *fe = nullptr;
    if (fm) {
        int fi = find_entry(*fm, fs_entry, wgsl::Stage::Fragment, err);
        if (fi < 0) return false;
        fe = &fm->entries[fi];
    }
    out = UserProgram{};
    out.vertex_inputs = ve.inputs;
    // inter-stage variables: the vertex outputs in location order, one
    // varying word per component (VsOut::var)
    std::map<uint32_t, std::pair<uint32_t, const wgsl::IoVar *>> var_at;   // location -> (word offset, output)
    uint32_t words = 0;
    for (const wgsl::IoVar &o : ve.outputs) {
        var_at[o.location] = {words, &o};
        words += o.ncomp;
    }
    if (words > (uint32_t)kMaxVar) {
        err = "the vertex outputs hold " + std::to_string(words) + " scalars; at most " + std::to_string(kMaxVar) +
              " inter-stage scalars are supported";
        return false;
    }
    out.nvar = words;
    if (fe) {
        for (const wgsl::IoVar &in : fe->inputs) {
            auto it = var_at.find(in.location);
            if (it == var_at.end() || it->second.second->ncomp != in.ncomp || it->second.second->scalar != in.scalar) {
                err = "fragment input @location(" + std::to_string(in.location) +
                      ") does not match a vertex output of the same type";
                return false;
            }
            uint32_t bits = ((1u << in.ncomp) - 1u) << it->second.first;
            if (in.interp == wgsl::IO_FLAT) {
                out.flat_mask |= bits;
            } else if (in.interp == wgsl::IO_PERSPECTIVE) {
                if (!extended) {
                    err = "perspective interpolation (the WGSL default) is todo!() in the reference "
                          "(fragment.rs:343-345); use @interpolate(linear) or WGC_SEMANTICS_EXTENDED";
                    return false;
                }
                out.persp_mask |= bits;
            }
        }
        bool color = false;
        for (const wgsl::IoVar &o : fe->outputs) {
            if (o.location != 0 || o.ncomp != 4 || o.scalar != 'f') {
                err = "the fragment shader must write one vec4<f32> at @location(0) (one colour target)";
                return false;
            }
            color = true;
        }
        if (!color) {
            err = "the fragment shader writes no colour";
            return false;
        }
        out.discards = fe->may_discard;
        out.frag_depth_tested = fe->frag_depth && fe->frag_depth_tested;
        out.early_depth = fe->early_depth;
    }
    // pipeline resource slots: the union of both stages' bindings
    auto add_slots = [&](const wgsl::Module &m, const wgsl::EntryPoint &e, std::vector<uint32_t> &map) -> bool {
        map.assign(m.resources.size(), 0u);
        for (uint32_t s : e.resources) {
            const wgsl::Resource &r = m.resources[s];
            size_t k = 0;
            for (; k < out.slots.size(); ++k)
                if (out.slots[k].group == r.group && out.slots[k].binding == r.binding) break;
            if (k == out.slots.size()) {
                if (out.slots.size() >= (size_t)kMaxRes) {
                    err = "more than " + std::to_string(kMaxRes) + " resource bindings";
                    return false;
                }
                UserProgram::Slot sl;
                sl.group = r.group;
                sl.binding = r.binding;
                sl.kind = r.kind;
                out.slots.push_back(sl);
            } else if (out.slots[k].kind != r.kind) {
                err = "the stages bind @group(" + std::to_string(r.group) + ") @binding(" + std::to_string(r.binding) +
                      ") as different kinds of resources";
                return false;
            }
            out.slots[k].min_size = std::max(out.slots[k].min_size, r.min_size);
            map[s] = (uint32_t)k;
        }
        return true;
    };
    std::vector<uint32_t> vmap, fmap;
    if (!add_slots(vm, ve, vmap)) return false;
    if (fe && !add_slots(*fm, *fe, fmap)) return false;
    // generated glue
    bool same = fm == &vm;
    std::string g = "#pragma once\n#include \"wgsl_prelude.h\"\nnamespace wgc {\nnamespace umv {\n" + vm.code + "}\n";
    if (fe && !same) g += "namespace umf {\n" + fm->code + "}\n";
    g += slot_map("vs_slot", vmap);
    if (fe) g += slot_map("fs_slot", fmap);
    std::vector<uint32_t> offs(16, 0u);
    for (auto &kv : var_at)
        if (kv.first < 16) offs[kv.first] = kv.second.first;
    g += slot_map("var_offset", offs);
    g += "struct VsEnv {\n"
         "    const DrawParams &d;\n"
         "    uint32_t inst, vert;\n"
         "    VsOut &o;\n"
         "    bool ok;\n"
         "    __device__ uint32_t vertex_index() const { return vert; }\n"
         "    __device__ uint32_t instance_index() const { return inst; }\n"
         "    __device__ bool front_facing() const { return true; }\n"
         "    __device__ void position(float *p) const { p[0] = p[1] = p[2] = 0.0f; p[3] = 1.0f; }\n"
         // vertex.rs:139-152: the attribute's bytes, unconverted; the load is
         // issued whether or not the attribute lies inside the buffer (outside,
         // it reads the buffer's start and fails), so that no branch on the
         // address keeps the vertices' loads apart
         "    __device__ void in_location(uint32_t l, uint32_t *w, uint32_t n)\n"
         "    {\n"
         "        if (l >= d.n_attr || !d.attr[l].data || d.attr[l].size < 4ull * n) {\n"
         "            ok = false;\n"
         "            for (uint32_t k = 0; k < n; ++k) w[k] = 0u;\n"
         "            return;\n"
         "        }\n"
         "        const AttrBind &at = d.attr[l];\n"
         "        uint64_t start = at.step_instance ? (uint64_t)inst * at.stride : (uint64_t)vert * at.stride;\n"
         "        uint64_t off = start + at.offset;\n"
         "        bool in = off + 4ull * n <= at.size;\n"
         "        ok = ok && in;\n"
         "        const uint32_t *p = reinterpret_cast<const uint32_t *>(at.data + (in ? off : 0ull));\n"
         "        for (uint32_t k = 0; k < n; ++k) w[k] = in ? p[k] : 0u;\n"
         "    }\n"
         "    __device__ void out_position(const float *p) { for (int k = 0; k < 4; ++k) o.clip[k] = p[k]; }\n"
         "    __device__ void out_location(uint32_t l, const uint32_t *w, uint32_t n)\n"
         "    {\n"
         "        uint32_t b = var_offset(l);\n"
         "        for (uint32_t k = 0; k < n; ++k) o.var[b + k] = __uint_as_float(w[k]);\n"
         "    }\n"
         "    __device__ void out_frag_depth(float) {}\n" +
         env_resources("vs_slot") +
         "};\n"
         "__device__ __forceinline__ uint32_t user_vertex(const DrawParams &d, uint32_t instance, uint32_t vertex, VsOut &o)\n"
         "{\n"
         "    for (int k = 0; k < kMaxVar; ++k) o.var[k] = 0.0f;\n"
         "    VsEnv env{d, instance, vertex, o, true};\n"
         "    uint32_t rc = umv::" + ve.symbol + "(env);\n"
         "    if (!env.ok) return ERR_VERTEX;\n"
         "    return rc ? (uint32_t)ERR_SHADER : 0u;\n"
         "}\n";
    if (fe) {
        g += "template <class Pos>\n"
             "struct FsEnv {\n"
             "    const DrawParams &d;\n"
             "    const Pos &pos;\n"
             "    bool ff;\n"
             "    uint32_t prim;\n"
             "    const float *var;\n"
             "    float *color;\n"
             "    float *depth;\n"
             "    __device__ uint32_t vertex_index() const { return 0u; }\n"
             "    __device__ uint32_t instance_index() const { return 0u; }\n"
             "    __device__ uint32_t primitive_index() const { return prim; }\n"
             "    __device__ bool front_facing() const { return ff; }\n"
             "    __device__ void position(float *p) const { pos(p); }\n"
             "    __device__ void in_location(uint32_t l, uint32_t *w, uint32_t n) const\n"
             "    {\n"
             "        uint32_t b = var_offset(l);\n"
             "        for (uint32_t k = 0; k < n; ++k) w[k] = __float_as_uint(var[b + k]);\n"
             "    }\n"
             "    __device__ void out_position(const float *) {}\n"
             "    __device__ void out_location(uint32_t l, const uint32_t *w, uint32_t n)\n"
             "    {\n"
             "        if (l == 0u)\n"
             "            for (uint32_t k = 0; k < n && k < 4u; ++k) color[k] = __uint_as_float(w[k]);\n"
             "    }\n"
             // a frag_depth member after the colour comes after the late
             // test and changes nothing (fragment.rs:457-488)
             + std::string(out.frag_depth_tested ? "    __device__ void out_frag_depth(float z) { *depth = z; }\n"
                                                 : "    __device__ void out_frag_depth(float) {}\n") +
             env_resources("fs_slot") +
             "};\n"
             "template <class Pos>\n"
             "__device__ __forceinline__ uint32_t user_fragment(const DrawParams &d, const Pos &pos, bool front_facing,\n"
             "                                                  uint32_t primitive_index, const float *var, float out[4],\n"
             "                                                  float *frag_depth)\n"
             "{\n"
             "    FsEnv<Pos> env{d, pos, front_facing, primitive_index, var, out, frag_depth};\n"
             "    return " + std::string(same ? "umv::" : "umf::") + fe->symbol + "(env);\n"
             "}\n";
    } else {
        g += "template <class Pos>\n"
             "__device__ __forceinline__ uint32_t user_fragment(const DrawParams &, const Pos &, bool, uint32_t,\n"
             "                                                  const float *, float out[4], float *)\n"
             "{\n    out[0] = out[1] = out[2] = 0.0f;\n    out[3] = 1.0f;\n    return 0u;\n}\n";
    }
    g += "}  // namespace wgc\n";
    // the glue follows the kernels: it uses their staged-triangle type
    std::string src = "#define WGC_USER_PROGRAM 1\n#include \"kernels_dev.h\"\n#include \"wgc_user.h\"\n";
    std::vector<std::string> names = {
        "wgc::k_setup<5, false>",          "wgc::k_setup<5, true>",
        "wgc::k_clip<5, false>",           "wgc::k_clip<5, true>",
        "wgc::k_tiles<5, false, 16, 256>", "wgc::k_tiles<5, true, 16, 256>",
        "wgc::k_raster<5, false, 16, 256>", "wgc::k_raster<5, true, 16, 256>",
        "wgc::k_lp_visibility<5, 1>",      "wgc::k_lp_visibility<5, 2>",
        "wgc::k_lp_shade<5, 1>",           "wgc::k_lp_shade<5, 2>",
        "wgc::k_lp_ordered<5, 1>",         "wgc::k_lp_ordered<5, 2>",
        "wgc::k_setup<5, false, true>",
        "wgc::k_lp_prep<5, 1>",            "wgc::k_lp_prep<5, 2>",
    };
    Compiled c;
    if (!rtc_build(hip_device, src, {{"wgc_user.h", g}}, names, c, err)) return false;
    if (hip_device < 0) return true;
    out.code = c.code;
    UserKernels &k = out.kernels;
    k.module = c.code->module;
    for (int m = 0; m < 2; ++m) {
        k.setup[m] = c.fns[0 + m];
        k.clip[m] = c.fns[2 + m];
        k.tiles[m] = c.fns[4 + m];
        k.raster[m] = c.fns[6 + m];
        k.lp_vis[m] = c.fns[8 + m];
        k.lp_shade[m] = c.fns[10 + m];
        k.lp_ordered[m] = c.fns[12 + m];
        k.occ_tiles[m] = occupancy(k.tiles[m], 256);
        k.occ_raster[m] = occupancy(k.raster[m], 256);
    }
    k.setup_clip = c.fns[14];
    k.lp_prep[0] = c.fns[15];
    k.lp_prep[1] = c.fns[16];
    k.occ_setup_clip = occupancy(k.setup_clip, 256);
    return true;
}

bool build_entry_run(int hip_device, const wgsl::Module &m, const std::string &entry, EntryRun &out,
                     std::string &err)
{
    int ei = -1;
    for (size_t i = 0; i < m.entries.size(); ++i)
        if (m.entries[i].name == entry || (entry.empty() && m.entries.size() == 1)) ei = (int)i;
    if (ei < 0) {
        err = entry.empty() ? "the module has several entry points; name one" : "no entry point '" + entry + "'";
        return false;
    }
    const wgsl::EntryPoint &e = m.entries[ei];
    std::vector<uint32_t> ident(m.resources.size());
    for (size_t k = 0; k < ident.size(); ++k) ident[k] = (uint32_t)k;
    std::string g = "#pragma once\n#include \"wgsl_prelude.h\"\nnamespace wgc {\nnamespace um {\n" + m.code + "}\n" +
                    slot_map("run_slot", ident) +
                    "struct RunEnv {\n"
                    "    const DrawParams &d;\n"
                    "    uint32_t index;\n"
                    "    uint32_t *o;\n"
                    "    __device__ uint32_t vertex_index() const { return index; }\n"
                    "    __device__ uint32_t instance_index() const { return 0u; }\n"
                    "    __device__ uint32_t primitive_index() const { return 0u; }\n"
                    "    __device__ bool front_facing() const { return true; }\n"
                    "    __device__ void position(float *p) const { p[0] = (float)index; p[1] = 0.0f; p[2] = 0.0f; p[3] = 1.0f; }\n"
                    "    __device__ void in_location(uint32_t, uint32_t *w, uint32_t n) const { for (uint32_t k = 0; k < n; ++k) w[k] = 0u; }\n"
                    "    __device__ void out_position(const float *p) { for (int k = 0; k < 4; ++k) o[1 + k] = __float_as_uint(p[k]); }\n"
                    "    __device__ void out_frag_depth(float z) { o[1] = __float_as_uint(z); }\n"
                    "    __device__ void out_location(uint32_t l, const uint32_t *w, uint32_t n)\n"
                    "    {\n"
                    "        if (l < 16u)\n"
                    "            for (uint32_t k = 0; k < n; ++k) o[5 + 4 * l + k] = w[k];\n"
                    "    }\n" +
                    env_resources("run_slot") +
                    "};\n"
                    "__global__ void k_run(DrawParams d, uint32_t first, uint32_t count, uint32_t *out)\n"
                    "{\n"
                    "    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;\n"
                    "    if (i >= count) return;\n"
                    "    uint32_t *o = out + (size_t)i * " + u(kRunWords) + ";\n"
                    "    for (uint32_t k = 0; k < " + u(kRunWords) + "; ++k) o[k] = 0u;\n"
                    "    RunEnv env{d, first + i, o};\n"
                    "    o[0] = um::" + e.symbol + "(env);\n"
                    "}\n"
                    "}  // namespace wgc\n";
    std::string src = "#include \"wgc_core.h\"\n#include \"wgc_run.h\"\n";
    Compiled c;
    if (!rtc_build(hip_device, src, {{"wgc_run.h", g}}, {"wgc::k_run"}, c, err)) return false;
    out.code = c.code;
    out.fn = c.fns[0];
    out.slots = m.resources;
    out.stage = e.stage;
    return true;
}

std::string reflection_json(const wgsl::Module &m)
{
    auto str = [](const std::string &s) {
        std::string o = "\"";
        for (char c : s) {
            if (c == '"' || c == '\\') o += '\\';
            o += c;
        }
        return o + "\"";
    };
    auto io = [&](const std::vector<wgsl::IoVar> &v) {
        std::string o = "[";
        for (size_t i = 0; i < v.size(); ++i)
            o += std::string(i ? "," : "") + "{\"location\":" + std::to_string(v[i].location) + ",\"ncomp\":" +
                 std::to_string(v[i].ncomp) + ",\"scalar\":\"" + v[i].scalar + "\",\"interp\":" +
                 std::to_string(v[i].interp) + "}";
        return o + "]";
    };
    static const char *kinds[] = {"uniform", "storage", "texture", "sampler"};
    std::string o = "{\"resources\":[";
    for (size_t i = 0; i < m.resources.size(); ++i) {
        const wgsl::Resource &r = m.resources[i];
        o += std::string(i ? "," : "") + "{\"name\":" + str(r.name) + ",\"group\":" + std::to_string(r.group) +
             ",\"binding\":" + std::to_string(r.binding) + ",\"kind\":\"" + kinds[(int)r.kind] +
             "\",\"min_size\":" + std::to_string(r.min_size) + "}";
    }
    o += "],\"entries\":[";
    for (size_t i = 0; i < m.entries.size(); ++i) {
        const wgsl::EntryPoint &e = m.entries[i];
        std::string res = "[";
        for (size_t k = 0; k < e.resources.size(); ++k) res += (k ? "," : "") + std::to_string(e.resources[k]);
        res += "]";
        o += std::string(i ? "," : "") + "{\"name\":" + str(e.name) + ",\"stage\":\"" +
             (e.stage == wgsl::Stage::Vertex ? "vertex" : "fragment") + "\",\"symbol\":" + str(e.symbol) +
             ",\"inputs\":" + io(e.inputs) + ",\"outputs\":" + io(e.outputs) +
             ",\"vertex_index\":" + (e.vertex_index ? "true" : "false") +
             ",\"instance_index\":" + (e.instance_index ? "true" : "false") +
             ",\"front_facing\":" + (e.front_facing ? "true" : "false") +
             ",\"position_in\":" + (e.position_in ? "true" : "false") +
             ",\"position_out\":" + (e.position_out ? "true" : "false") +
             ",\"frag_depth\":" + (e.frag_depth ? "true" : "false") +
             ",\"frag_depth_tested\":" + (e.frag_depth_tested ? "true" : "false") +
             ",\"primitive_index\":" + (e.primitive_index ? "true" : "false") +
             ",\"early_depth_test\":" + std::to_string(e.early_depth) +
             ",\"may_discard\":" + (e.may_discard ? "true" : "false") + ",\"resources\":" + res + "}";
    }
    return o + "]}";
}

}  // namespace wgc
