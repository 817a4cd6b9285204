// runtime.cpp -- libwgc C ABI: devices, resources, command recording and the
// queue that turns recorded render passes into kernel launches.
//
// Mirrors the reference's backend objects (wgpu-cpu/src/{device,buffer,
// texture,sampler,bind_group,shader,pipeline,command}.rs and
// render_pass/mod.rs): handles are reference counted like its Arc clones,
// and one HIP stream per device takes the place of its engine thread
// (engine.rs:26-46).
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/wgc.h"
#include "user_program.h"
#include "wgc_core.h"
#include "wgc_launch.h"
#include "wgsl.h"

using namespace wgc;

// ------------------------------------------------------------------
// errors
// ------------------------------------------------------------------

static thread_local std::string g_last_error;

static int32_t set_error(int32_t code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

static int32_t hip_fail(hipError_t e, const char *what)
{
    return set_error(WGC_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(expr, what)                              \
    do {                                                 \
        hipError_t _e = (expr);                          \
        if (_e != hipSuccess) return hip_fail(_e, what); \
    } while (0)

// Every transfer and fill goes through the device's own (non-blocking)
// stream and is waited for: null-stream memsets and pageable copies are not
// ordered with that stream and may complete after later work on it.
static hipError_t stream_copy(hipStream_t st, void *dst, const void *src, size_t n, hipMemcpyKind kind)
{
    if (!n) return hipSuccess;
    hipError_t e = hipMemcpyAsync(dst, src, n, kind, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e;
}

static hipError_t stream_zero(hipStream_t st, void *dst, size_t n)
{
    if (!n) return hipSuccess;
    hipError_t e = hipMemsetAsync(dst, 0, n, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e;
}

// Zero fill ordered on the device stream without a host wait: later work on
// the stream (submissions, reads) sees it, and nothing else touches device
// memory.  Never called while the stream is being captured.
static hipError_t stream_zero_async(hipStream_t st, void *dst, size_t n)
{
    if (!n) return hipSuccess;
    return hipMemsetAsync(dst, 0, n, st);
}

// First-fit allocator over [0, capacity) in 256-byte granules with
// coalescing of neighbouring free ranges.
class RangeAllocator {
public:
    static constexpr uint64_t kAlign = 256;
    void init(uint64_t capacity)
    {
        free_.clear();
        cap_ = capacity / kAlign * kAlign;
        if (cap_) free_[0] = cap_;
        used_ = 0;
    }
    bool alloc(uint64_t size, uint64_t &off)
    {
        size = std::max<uint64_t>(kAlign, (size + kAlign - 1) / kAlign * kAlign);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < size) continue;
            off = it->first;
            uint64_t rest = it->second - size;
            free_.erase(it);
            if (rest) free_[off + size] = rest;
            used_ += size;
            return true;
        }
        return false;
    }
    void release(uint64_t off, uint64_t size)
    {
        size = std::max<uint64_t>(kAlign, (size + kAlign - 1) / kAlign * kAlign);
        used_ -= size;
        auto next = free_.lower_bound(off);
        if (next != free_.end() && off + size == next->first) {
            size += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += size;
                return;
            }
        }
        free_[off] = size;
    }
    uint64_t used() const { return used_; }
    uint64_t capacity() const { return cap_; }
    uint64_t largest_free() const
    {
        uint64_t m = 0;
        for (auto &kv : free_) m = std::max(m, kv.second);
        return m;
    }

private:
    std::map<uint64_t, uint64_t> free_;
    uint64_t cap_ = 0, used_ = 0;
};

// Blocks of device or pinned host memory carved by RangeAllocator; grows by
// whole blocks, never frees a block before the device goes away (freeing
// would synchronise the device).
struct BlockPool {
    struct Block {
        uint8_t *base;
        RangeAllocator ra;
    };
    bool pinned = false;
    uint64_t block_size = 0;
    std::vector<Block> blocks;

    hipError_t alloc(uint64_t size, uint8_t *&ptr)
    {
        for (auto &b : blocks) {
            uint64_t off;
            if (b.ra.alloc(size, off)) {
                ptr = b.base + off;
                return hipSuccess;
            }
        }
        uint64_t bs = std::max<uint64_t>(block_size, (size + RangeAllocator::kAlign - 1) / RangeAllocator::kAlign *
                                                         RangeAllocator::kAlign);
        void *p = nullptr;
        hipError_t e = pinned ? hipHostMalloc(&p, bs, hipHostMallocDefault) : hipMalloc(&p, bs);
        if (e != hipSuccess) return e;
        blocks.push_back(Block{(uint8_t *)p, RangeAllocator()});
        blocks.back().ra.init(bs);
        uint64_t off = 0;
        blocks.back().ra.alloc(size, off);
        ptr = (uint8_t *)p + off;
        return hipSuccess;
    }
    void release(uint8_t *ptr, uint64_t size)
    {
        for (auto &b : blocks)
            if (ptr >= b.base && ptr < b.base + b.ra.capacity()) {
                b.ra.release((uint64_t)(ptr - b.base), size);
                return;
            }
    }
    void destroy()
    {
        for (auto &b : blocks) {
            if (pinned) (void)hipHostFree(b.base);
            else (void)hipFree(b.base);
        }
        blocks.clear();
    }
};

extern "C" uint32_t wgc_abi_version(void) { return WGC_ABI_VERSION; }
extern "C" const char *wgc_last_error(void) { return g_last_error.c_str(); }

// ------------------------------------------------------------------
// static scratch of the allocation-free submission
// ------------------------------------------------------------------

// wgc_queue_submit_noalloc lays its launch groups' scratch (zero region,
// kept-zero region, bins, records, lists) out in this array instead of the
// device's growable arena.  Every GPU the library runs on holds its own copy
// until the process ends.  In host code `scratch` names a host-side stand-in,
// not a device address: host code only ever uses the address that
// hipGetSymbolAddress returns for the current device (WgcDevice_T::scratch).
static constexpr size_t kScratchBytes = 128ull << 20;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

static std::mutex g_scratch_mu;
static std::map<int, uint8_t *> g_scratch_address;   // HIP ordinal -> address of `scratch` there

// The address of `scratch` on `hip_device`, which must be current.
static hipError_t scratch_address(int hip_device, uint8_t *&out)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_address.find(hip_device);
    if (it != g_scratch_address.end()) {
        out = it->second;
        return hipSuccess;
    }
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_address[hip_device] = (uint8_t *)p;
    out = (uint8_t *)p;
    return hipSuccess;
}

// ------------------------------------------------------------------
// objects
// ------------------------------------------------------------------

struct Object {
    std::atomic<int> refs{1};
    virtual ~Object() = default;
};

static void retain(Object *o)
{
    if (o) o->refs.fetch_add(1, std::memory_order_relaxed);
}

static void release(Object *o)
{
    if (o && o->refs.fetch_sub(1, std::memory_order_acq_rel) == 1) delete o;
}

struct WgcDevice_T;

// Device memory pool that resources are sub-allocated from (serving: targets
// that come and go at request rate, without hipMalloc / hipFree per request).
struct WgcHeap_T : Object {
    WgcDevice_T *dev = nullptr;
    uint8_t *base = nullptr;
    std::mutex mu;
    RangeAllocator ra;
    ~WgcHeap_T() override;
};

// Mapping of a buffer into host memory (buffer.rs:112-167): the range is
// copied into pinned memory on the device stream after the work submitted
// before map_async, and is mapped once that copy has completed.
struct MapState {
    uint32_t mode = 0;           // WgcMapMode; 0: not mapped
    bool pending = false;        // requested, copy not yet complete
    uint64_t offset = 0, size = 0;
    uint8_t *host = nullptr;     // pinned
    hipEvent_t done = nullptr;
    WgcBufferMapCallback callback = nullptr;
    void *userdata = nullptr;
};

struct WgcBuffer_T : Object {
    WgcDevice_T *dev = nullptr;
    WgcHeap_T *heap = nullptr;   // set when sub-allocated from a heap
    uint8_t *ptr = nullptr;
    uint64_t size = 0;
    MapState map;
    ~WgcBuffer_T() override;
};

struct WgcTexture_T : Object {
    WgcDevice_T *dev = nullptr;
    WgcHeap_T *heap = nullptr;
    uint8_t *ptr = nullptr;
    uint32_t width = 0, height = 0, format = 0;
    uint64_t size = 0;
    ~WgcTexture_T() override;
};

struct WgcSampler_T : Object {
    WgcSamplerDescriptor desc{};
};

struct BindingSlot {
    uint32_t binding = 0, kind = 0;
    WgcBuffer_T *buffer = nullptr;
    uint64_t offset = 0, size = 0;
    WgcTexture_T *texture = nullptr;
    WgcSampler_T *sampler = nullptr;
};

struct WgcBindGroup_T : Object {
    std::vector<BindingSlot> entries;
    ~WgcBindGroup_T() override
    {
        for (auto &e : entries) {
            release(e.buffer);
            release(e.texture);
            release(e.sampler);
        }
    }
    const BindingSlot *find(uint32_t binding) const
    {
        for (auto &e : entries)
            if (e.binding == binding) return &e;
        return nullptr;
    }
};

struct WgcShaderModule_T : Object {
    uint32_t program = 0, interpolation = 0;
    std::shared_ptr<wgsl::Module> wgsl;               // PROG_USER
    std::map<std::string, std::shared_ptr<EntryRun>> runs;
    std::mutex mu;
};

struct WgcRenderPipeline_T : Object {
    WgcRenderPipelineDescriptor desc{};
    uint32_t program = 0, interp = 0, nvar = 0;
    bool has_fragment = false;
    std::shared_ptr<UserProgram> user;                // PROG_USER
    std::vector<uint32_t> vb_stride, vb_step;
    AttrDev attr[kMaxAttr]{};
    uint32_t n_attr = 0;
};

struct BufferBinding {
    WgcBuffer_T *buffer = nullptr;
    uint64_t offset = 0, size = 0;
    uint32_t format = 0;
};

enum class CmdKind { SetPipeline, SetBindGroup, SetVertexBuffer, SetIndexBuffer, SetViewport,
                     SetScissor, SetBlendConstant, SetStencilReference, Draw, DrawIndexed };

struct SubCommand {
    CmdKind kind;
    Object *obj = nullptr;     // pipeline / bind group / buffer (one reference held)
    uint32_t slot = 0, format = 0;
    uint64_t offset = 0, size = 0;
    float viewport[6]{};
    uint32_t rect[4]{};
    uint32_t first = 0, count = 0, first_instance = 0, instance_count = 0;
    int32_t base_vertex = 0;
};

enum CopyKind : uint32_t {
    COPY_NONE = 0, COPY_B2B, COPY_T2B, COPY_B2T, COPY_T2T, CLEAR_BUFFER, CLEAR_TEXTURE,
};

// A render pass, or (copy_kind != COPY_NONE) a copy or clear recorded on the
// encoder between passes.
struct PassCommand {
    uint32_t copy_kind = COPY_NONE;
    Object *copy_src = nullptr, *copy_dst = nullptr;
    uint64_t src_off = 0, dst_off = 0, copy_bytes = 0;   // byte offsets; B2B / clear size
    uint32_t src_pitch = 0, dst_pitch = 0;                // buffer bytes per row (texel copies)
    uint32_t sx = 0, sy = 0, dx = 0, dy = 0, cw = 0, ch = 0;   // texel origins and extent
    WgcTexture_T *color = nullptr;
    uint32_t color_load = 0;
    double clear_color[4]{};
    WgcTexture_T *depth = nullptr;
    uint32_t has_depth_ops = 0, depth_load = 0;
    float clear_depth = 0.0f;
    std::vector<SubCommand> cmds;

    void drop()
    {
        release(copy_src);
        release(copy_dst);
        copy_src = copy_dst = nullptr;
        release(color);
        release(depth);
        for (auto &c : cmds) release(c.obj);
        cmds.clear();
        color = depth = nullptr;
    }
};

// One rasterising draw with the pass it belongs to.
struct DrawWork {
    DrawParams d;
    const PassCommand *pass;
};

enum class StepKind { Group, Clear, Copy, Lines };

struct Step {
    StepKind kind = StepKind::Group;
    bool multi = false;
    std::vector<DrawParams> draws;   // the group's draws (scratch filled when laid out)
    GroupLaunch gl{};
    BatchArgs batch{};
    size_t table_off = 0;            // offset of this group's tables in the plan's device block
    uint32_t stage = 0;              // steps of one stage are independent and run concurrently
    size_t arena_off = 0;            // this group's scratch within the arena (past the zero region)
    size_t zone_off = 0;             // this group's words in the arena's zero region (not shared
                                     // with the plan's other groups; cleared when the submission ends)
    size_t keep_off = 0;             // this group's split-tile winners in the kept-zero region
    // clear
    uint32_t *color = nullptr;
    uint32_t rgba = 0;
    float *depth = nullptr;
    float dv = 0.0f;
    uint64_t n = 0;
    // copy: rows of row_bytes (or a zero fill of row_bytes when src is null)
    const uint8_t *copy_src = nullptr;
    uint8_t *copy_dst = nullptr;
    uint64_t src_pitch = 0, dst_pitch = 0, row_bytes = 0;
    uint32_t rows = 0;
};

struct Plan {
    WgcDevice_T *dev = nullptr;      // set (and retained) while the plan holds pooled memory or a graph
    uint64_t arena_gen = UINT64_MAX;
    size_t zone_used = 0;            // bytes of the arena's zero region its launch groups use
    size_t keep_used = 0;            // bytes of the kept-zero region its launch groups use
    uint32_t raster_cfg = UINT32_MAX;
    bool batching = false;
    bool noalloc = false;            // laid out in the device's static scratch (no tables, no graph)
    std::vector<Step> steps;
    uint8_t *d_tables = nullptr;     // from the device's table pool
    uint8_t *h_tables = nullptr;     // pinned staging of the upload, from the staging pool
    size_t tables_size = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t prims = 0, launches = 0;
    uint64_t submits = 0;            // the graph is captured on the second submission
    int hip_device = 0;

    void reset();
    ~Plan() { reset(); }
};

struct WgcCommandBuffer_T : Object {
    std::vector<PassCommand> passes;
    std::unique_ptr<Plan> plan{new Plan()};  // launch plan cached for resubmission
    // the same passes laid out in the static scratch, for the allocation-free
    // submission (built on its first use)
    std::unique_ptr<Plan> plan_noalloc{new Plan()};
    ~WgcCommandBuffer_T() override
    {
        plan.reset();
        plan_noalloc.reset();
        for (auto &p : passes) p.drop();
    }
};

struct WgcCommandEncoder_T : Object {
    WgcDevice_T *dev = nullptr;
    std::vector<PassCommand> passes;
    bool pass_open = false;
    ~WgcCommandEncoder_T() override
    {
        for (auto &p : passes) p.drop();
    }
};

struct WgcRenderPass_T : Object {
    WgcCommandEncoder_T *enc = nullptr;
    PassCommand pass;
    ~WgcRenderPass_T() override { pass.drop(); }
};

// Device-side counters of a submission; summed into a pinned host slot at
// its end (one 16-byte store), so that finishing one submission can be
// observed without draining the stream.
struct DevCounters {
    unsigned long long frags[kFragSlots];
    uint32_t error;
    uint32_t pad[63];
};
struct HostCounters {
    uint32_t frags_lo, frags_hi, error, pad;   // written by k_end_submission
    uint32_t pad2[12];
};

struct Inflight {
    uint64_t index;
    hipEvent_t done;
    std::vector<Object *> keep;  // command buffers used by the submission
    HostCounters *slot;          // pinned host copy of the counters
    uint64_t prims, launches;
};

// Pooled memory whose release, and graphs whose destruction, wait until the
// submissions before them have executed (they may still read the memory or
// run the graph).
struct Deferred {
    uint64_t after;              // submission index that must have retired
    BlockPool *pool;
    uint8_t *ptr;
    uint64_t size;
    hipGraph_t graph;
    hipGraphExec_t exec;

    void finish()
    {
        if (pool) pool->release(ptr, size);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

struct WgcDevice_T : Object {
    int hip_device = 0;
    uint32_t semantics = 0;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;     // second stream for the independent groups of a stage
    hipEvent_t fork = nullptr, join = nullptr;
    std::mutex mu;
    // per-draw scratch arena, reused by every draw (they run in stream order);
    // its first zone_size bytes are zero when a plan starts (the kernel that
    // ends a submission clears what the submission's plan used there)
    uint8_t *arena = nullptr;
    size_t arena_size = 0;
    size_t zone_size = 0;
    size_t keep_size = 0;   // then keep_size bytes that every use leaves zero (split-tile winners)
    // persistent device words
    DevCounters *d_counters = nullptr;
    bool counters_dirty = false;            // a submission stopped before taking its counters
    uint32_t *d_error = nullptr;            // &d_counters->error
    unsigned long long *d_frags = nullptr;  // d_counters->frags
    float *d_srgb = nullptr;
    uint64_t next_index = 1;
    uint64_t last_done = 0;
    std::vector<Inflight> inflight;
    std::vector<hipEvent_t> event_pool;
    std::vector<HostCounters *> slot_pool;   // free pinned counter slots
    std::vector<HostCounters *> slot_blocks;
    BlockPool table_pool;                   // device draw tables of plans
    BlockPool staging_pool;                 // pinned staging of their uploads
    std::mutex deferred_mu;                 // plans are reset from any thread
    std::vector<Deferred> deferred;
    // address of the static scratch array on this device (allocation-free
    // submissions lay their plans out in it)
    uint8_t *scratch = nullptr;
    std::vector<WgcBuffer_T *> pending_maps;   // held until their copies complete
    uint32_t pending_error = 0;             // error bits of retired submissions not yet reported
    uint64_t last_frags = 0;                // fragments of the last retired submission
    uint64_t total_frags = 0, total_submissions = 0;
    hipEvent_t ts[64]{};
    bool ts_recorded[64]{};
    bool timing = false;
    KernelTimer timer;
    std::map<std::string, std::pair<double, uint64_t>> ktimes;
    uint64_t last_prims = 0, last_launches = 0;
    uint64_t arena_gen = 0;        // bumped whenever memory a cached plan points at moves
    uint32_t cu_count = 256;        // compute units of the device
    uint32_t raster_cfg = RASTER_16x16_256;
    bool batching = true;          // merge independent single-draw passes into launch groups
    // replay cached plans as hipGraphs (WGC_OPTION_GRAPHS).  Off by default:
    // on this ROCm a graph launch costs 22-27 us of host time per
    // submission against 7 us for the same launches made directly, and the
    // GPU runs the replayed launches slower (the 1080p pass 47.1 against
    // 41.3 us per frame, the 256x256 pass 32.2 against 26.4)
    bool use_graphs = false;
    uint32_t *frag_px = nullptr;   // optional per-pixel fragment counts (debug)
    size_t frag_px_size = 0;

    ~WgcDevice_T() override
    {
        (void)hipSetDevice(hip_device);
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto &f : inflight) {
            (void)hipEventDestroy(f.done);
            for (auto *o : f.keep) release(o);
        }
        for (auto &x : deferred) {   // the stream has drained: nothing reads them any more
            x.pool = nullptr;
            x.finish();
        }
        deferred.clear();
        table_pool.destroy();
        staging_pool.destroy();
        for (auto *b : slot_blocks) (void)hipHostFree(b);
        for (auto e : event_pool) (void)hipEventDestroy(e);
        for (auto e : ts)
            if (e) (void)hipEventDestroy(e);
        for (auto &p : timer.pending) {
            (void)hipEventDestroy(p.begin);
            (void)hipEventDestroy(p.end);
        }
        for (auto e : timer.pool) (void)hipEventDestroy(e);
        if (arena) (void)hipFree(arena);
        if (d_counters) (void)hipFree(d_counters);
        if (d_srgb) (void)hipFree(d_srgb);
        if (frag_px) (void)hipFree(frag_px);
        for (auto *b : pending_maps) release(b);
        if (fork) (void)hipEventDestroy(fork);
        if (join) (void)hipEventDestroy(join);
        if (side) (void)hipStreamDestroy(side);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

// Gives a sub-allocation back to its heap.  Every submission that used the
// resource held a reference until it executed, so the bytes are free now.
static void heap_give_back(WgcHeap_T *heap, uint8_t *ptr, uint64_t size)
{
    {
        std::lock_guard<std::mutex> lk(heap->mu);
        heap->ra.release((uint64_t)(ptr - heap->base), size);
    }
    release(heap);
}

WgcBuffer_T::~WgcBuffer_T()
{
    if (map.host || map.done) {
        (void)hipSetDevice(dev->hip_device);
        if (map.done) {
            (void)hipEventSynchronize(map.done);
            (void)hipEventDestroy(map.done);
        }
        if (map.host) (void)hipHostFree(map.host);
    }
    if (heap) {
        heap_give_back(heap, ptr, size);
    } else if (ptr) {
        (void)hipSetDevice(dev->hip_device);
        (void)hipFree(ptr);
    }
    release(dev);
}

WgcTexture_T::~WgcTexture_T()
{
    if (heap) {
        heap_give_back(heap, ptr, size);
    } else if (ptr) {
        (void)hipSetDevice(dev->hip_device);
        (void)hipFree(ptr);
    }
    release(dev);
}

WgcHeap_T::~WgcHeap_T()
{
    if (base) {
        (void)hipSetDevice(dev->hip_device);
        (void)hipFree(base);
    }
    release(dev);
}

// Releases a plan's pooled memory once the submissions enqueued so far have
// executed, and the graph.
void Plan::reset()
{
    if (dev) {
        // every submission enqueued so far may still read the tables or run
        // the graph: they go once those have executed
        {
            std::lock_guard<std::mutex> lk(dev->deferred_mu);
            uint64_t after = dev->next_index - 1;
            if (d_tables)
                dev->deferred.push_back(Deferred{after, &dev->table_pool, d_tables, tables_size, nullptr, nullptr});
            if (h_tables)
                dev->deferred.push_back(Deferred{after, &dev->staging_pool, h_tables, tables_size, nullptr, nullptr});
            if (graph || exec) dev->deferred.push_back(Deferred{after, nullptr, nullptr, 0, graph, exec});
        }
        release(dev);   // may destroy the device (and with it what was deferred)
    } else {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
    dev = nullptr;
    exec = nullptr;
    graph = nullptr;
    d_tables = nullptr;
    h_tables = nullptr;
    tables_size = 0;
    submits = 0;
    steps.clear();
    arena_gen = UINT64_MAX;
}

// ------------------------------------------------------------------
// device
// ------------------------------------------------------------------

static float srgb_entry(int i)
{
    double c = (double)i / 255.0;
    double l = (c <= 0.04045) ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4);
    return (float)l;
}

static hipError_t take_slot(WgcDevice_T *d, HostCounters *&slot);
static constexpr int kPreparedEvents = 64;

extern "C" int32_t wgc_device_create(const WgcDeviceDescriptor *desc, WgcDevice *out)
{
    if (!desc || !out) return set_error(WGC_ERR_INVALID, "null argument");
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n), "hipGetDeviceCount");
    if (desc->hip_device < 0 || desc->hip_device >= n)
        return set_error(WGC_ERR_INVALID, "no such HIP device");
    if (desc->semantics > WGC_SEMANTICS_EXTENDED) return set_error(WGC_ERR_INVALID, "bad semantics");
    auto *d = new WgcDevice_T();
    d->hip_device = desc->hip_device;
    d->semantics = desc->semantics;
    hipError_t e = hipSetDevice(d->hip_device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&d->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&d->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&d->join, hipEventDisableTiming);
    if (e == hipSuccess) {
        int cus = 0;
        e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, d->hip_device);
        if (e == hipSuccess && cus > 0) d->cu_count = (uint32_t)cus;
    }
    if (e == hipSuccess) e = scratch_address(d->hip_device, d->scratch);
    if (e == hipSuccess) e = hipMalloc(&d->d_counters, sizeof(DevCounters));
    if (e == hipSuccess) e = hipMalloc(&d->d_srgb, 256 * sizeof(float));
    if (e == hipSuccess) e = stream_zero(d->stream, d->d_counters, sizeof(DevCounters));
    if (e == hipSuccess) {
        d->d_error = &d->d_counters->error;
        d->d_frags = d->d_counters->frags;
        d->table_pool.block_size = 16ull << 20;
        d->staging_pool.pinned = true;
        d->staging_pool.block_size = 16ull << 20;
    }
    if (e == hipSuccess) {
        float lut[256];
        for (int i = 0; i < 256; ++i) lut[i] = srgb_entry(i);
        e = stream_copy(d->stream, d->d_srgb, lut, sizeof lut, hipMemcpyHostToDevice);
    }
    // the first block of pinned counter slots and a pool of completion
    // events, so that the allocation-free submission finds them (it waits
    // for a submission to retire rather than creating more)
    if (e == hipSuccess) {
        HostCounters *slot = nullptr;
        e = take_slot(d, slot);
        if (e == hipSuccess) d->slot_pool.push_back(slot);
    }
    for (int i = 0; i < kPreparedEvents && e == hipSuccess; ++i) {
        hipEvent_t ev = nullptr;
        e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e == hipSuccess) d->event_pool.push_back(ev);
    }
    if (e != hipSuccess) {
        release(d);
        return hip_fail(e, "device setup");
    }
    *out = d;
    return WGC_OK;
}

static int32_t retire(WgcDevice_T *d, uint64_t upto, bool wait, uint64_t timeout_ns);

// Waits for the device's submissions first, so that they release the command
// buffers (and plans) they hold.
extern "C" void wgc_device_release(WgcDevice device)
{
    if (!device) return;
    {
        std::lock_guard<std::mutex> lk(device->mu);
        (void)hipSetDevice(device->hip_device);
        (void)retire(device, UINT64_MAX, true, 0);
    }
    release(device);
}

static hipError_t take_slot(WgcDevice_T *d, HostCounters *&slot)
{
    if (d->slot_pool.empty()) {
        void *p = nullptr;
        hipError_t e = hipHostMalloc(&p, 64 * sizeof(HostCounters), hipHostMallocDefault);
        if (e != hipSuccess) return e;
        d->slot_blocks.push_back((HostCounters *)p);
        for (int i = 0; i < 64; ++i) d->slot_pool.push_back((HostCounters *)p + i);
    }
    slot = d->slot_pool.back();
    d->slot_pool.pop_back();
    return hipSuccess;
}

// Drains finished submissions; with wait, blocks for `upto`.
static int32_t retire(WgcDevice_T *d, uint64_t upto, bool wait, uint64_t timeout_ns)
{
    auto t0 = std::chrono::steady_clock::now();
    while (!d->inflight.empty() && d->inflight.front().index <= upto) {
        Inflight &f = d->inflight.front();
        hipError_t q = hipEventQuery(f.done);
        if (q == hipErrorNotReady) {
            if (!wait) return WGC_OK;
            if (timeout_ns) {
                auto el = std::chrono::duration_cast<std::chrono::nanoseconds>(
                              std::chrono::steady_clock::now() - t0).count();
                if ((uint64_t)el >= timeout_ns) return set_error(WGC_ERR_TIMEOUT, "poll timed out");
                std::this_thread::sleep_for(std::chrono::microseconds(20));
                continue;
            }
            q = hipEventSynchronize(f.done);
        }
        if (q != hipSuccess) return hip_fail(q, "submission");
        d->last_done = f.index;
        d->event_pool.push_back(f.done);
        d->pending_error |= f.slot->error;
        uint64_t frags = (uint64_t)f.slot->frags_lo | ((uint64_t)f.slot->frags_hi << 32);
        d->last_frags = frags;
        d->total_frags += frags;
        d->total_submissions += 1;
        d->last_prims = f.prims;
        d->last_launches = f.launches;
        d->slot_pool.push_back(f.slot);
        std::vector<Object *> keep = std::move(f.keep);
        d->inflight.erase(d->inflight.begin());
        for (auto *o : keep) release(o);  // may reset plans, which defer their memory
    }
    // memory whose last reader has executed goes back to its pool, and
    // graphs that no submission runs any more are destroyed
    std::lock_guard<std::mutex> lk(d->deferred_mu);
    for (size_t i = 0; i < d->deferred.size();) {
        Deferred &x = d->deferred[i];
        if (x.after <= d->last_done) {
            x.finish();
            x = d->deferred.back();
            d->deferred.pop_back();
        } else {
            ++i;
        }
    }
    return WGC_OK;
}

static int32_t resolve_timers(WgcDevice_T *d)
{
    for (auto &p : d->timer.pending) {
        float ms = 0.0f;
        HIP_TRY(hipEventSynchronize(p.end), "timer");
        HIP_TRY(hipEventElapsedTime(&ms, p.begin, p.end), "timer");
        auto &slot = d->ktimes[p.name];
        slot.first += ms;
        slot.second += 1;
        d->timer.pool.push_back(p.begin);
        d->timer.pool.push_back(p.end);
    }
    d->timer.pending.clear();
    return WGC_OK;
}

// Reports (once) the error bits of the submissions retired so far.
static int32_t check_device_error(WgcDevice_T *d)
{
    uint32_t err = d->pending_error;
    if (!err) return WGC_OK;
    d->pending_error = 0;
    std::string msg = "render pass failed on the device:";
    if (err & ERR_INDEX) msg += " index out of the index buffer (reference: panic);";
    if (err & ERR_VERTEX) msg += " vertex fetch out of the vertex buffer (reference: panic);";
    if (err & ERR_W_ZERO) msg += " clip position w == 0 (reference: expect(\"w=0\"));";
    if (err & ERR_CLIP_OVERFLOW) msg += " clipped sub-triangle buffer full;";
    if (err & ERR_RECORD_OVERFLOW) msg += " triangle record buffer full;";
    if (err & ERR_COORD_RANGE) msg += " framebuffer coordinate beyond 2^30;";
    if (err & ERR_PRIM_RANGE) msg += " more than 2^25 primitives in one draw;";
    if (err & ERR_SHADER)
        msg += " a shader invocation aborted: integer division by zero, overflow or an index out of bounds "
               "(reference: panic);";
    return set_error(WGC_ERR_EXECUTION, msg);
}

struct MapCall {
    WgcBufferMapCallback fn;
    void *userdata;
    int32_t status;
};
static std::vector<MapCall> finish_maps(WgcDevice_T *d, bool wait);

extern "C" int32_t wgc_device_poll(WgcDevice device, uint64_t index, uint64_t timeout_ns)
{
    if (!device) return set_error(WGC_ERR_INVALID, "null device");
    std::vector<MapCall> calls;
    int32_t r;
    {
        std::lock_guard<std::mutex> lk(device->mu);
        HIP_TRY(hipSetDevice(device->hip_device), "hipSetDevice");
        r = retire(device, index, true, timeout_ns);
        if (r == WGC_OK && device->timing) r = resolve_timers(device);
        // mapping requests: a blocking poll waits for them like for submissions
        if (!device->pending_maps.empty()) calls = finish_maps(device, r == WGC_OK && timeout_ns != 1);
        if (r == WGC_OK) r = check_device_error(device);
    }
    for (auto &c : calls) c.fn(c.status, c.userdata);   // the map callbacks, outside the lock
    return r;
}

static int32_t drain(WgcDevice_T *d)
{
    int32_t r = retire(d, UINT64_MAX, true, 0);
    if (r != WGC_OK) return r;
    HIP_TRY(hipStreamSynchronize(d->stream), "hipStreamSynchronize");
    return WGC_OK;
}

// Totals over every retired submission; does not wait.
extern "C" int32_t wgc_device_totals(WgcDevice device, uint64_t *fragments, uint64_t *submissions)
{
    if (!device) return set_error(WGC_ERR_INVALID, "null device");
    std::lock_guard<std::mutex> lk(device->mu);
    HIP_TRY(hipSetDevice(device->hip_device), "hipSetDevice");
    int32_t r = retire(device, UINT64_MAX, false, 0);
    if (r != WGC_OK) return r;
    if (fragments) *fragments = device->total_frags;
    if (submissions) *submissions = device->total_submissions;
    return WGC_OK;
}

extern "C" int32_t wgc_device_last_pass_stats(WgcDevice device, WgcPassStats *out)
{
    if (!device || !out) return set_error(WGC_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(device->mu);
    HIP_TRY(hipSetDevice(device->hip_device), "hipSetDevice");
    int32_t r = drain(device);
    if (r != WGC_OK) return r;
    out->fragments = device->last_frags;
    out->primitives = device->last_prims;
    out->kernel_launches = device->last_launches;
    return WGC_OK;
}

// Optional per-pixel fragment counts: allocates a width*height u32 buffer
// that every later pass adds its emitted fragments to (debug aid).
extern "C" int32_t wgc_device_debug_fragment_counts(WgcDevice device, uint32_t width,
                                                    uint32_t height, uint32_t *host_out)
{
    if (!device) return set_error(WGC_ERR_INVALID, "null device");
    std::lock_guard<std::mutex> lk(device->mu);
    HIP_TRY(hipSetDevice(device->hip_device), "hipSetDevice");
    int32_t r = drain(device);
    if (r != WGC_OK) return r;
    size_t need = (size_t)width * height * 4u;
    if (host_out) {
        if (!device->frag_px || need != device->frag_px_size)
            return set_error(WGC_ERR_INVALID, "fragment counts not enabled for this size");
        HIP_TRY(stream_copy(device->stream, host_out, device->frag_px, need, hipMemcpyDeviceToHost), "readback");
        return WGC_OK;
    }
    device->arena_gen++;  // cached plans point at the counts buffer
    if (device->frag_px) {
        HIP_TRY(hipFree(device->frag_px), "hipFree");
        device->frag_px = nullptr;
        device->frag_px_size = 0;
    }
    if (need) {
        HIP_TRY(hipMalloc(&device->frag_px, need), "hipMalloc");
        HIP_TRY(stream_zero(device->stream, device->frag_px, need), "hipMemset");
        device->frag_px_size = need;
    }
    return WGC_OK;
}

// ------------------------------------------------------------------
// buffers and textures
// ------------------------------------------------------------------

extern "C" int32_t wgc_buffer_create(WgcDevice device, uint64_t size, WgcBuffer *out)
{
    if (!device || !out) return set_error(WGC_ERR_INVALID, "null argument");
    HIP_TRY(hipSetDevice(device->hip_device), "hipSetDevice");
    auto *b = new WgcBuffer_T();
    b->dev = device;
    retain(device);
    b->size = size;
    if (size) {
        hipError_t e = hipMalloc(&b->ptr, size);
        if (e == hipSuccess) {
            std::lock_guard<std::mutex> lk(device->mu);
            e = stream_zero_async(device->stream, b->ptr, size);
        }
        if (e != hipSuccess) {
            release(b);
            return set_error(WGC_ERR_OUT_OF_MEMORY, std::string("buffer: ") + hipGetErrorString(e));
        }
    }
    *out = b;
    return WGC_OK;
}

extern "C" void wgc_buffer_release(WgcBuffer buffer) { release(buffer); }

// Queue writes are ordered after every earlier submission on the device
// stream and before every later one; the caller's memory is free to reuse on
// return, like the reference's synchronous write (device.rs:332-344).  Up to
// 64 MiB the bytes go through pooled pinned staging and the host does not
// wait (the staging returns to its pool once the next submission has
// executed); larger writes copy from the caller's memory and wait.
static hipError_t staged_upload(WgcDevice_T *d, uint8_t *dst, const void *src, uint64_t n)
{
    if (n <= (64ull << 20)) {
        uint8_t *stage = nullptr;
        hipError_t e = d->staging_pool.alloc(n, stage);
        if (e != hipSuccess) return e;
        memcpy(stage, src, n);
        e = hipMemcpyAsync(dst, stage, n, hipMemcpyHostToDevice, d->stream);
        d->deferred.push_back(Deferred{d->next_index, &d->staging_pool, stage, n});
        return e;
    }
    hipError_t e = hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, d->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(d->stream);
    return e;
}

extern "C" int32_t wgc_queue_write_buffer(WgcDevice device, WgcBuffer buffer, uint64_t offset,
                                          const void *data, uint64_t size)
{
    if (!device || !buffer || (!data && size)) return set_error(WGC_ERR_INVALID, "null argument");
    if (offset > buffer->size || size > buffer->size - offset)
        return set_error(WGC_ERR_INVALID, "write_buffer out of range (reference: slice panic)");
    if (!size) return WGC_OK;
    std::lock_guard<std::mutex> lk(device->mu);
    HIP_TRY(hipSetDevice(device->hip_device), "hipSetDevice");
    HIP_TRY(staged_upload(device, buffer->ptr + offset, data, size), "write_buffer");
    return WGC_OK;
}

extern "C" int32_t wgc_buffer_read(WgcBuffer buffer, uint64_t offset, void *dst, uint64_t size)
{
    if (!buffer || (!dst && size)) return set_error(WGC_ERR_INVALID, "null argument");
    if (offset > buffer->size || size > buffer->size - offset)
        return set_error(WGC_ERR_INVALID, "read out of range");
    WgcDevice_T *d = buffer->dev;
    std::lock_guard<std::mutex> lk(d->mu);
    HIP_TRY(hipSetDevice(d->hip_device), "hipSetDevice");
    int32_t r = drain(d);
    if (r != WGC_OK) return r;
    HIP_TRY(stream_copy(d->stream, dst, buffer->ptr + offset, size, hipMemcpyDeviceToHost), "buffer read");
    return WGC_OK;
}

// ---- mapping (buffer.rs:112-167) ----

extern "C" int32_t wgc_buffer_map_async(WgcBuffer buffer, uint32_t mode, uint64_t offset, uint64_t size,
                                        WgcBufferMapCallback callback, void *userdata)
{
    if (!buffer) return set_error(WGC_ERR_INVALID, "null buffer");
    if (mode != WGC_MAP_READ && mode != WGC_MAP_WRITE) return set_error(WGC_ERR_INVALID, "bad map mode");
    if (buffer->map.mode || buffer->map.pending)
        return set_error(WGC_ERR_INVALID, "Buffer already mapped (buffer.rs:119-124)");
    if (offset % 8u || size % 4u) return set_error(WGC_ERR_INVALID, "map offset must be a multiple of 8, size of 4");
    if (offset > buffer->size || size > buffer->size - offset) return set_error(WGC_ERR_INVALID, "map range outside the buffer");
    WgcDevice_T *d = buffer->dev;
    std::lock_guard<std::mutex> lk(d->mu);
    HIP_TRY(hipSetDevice(d->hip_device), "hipSetDevice");
    MapState &m = buffer->map;
    if (size) HIP_TRY(hipHostMalloc((void **)&m.host, size, hipHostMallocDefault), "map staging");
    hipError_t e = hipEventCreateWithFlags(&m.done, hipEventDisableTiming);
    // a write mapping starts from the buffer's contents as well
    if (e == hipSuccess && size) e = hipMemcpyAsync(m.host, buffer->ptr + offset, size, hipMemcpyDeviceToHost, d->stream);
    if (e == hipSuccess) e = hipEventRecord(m.done, d->stream);
    if (e != hipSuccess) {
        if (m.done) (void)hipEventDestroy(m.done);
        if (m.host) (void)hipHostFree(m.host);
        m = MapState{};
        return hip_fail(e, "map_async");
    }
    m.pending = true;
    m.mode = mode;
    m.offset = offset;
    m.size = size;
    m.callback = callback;
    m.userdata = userdata;
    retain(buffer);
    d->pending_maps.push_back(buffer);
    return WGC_OK;
}

extern "C" int32_t wgc_buffer_get_mapped_range(WgcBuffer buffer, uint64_t offset, uint64_t size, void **data)
{
    if (!buffer || !data) return set_error(WGC_ERR_INVALID, "null argument");
    const MapState &m = buffer->map;
    if (!m.mode || m.pending) return set_error(WGC_ERR_INVALID, "Buffer is not mapped (buffer.rs:162)");
    if (offset < m.offset || offset - m.offset > m.size || size > m.size - (offset - m.offset))
        return set_error(WGC_ERR_INVALID, "range outside the mapped range");
    *data = m.host + (offset - m.offset);
    return WGC_OK;
}

extern "C" int32_t wgc_buffer_unmap(WgcBuffer buffer)
{
    if (!buffer) return set_error(WGC_ERR_INVALID, "null buffer");
    MapState &m = buffer->map;
    if (!m.mode) return set_error(WGC_ERR_INVALID, "Buffer is not mapped (buffer.rs:171)");
    WgcDevice_T *d = buffer->dev;
    WgcBufferMapCallback cb = nullptr;
    void *ud = nullptr;
    {
        std::lock_guard<std::mutex> lk(d->mu);
        HIP_TRY(hipSetDevice(d->hip_device), "hipSetDevice");
        if (m.pending) {
            // unmapped before the mapping completed: the request is aborted
            cb = m.callback;
            ud = m.userdata;
            for (size_t i = 0; i < d->pending_maps.size(); ++i)
                if (d->pending_maps[i] == buffer) {
                    d->pending_maps.erase(d->pending_maps.begin() + i);
                    release(buffer);   // the pending list's reference (the caller holds its own)
                    break;
                }
            (void)hipEventSynchronize(m.done);
        } else if (m.mode == WGC_MAP_WRITE && m.size) {
            // the written bytes go back to the buffer, ordered before later submissions
            HIP_TRY(stream_copy(d->stream, buffer->ptr + m.offset, m.host, m.size, hipMemcpyHostToDevice), "unmap");
        }
        (void)hipEventDestroy(m.done);
        if (m.host) (void)hipHostFree(m.host);
        m = MapState{};
    }
    if (cb) cb(WGC_ERR_INVALID, ud);
    return WGC_OK;
}

// Completes the mapping requests whose copies have finished (waiting for
// them when `wait`); returns the callbacks to run once the device lock is
// released.
static std::vector<MapCall> finish_maps(WgcDevice_T *d, bool wait)
{
    std::vector<MapCall> calls;
    std::vector<WgcBuffer_T *> still;
    for (WgcBuffer_T *b : d->pending_maps) {
        hipError_t e = wait ? hipEventSynchronize(b->map.done) : hipEventQuery(b->map.done);
        if (e == hipErrorNotReady) {
            still.push_back(b);
            continue;
        }
        b->map.pending = false;
        if (b->map.callback)
            calls.push_back(MapCall{b->map.callback, b->map.userdata, e == hipSuccess ? WGC_OK : WGC_ERR_DEVICE});
        b->map.callback = nullptr;
        release(b);
    }
    d->pending_maps.swap(still);
    return calls;
}

static uint32_t texel_bytes(uint32_t format)
{
    switch (format) {
    case WGC_TEXTURE_FORMAT_RGBA8_UNORM:
    case WGC_TEXTURE_FORMAT_RGBA8_UNORM_SRGB:
    case WGC_TEXTURE_FORMAT_BGRA8_UNORM:
    case WGC_TEXTURE_FORMAT_BGRA8_UNORM_SRGB:
    case WGC_TEXTURE_FORMAT_DEPTH32_FLOAT:
        return 4;
    default:
        return 0;
    }
}

extern "C" int32_t wgc_texture_create(WgcDevice device, uint32_t width, uint32_t height,
                                      uint32_t format, WgcTexture *out)
{
    if (!device || !out) return set_error(WGC_ERR_INVALID, "null argument");
    uint32_t tb = texel_bytes(format);
    if (!tb) return set_error(WGC_ERR_UNSUPPORTED, "unsupported texture format");
    // tile coordinates are 16-bit with 8-pixel tiles
    if (width == 0 || height == 0 || width > 65535u * 8u || height > 65535u * 8u)
        return set_error(WGC_ERR_INVALID, "bad texture size");
    HIP_TRY(hipSetDevice(device->hip_device), "hipSetDevice");
    auto *t = new WgcTexture_T();
    t->dev = device;
    retain(device);
    t->width = width;
    t->height = height;
    t->format = format;
    t->size = (uint64_t)width * height * tb;
    hipError_t e = hipMalloc(&t->ptr, t->size);
    if (e == hipSuccess) {
        std::lock_guard<std::mutex> lk(device->mu);
        e = stream_zero_async(device->stream, t->ptr, t->size);
    }
    if (e != hipSuccess) {
        release(t);
        return set_error(WGC_ERR_OUT_OF_MEMORY, std::string("texture: ") + hipGetErrorString(e));
    }
    *out = t;
    return WGC_OK;
}

extern "C" void wgc_texture_release(WgcTexture texture) { release(texture); }

// ------------------------------------------------------------------
// heaps
// ------------------------------------------------------------------

extern "C" int32_t wgc_heap_create(WgcDevice device, uint64_t size, WgcHeap *out)
{
    if (!device || !out || !size) return set_error(WGC_ERR_INVALID, "bad heap arguments");
    HIP_TRY(hipSetDevice(device->hip_device), "hipSetDevice");
    auto *h = new WgcHeap_T();
    h->dev = device;
    retain(device);
    size = (size + RangeAllocator::kAlign - 1) / RangeAllocator::kAlign * RangeAllocator::kAlign;
    if (hipMalloc(&h->base, size) != hipSuccess) {
        h->base = nullptr;
        release(h);
        return set_error(WGC_ERR_OUT_OF_MEMORY, "heap: cannot allocate " + std::to_string(size) + " bytes");
    }
    h->ra.init(size);
    *out = h;
    return WGC_OK;
}

extern "C" void wgc_heap_release(WgcHeap heap) { release(heap); }

extern "C" int32_t wgc_heap_stats(WgcHeap heap, uint64_t *used, uint64_t *capacity, uint64_t *largest_free)
{
    if (!heap) return set_error(WGC_ERR_INVALID, "null heap");
    std::lock_guard<std::mutex> lk(heap->mu);
    if (used) *used = heap->ra.used();
    if (capacity) *capacity = heap->ra.capacity();
    if (largest_free) *largest_free = heap->ra.largest_free();
    return WGC_OK;
}

static int32_t heap_take(WgcHeap_T *heap, uint64_t size, uint8_t *&ptr)
{
    std::lock_guard<std::mutex> lk(heap->mu);
    uint64_t off = 0;
    if (!heap->ra.alloc(size, off))
        return set_error(WGC_ERR_OUT_OF_MEMORY, "heap full (" + std::to_string(heap->ra.used()) + " of " +
                                                    std::to_string(heap->ra.capacity()) + " bytes in use)");
    ptr = heap->base + off;
    return WGC_OK;
}

extern "C" int32_t wgc_buffer_create_in_heap(WgcHeap heap, uint64_t size, WgcBuffer *out)
{
    if (!heap || !out || !size) return set_error(WGC_ERR_INVALID, "bad argument");
    WgcDevice_T *device = heap->dev;
    uint8_t *ptr = nullptr;
    int32_t r = heap_take(heap, size, ptr);
    if (r != WGC_OK) return r;
    auto *b = new WgcBuffer_T();
    b->dev = device;
    retain(device);
    b->heap = heap;
    retain(heap);
    b->ptr = ptr;
    b->size = size;
    std::lock_guard<std::mutex> lk(device->mu);
    HIP_TRY(hipSetDevice(device->hip_device), "hipSetDevice");
    hipError_t e = stream_zero_async(device->stream, ptr, size);
    if (e != hipSuccess) {
        release(b);
        return hip_fail(e, "buffer zero fill");
    }
    *out = b;
    return WGC_OK;
}

extern "C" int32_t wgc_texture_create_in_heap(WgcHeap heap, uint32_t width, uint32_t height, uint32_t format,
                                              WgcTexture *out)
{
    if (!heap || !out) return set_error(WGC_ERR_INVALID, "null argument");
    uint32_t tb = texel_bytes(format);
    if (!tb) return set_error(WGC_ERR_UNSUPPORTED, "unsupported texture format");
    if (width == 0 || height == 0 || width > 65535u * 8u || height > 65535u * 8u)
        return set_error(WGC_ERR_INVALID, "bad texture size");
    WgcDevice_T *device = heap->dev;
    uint64_t size = (uint64_t)width * height * tb;
    uint8_t *ptr = nullptr;
    int32_t r = heap_take(heap, size, ptr);
    if (r != WGC_OK) return r;
    auto *t = new WgcTexture_T();
    t->dev = device;
    retain(device);
    t->heap = heap;
    retain(heap);
    t->ptr = ptr;
    t->width = width;
    t->height = height;
    t->format = format;
    t->size = size;
    std::lock_guard<std::mutex> lk(device->mu);
    HIP_TRY(hipSetDevice(device->hip_device), "hipSetDevice");
    hipError_t e = stream_zero_async(device->stream, ptr, size);
    if (e != hipSuccess) {
        release(t);
        return hip_fail(e, "texture zero fill");
    }
    *out = t;
    return WGC_OK;
}

extern "C" int32_t wgc_queue_write_texture(WgcDevice device, WgcTexture texture, const void *data,
                                           uint64_t data_size, uint64_t offset, uint32_t bytes_per_row)
{
    if (!device || !texture || !data) return set_error(WGC_ERR_INVALID, "null argument");
    // device.rs:406-430: rows of bytes_per_row bytes, copied back to back
    uint64_t bpr = bytes_per_row ? bytes_per_row : data_size;
    uint64_t rows = bytes_per_row ? texture->height : 1u;
    uint64_t n = bpr * rows;
    if (offset > data_size || n > data_size - offset)
        return set_error(WGC_ERR_INVALID, "write_texture source too small (reference: slice panic)");
    if (n > texture->size)
        return set_error(WGC_ERR_INVALID, "write_texture past the texture (reference: slice panic)");
    std::lock_guard<std::mutex> lk(device->mu);
    HIP_TRY(hipSetDevice(device->hip_device), "hipSetDevice");
    HIP_TRY(staged_upload(device, texture->ptr, (const uint8_t *)data + offset, n), "write_texture");
    return WGC_OK;
}

extern "C" int32_t wgc_texture_read(WgcTexture texture, void *dst, uint64_t size)
{
    if (!texture || !dst) return set_error(WGC_ERR_INVALID, "null argument");
    if (size != texture->size) return set_error(WGC_ERR_INVALID, "size must equal the texture size");
    WgcDevice_T *d = texture->dev;
    std::lock_guard<std::mutex> lk(d->mu);
    HIP_TRY(hipSetDevice(d->hip_device), "hipSetDevice");
    int32_t r = drain(d);
    if (r != WGC_OK) return r;
    HIP_TRY(stream_copy(d->stream, dst, texture->ptr, size, hipMemcpyDeviceToHost), "texture read");
    return WGC_OK;
}

// ------------------------------------------------------------------
// samplers, bind groups, shaders, pipelines
// ------------------------------------------------------------------

extern "C" int32_t wgc_sampler_create(WgcDevice device, const WgcSamplerDescriptor *desc, WgcSampler *out)
{
    if (!device || !desc || !out) return set_error(WGC_ERR_INVALID, "null argument");
    if (desc->address_mode_u > 2 || desc->address_mode_v > 2)
        return set_error(WGC_ERR_UNSUPPORTED, "address mode (ClampToBorder is todo!() in the reference)");
    auto *s = new WgcSampler_T();
    s->desc = *desc;
    *out = s;
    return WGC_OK;
}

extern "C" void wgc_sampler_release(WgcSampler sampler) { release(sampler); }

extern "C" int32_t wgc_bind_group_create(WgcDevice device, uint32_t n, const WgcBindGroupEntry *entries,
                                         WgcBindGroup *out)
{
    if (!device || !out || (n && !entries)) return set_error(WGC_ERR_INVALID, "null argument");
    auto *g = new WgcBindGroup_T();
    for (uint32_t i = 0; i < n; ++i) {
        const WgcBindGroupEntry &e = entries[i];
        BindingSlot s;
        s.binding = e.binding;
        s.kind = e.kind;
        if (e.kind == WGC_BINDING_BUFFER) {
            if (!e.buffer || e.offset > e.buffer->size) {
                release(g);
                return set_error(WGC_ERR_INVALID, "bad buffer binding");
            }
            s.buffer = e.buffer;
            retain(e.buffer);
            s.offset = e.offset;
            s.size = e.size ? e.size : e.buffer->size - e.offset;
            if (s.size > e.buffer->size - e.offset) {
                release(g);
                return set_error(WGC_ERR_INVALID, "buffer binding past the end");
            }
        } else if (e.kind == WGC_BINDING_TEXTURE) {
            if (!e.texture) { release(g); return set_error(WGC_ERR_INVALID, "null texture"); }
            s.texture = e.texture;
            retain(e.texture);
        } else if (e.kind == WGC_BINDING_SAMPLER) {
            if (!e.sampler) { release(g); return set_error(WGC_ERR_INVALID, "null sampler"); }
            s.sampler = e.sampler;
            retain(e.sampler);
        } else {
            release(g);
            return set_error(WGC_ERR_INVALID, "bad binding kind");
        }
        g->entries.push_back(s);
    }
    *out = g;
    return WGC_OK;
}

extern "C" void wgc_bind_group_release(WgcBindGroup group) { release(group); }

extern "C" int32_t wgc_shader_module_create_builtin(WgcDevice device, uint32_t program,
                                                    uint32_t interpolation, WgcShaderModule *out)
{
    if (!device || !out) return set_error(WGC_ERR_INVALID, "null argument");
    if (program < PROG_MVP_COLOR || program > PROG_CLIP_COLOR)
        return set_error(WGC_ERR_INVALID, "unknown built-in program");
    if (interpolation > WGC_INTERPOLATE_FLAT) return set_error(WGC_ERR_INVALID, "bad interpolation");
    auto *m = new WgcShaderModule_T();
    m->program = program;
    m->interpolation = interpolation;
    *out = m;
    return WGC_OK;
}

extern "C" void wgc_shader_module_release(WgcShaderModule module) { release(module); }

static bool is_unsupported_message(const std::string &err)
{
    return err.find("todo!()") != std::string::npos || err.find("not supported") != std::string::npos ||
           err.find("not implemented") != std::string::npos;
}

extern "C" int32_t wgc_shader_module_create_wgsl(WgcDevice device, const char *source, size_t length,
                                                 WgcShaderModule *out)
{
    if (!device || !source || !out) return set_error(WGC_ERR_INVALID, "null argument");
    auto m = std::make_shared<wgsl::Module>();
    std::string err;
    if (!wgsl::translate(std::string(source, length), *m, err))
        return set_error(is_unsupported_message(err) ? WGC_ERR_UNSUPPORTED : WGC_ERR_INVALID, err);
    auto *sm = new WgcShaderModule_T();
    sm->program = PROG_USER;
    sm->wgsl = m;
    *out = sm;
    return WGC_OK;
}

extern "C" int32_t wgc_shader_module_run(WgcDevice device, WgcShaderModule module, const char *entry_point,
                                         uint32_t first_index, uint32_t count, const WgcBindGroup *groups,
                                         uint32_t group_count, uint32_t *results, size_t results_words)
{
    if (!device || !module || (!results && count)) return set_error(WGC_ERR_INVALID, "null argument");
    if (!module->wgsl) return set_error(WGC_ERR_INVALID, "not a WGSL module");
    if (results_words < (size_t)count * kRunWords) return set_error(WGC_ERR_INVALID, "results buffer too small");
    if (group_count > 4 || (group_count && !groups)) return set_error(WGC_ERR_INVALID, "bad bind groups");
    std::string name = entry_point ? entry_point : "";
    std::shared_ptr<EntryRun> run;
    {
        std::lock_guard<std::mutex> lk(module->mu);
        auto it = module->runs.find(name);
        if (it != module->runs.end()) {
            run = it->second;
        } else {
            run = std::make_shared<EntryRun>();
            std::string err;
            if (!build_entry_run(device->hip_device, *module->wgsl, name, *run, err))
                return set_error(err.compare(0, 3, "hip") == 0 ? WGC_ERR_DEVICE : WGC_ERR_INVALID, err);
            module->runs[name] = run;
        }
    }
    DrawParams d;
    memset(&d, 0, sizeof d);
    bool extended = device->semantics == WGC_SEMANTICS_EXTENDED;
    d.srgb = extended ? 1u : 0u;
    d.srgb_lut = device->d_srgb;
    if (run->slots.size() > (size_t)kMaxRes) return set_error(WGC_ERR_UNSUPPORTED, "too many resource bindings");
    for (size_t k = 0; k < run->slots.size(); ++k) {
        const wgsl::Resource &r = run->slots[k];
        const BindingSlot *b = r.group < group_count && groups[r.group] ? groups[r.group]->find(r.binding) : nullptr;
        if (!b) continue;   // unbound: only an error if the entry point reads it (reference: panic there)
        ResBind &rb = d.res[k];
        if (b->kind == WGC_BINDING_BUFFER) {
            if (b->size < r.min_size) return set_error(WGC_ERR_INVALID, "bound buffer smaller than " + r.name);
            rb.ptr = b->buffer->ptr + b->offset;
            rb.size = b->size;
        } else if (b->kind == WGC_BINDING_TEXTURE) {
            rb.ptr = b->texture->ptr;
            rb.w = b->texture->width;
            rb.h = b->texture->height;
            rb.srgb = b->texture->format == WGC_TEXTURE_FORMAT_RGBA8_UNORM_SRGB;
        } else {
            const WgcSamplerDescriptor &sd = b->sampler->desc;
            rb.addr_u = sd.address_mode_u;
            rb.addr_v = sd.address_mode_v;
            rb.bilinear = (extended && (sd.mag_filter == WGC_FILTER_LINEAR || sd.min_filter == WGC_FILTER_LINEAR)) ? 1u : 0u;
        }
    }
    if (!count) return WGC_OK;
    std::lock_guard<std::mutex> lk(device->mu);
    HIP_TRY(hipSetDevice(device->hip_device), "hipSetDevice");
    int32_t rc = drain(device);
    if (rc != WGC_OK) return rc;
    uint32_t *d_out = nullptr;
    size_t bytes = (size_t)count * kRunWords * 4u;
    HIP_TRY(hipMalloc(&d_out, bytes), "run results");
    void *args[] = {&d, &first_index, &count, &d_out};
    hipError_t e = hipModuleLaunchKernel(run->fn, (count + 63u) / 64u, 1, 1, 64, 1, 1, 0, device->stream, args, nullptr);
    if (e == hipSuccess) e = stream_copy(device->stream, results, d_out, bytes, hipMemcpyDeviceToHost);
    (void)hipFree(d_out);
    if (e != hipSuccess) return hip_fail(e, "entry-point run");
    return WGC_OK;
}

extern "C" int32_t wgc_wgsl_check_pipeline(const char *source, size_t length, const char *vertex_entry,
                                           const char *fragment_entry, uint32_t semantics)
{
    if (!source) return set_error(WGC_ERR_INVALID, "null argument");
    wgsl::Module m;
    std::string err;
    if (!wgsl::translate(std::string(source, length), m, err))
        return set_error(is_unsupported_message(err) ? WGC_ERR_UNSUPPORTED : WGC_ERR_INVALID, err);
    UserProgram up;
    bool has_fs = false;
    for (auto &e : m.entries) has_fs = has_fs || e.stage == wgsl::Stage::Fragment;
    if (!build_user_program(-1, m, vertex_entry ? vertex_entry : "", has_fs ? &m : nullptr,
                            fragment_entry ? fragment_entry : "", semantics == WGC_SEMANTICS_EXTENDED, up, err))
        return set_error(err.compare(0, 5, "hipRT") == 0 ? WGC_ERR_DEVICE : WGC_ERR_INVALID, err);
    return WGC_OK;
}

extern "C" int32_t wgc_wgsl_translate(const char *source, size_t length, uint32_t what, char *out, size_t capacity,
                                      size_t *needed)
{
    if (!source || what > 1) return set_error(WGC_ERR_INVALID, "bad argument");
    wgsl::Module m;
    std::string err;
    if (!wgsl::translate(std::string(source, length), m, err))
        return set_error(is_unsupported_message(err) ? WGC_ERR_UNSUPPORTED : WGC_ERR_INVALID, err);
    std::string text = what == 0 ? m.code : reflection_json(m);
    if (needed) *needed = text.size() + 1;
    if (out && capacity) {
        size_t n = std::min(capacity - 1, text.size());
        memcpy(out, text.data(), n);
        out[n] = '\0';
    }
    return WGC_OK;
}

extern "C" int32_t wgc_render_pipeline_create(WgcDevice device, const WgcRenderPipelineDescriptor *desc,
                                              WgcRenderPipeline *out)
{
    if (!device || !desc || !out || !desc->vertex_module) return set_error(WGC_ERR_INVALID, "null argument");
    if (desc->topology > WGC_TOPOLOGY_TRIANGLE_STRIP) return set_error(WGC_ERR_INVALID, "bad topology");
    if (desc->strip_index_format != 0 && desc->strip_index_format != WGC_INDEX_UINT16 &&
        desc->strip_index_format != WGC_INDEX_UINT32)
        return set_error(WGC_ERR_INVALID, "bad strip_index_format");
    if (desc->strip_index_format && desc->topology != WGC_TOPOLOGY_LINE_STRIP &&
        desc->topology != WGC_TOPOLOGY_TRIANGLE_STRIP)
        return set_error(WGC_ERR_INVALID, "strip_index_format set, but not a strip topology (state.rs:306-310)");
    if (desc->vertex_buffer_count > (uint32_t)kMaxVb) return set_error(WGC_ERR_UNSUPPORTED, "too many vertex buffers");
    bool user = desc->vertex_module->program == PROG_USER;
    if (desc->fragment_module && (desc->fragment_module->program == PROG_USER) != user)
        return set_error(WGC_ERR_INVALID, "a WGSL stage cannot be paired with a built-in program");
    if (!user && desc->fragment_module && desc->fragment_module->program != desc->vertex_module->program)
        return set_error(WGC_ERR_INVALID, "vertex and fragment entry points come from different programs");
    auto *p = new WgcRenderPipeline_T();
    p->desc = *desc;
    p->desc.vertex_buffers = nullptr;
    p->program = desc->vertex_module->program;
    p->interp = desc->vertex_module->interpolation;
    p->has_fragment = desc->fragment_module != nullptr;
    // vertex.rs:60-85: shader locations -> (buffer, offset, size)
    bool seen[kMaxAttr] = {};
    for (uint32_t b = 0; b < desc->vertex_buffer_count; ++b) {
        const WgcVertexBufferLayout &l = desc->vertex_buffers[b];
        p->vb_stride.push_back(l.array_stride);
        p->vb_step.push_back(l.step_mode);
        for (uint32_t a = 0; a < l.attribute_count; ++a) {
            const WgcVertexAttribute &at = l.attributes[a];
            if (at.shader_location >= (uint32_t)kMaxAttr || at.format < 1 || at.format > 4 || (at.offset & 3u)) {
                release(p);
                return set_error(WGC_ERR_UNSUPPORTED, "vertex attribute");
            }
            if (seen[at.shader_location]) {
                release(p);
                return set_error(WGC_ERR_INVALID, "duplicate vertex attribute location (reference: panic)");
            }
            seen[at.shader_location] = true;
            p->attr[at.shader_location] = AttrDev{b, at.offset, at.format, 0u};
            p->n_attr = std::max(p->n_attr, at.shader_location + 1u);
        }
    }
    if (user) {
        auto up = std::make_shared<UserProgram>();
        std::string err;
        const WgcShaderModule_T *fm = desc->fragment_module;
        if (!build_user_program(device->hip_device, *desc->vertex_module->wgsl,
                                desc->vertex_entry ? desc->vertex_entry : "", fm ? fm->wgsl.get() : nullptr,
                                desc->fragment_entry ? desc->fragment_entry : "",
                                device->semantics == WGC_SEMANTICS_EXTENDED, *up, err)) {
            release(p);
            return set_error(err.compare(0, 5, "hipRT") == 0 || err.compare(0, 5, "hipMo") == 0 ? WGC_ERR_DEVICE
                                                                                             : WGC_ERR_INVALID,
                             err);
        }
        // vertex.rs:139-152 copies attribute_size bytes into the input: the
        // attribute must have exactly the input's size
        for (const wgsl::IoVar &in : up->vertex_inputs) {
            if (in.location >= p->n_attr || !seen[in.location] || p->attr[in.location].ncomp != in.ncomp) {
                release(p);
                return set_error(WGC_ERR_INVALID, "vertex input @location(" + std::to_string(in.location) +
                                                      ") has no vertex attribute of its size");
            }
        }
        p->nvar = up->nvar;
        p->user = up;
        if (desc->has_depth_stencil && desc->depth_format != WGC_TEXTURE_FORMAT_DEPTH32_FLOAT) {
            release(p);
            return set_error(WGC_ERR_UNSUPPORTED, "only Depth32Float depth is implemented");
        }
        if (desc->has_depth_stencil && (desc->depth_compare < 1 || desc->depth_compare > 8)) {
            release(p);
            return set_error(WGC_ERR_INVALID, "bad depth compare");
        }
        *out = p;
        return WGC_OK;
    }
    // the built-in programs' vertex inputs and varyings
    auto need = [&](uint32_t loc, uint32_t lo, uint32_t hi) {
        return loc < p->n_attr && seen[loc] && p->attr[loc].ncomp >= lo && p->attr[loc].ncomp <= hi;
    };
    switch (p->program) {
    case PROG_MVP_COLOR:
    case PROG_CLIP_COLOR:
        if (!need(0, 4, 4) || !need(1, 3, 4)) {
            release(p);
            return set_error(WGC_ERR_INVALID, "program needs @location(0) vec4f position, @location(1) vec3f/vec4f colour");
        }
        p->nvar = p->attr[1].ncomp;
        break;
    case PROG_MVP_TEXTURE:
        if (!need(0, 4, 4) || !need(1, 2, 2)) {
            release(p);
            return set_error(WGC_ERR_INVALID, "program needs @location(0) vec4f position, @location(1) vec2f uv");
        }
        p->nvar = 2;
        break;
    default:
        p->nvar = 4;
        break;
    }
    if (p->interp == WGC_INTERPOLATE_PERSPECTIVE && device->semantics == WGC_SEMANTICS_REFERENCE) {
        release(p);
        return set_error(WGC_ERR_UNSUPPORTED,
                         "perspective interpolation is todo!() in the reference (fragment.rs:343-345); "
                         "use WGC_SEMANTICS_EXTENDED");
    }
    if (desc->has_depth_stencil && desc->depth_format != WGC_TEXTURE_FORMAT_DEPTH32_FLOAT) {
        release(p);
        return set_error(WGC_ERR_UNSUPPORTED, "only Depth32Float depth is implemented");
    }
    if (desc->has_depth_stencil && (desc->depth_compare < 1 || desc->depth_compare > 8)) {
        release(p);
        return set_error(WGC_ERR_INVALID, "bad depth compare");
    }
    *out = p;
    return WGC_OK;
}

extern "C" void wgc_render_pipeline_release(WgcRenderPipeline pipeline) { release(pipeline); }

// ------------------------------------------------------------------
// recording
// ------------------------------------------------------------------

extern "C" int32_t wgc_command_encoder_create(WgcDevice device, WgcCommandEncoder *out)
{
    if (!device || !out) return set_error(WGC_ERR_INVALID, "null argument");
    auto *e = new WgcCommandEncoder_T();
    e->dev = device;
    *out = e;
    return WGC_OK;
}

extern "C" void wgc_command_encoder_release(WgcCommandEncoder encoder) { release(encoder); }

extern "C" int32_t wgc_command_encoder_begin_render_pass(WgcCommandEncoder enc, const WgcRenderPassDescriptor *desc,
                                                         WgcRenderPass *out)
{
    if (!enc || !desc || !out) return set_error(WGC_ERR_INVALID, "null argument");
    if (enc->pass_open) return set_error(WGC_ERR_INVALID, "a render pass is already open");
    if (desc->color_attachment_count > 1)
        return set_error(WGC_ERR_UNSUPPORTED, "more than one colour attachment");
    auto *p = new WgcRenderPass_T();
    p->enc = enc;
    retain(enc);
    uint32_t w = 0, h = 0;
    bool sized = false;
    // state.rs:84-96: all attachments must be the same size
    auto size_of = [&](WgcTexture_T *t) -> bool {
        if (!sized) { w = t->width; h = t->height; sized = true; return true; }
        return t->width == w && t->height == h;
    };
    if (desc->color_attachment_count == 1) {
        const WgcColorAttachment &c = desc->color_attachments[0];
        if (!c.view || c.view->format == WGC_TEXTURE_FORMAT_DEPTH32_FLOAT) {
            release(p);
            return set_error(WGC_ERR_INVALID, "bad colour attachment");
        }
        size_of(c.view);
        p->pass.color = c.view;
        retain(c.view);
        p->pass.color_load = c.load_op;
        for (int k = 0; k < 4; ++k) p->pass.clear_color[k] = c.clear_value[k];
    }
    if (desc->depth_stencil_attachment) {
        const WgcDepthStencilAttachment &dd = *desc->depth_stencil_attachment;
        if (!dd.view || dd.view->format != WGC_TEXTURE_FORMAT_DEPTH32_FLOAT) {
            release(p);
            return set_error(WGC_ERR_INVALID, "bad depth attachment");
        }
        if (!size_of(dd.view)) {
            release(p);
            return set_error(WGC_ERR_INVALID, "All render attachments must be the same size (state.rs:88-91)");
        }
        p->pass.depth = dd.view;
        retain(dd.view);
        p->pass.has_depth_ops = dd.has_depth_ops;
        p->pass.depth_load = dd.depth_load_op;
        p->pass.clear_depth = dd.depth_clear_value;
    }
    enc->pass_open = true;
    *out = p;
    return WGC_OK;
}

static int32_t push(WgcRenderPass pass, SubCommand c)
{
    if (!pass) return set_error(WGC_ERR_INVALID, "null pass");
    pass->pass.cmds.push_back(c);
    return WGC_OK;
}

extern "C" int32_t wgc_render_pass_set_pipeline(WgcRenderPass pass, WgcRenderPipeline pipeline)
{
    if (!pipeline) return set_error(WGC_ERR_INVALID, "null pipeline");
    SubCommand c{CmdKind::SetPipeline};
    c.obj = pipeline;
    retain(pipeline);
    return push(pass, c);
}

extern "C" int32_t wgc_render_pass_set_bind_group(WgcRenderPass pass, uint32_t index, WgcBindGroup group)
{
    SubCommand c{CmdKind::SetBindGroup};
    c.slot = index;
    c.obj = group;
    retain(group);
    return push(pass, c);
}

extern "C" int32_t wgc_render_pass_set_vertex_buffer(WgcRenderPass pass, uint32_t slot, WgcBuffer buffer,
                                                     uint64_t offset, uint64_t size)
{
    if (!buffer || slot >= (uint32_t)kMaxVb) return set_error(WGC_ERR_INVALID, "bad vertex buffer");
    if (offset > buffer->size) return set_error(WGC_ERR_INVALID, "vertex buffer offset past the end");
    SubCommand c{CmdKind::SetVertexBuffer};
    c.obj = buffer;
    retain(buffer);
    c.slot = slot;
    c.offset = offset;
    c.size = size ? size : buffer->size - offset;
    if (c.size > buffer->size - offset) { release(buffer); return set_error(WGC_ERR_INVALID, "vertex buffer range"); }
    return push(pass, c);
}

extern "C" int32_t wgc_render_pass_set_index_buffer(WgcRenderPass pass, WgcBuffer buffer, uint32_t format,
                                                    uint64_t offset, uint64_t size)
{
    if (!buffer || (format != WGC_INDEX_UINT16 && format != WGC_INDEX_UINT32))
        return set_error(WGC_ERR_INVALID, "bad index buffer");
    if (offset > buffer->size) return set_error(WGC_ERR_INVALID, "index buffer offset past the end");
    SubCommand c{CmdKind::SetIndexBuffer};
    c.obj = buffer;
    retain(buffer);
    c.format = format;
    c.offset = offset;
    c.size = size ? size : buffer->size - offset;
    if (c.size > buffer->size - offset) { release(buffer); return set_error(WGC_ERR_INVALID, "index buffer range"); }
    return push(pass, c);
}

extern "C" int32_t wgc_render_pass_set_viewport(WgcRenderPass pass, float x, float y, float w, float h,
                                                float min_depth, float max_depth)
{
    SubCommand c{CmdKind::SetViewport};
    c.viewport[0] = x; c.viewport[1] = y; c.viewport[2] = w; c.viewport[3] = h;
    c.viewport[4] = min_depth; c.viewport[5] = max_depth;
    return push(pass, c);
}

extern "C" int32_t wgc_render_pass_set_scissor_rect(WgcRenderPass pass, uint32_t x, uint32_t y, uint32_t w,
                                                    uint32_t h)
{
    SubCommand c{CmdKind::SetScissor};
    c.rect[0] = x; c.rect[1] = y; c.rect[2] = w; c.rect[3] = h;
    return push(pass, c);
}

extern "C" int32_t wgc_render_pass_set_blend_constant(WgcRenderPass pass, const double *)
{
    return push(pass, SubCommand{CmdKind::SetBlendConstant});
}

extern "C" int32_t wgc_render_pass_set_stencil_reference(WgcRenderPass pass, uint32_t)
{
    return push(pass, SubCommand{CmdKind::SetStencilReference});
}

extern "C" int32_t wgc_render_pass_draw(WgcRenderPass pass, uint32_t first_vertex, uint32_t vertex_count,
                                        uint32_t first_instance, uint32_t instance_count)
{
    SubCommand c{CmdKind::Draw};
    c.first = first_vertex;
    c.count = vertex_count;
    c.first_instance = first_instance;
    c.instance_count = instance_count;
    return push(pass, c);
}

extern "C" int32_t wgc_render_pass_draw_indexed(WgcRenderPass pass, uint32_t first_index, uint32_t index_count,
                                                int32_t base_vertex, uint32_t first_instance,
                                                uint32_t instance_count)
{
    SubCommand c{CmdKind::DrawIndexed};
    c.first = first_index;
    c.count = index_count;
    c.base_vertex = base_vertex;
    c.first_instance = first_instance;
    c.instance_count = instance_count;
    return push(pass, c);
}

extern "C" int32_t wgc_render_pass_end(WgcRenderPass pass)
{
    if (!pass) return set_error(WGC_ERR_INVALID, "null pass");
    WgcCommandEncoder_T *enc = pass->enc;
    enc->passes.push_back(std::move(pass->pass));
    pass->pass = PassCommand();
    enc->pass_open = false;
    release(enc);
    pass->enc = nullptr;
    release(pass);
    return WGC_OK;
}

// ---- copies and clears between passes ----

static int32_t push_copy(WgcCommandEncoder_T *enc, PassCommand &&pc)
{
    if (enc->pass_open) {
        pc.drop();
        return set_error(WGC_ERR_INVALID, "a render pass is open on this encoder");
    }
    enc->passes.push_back(std::move(pc));
    return WGC_OK;
}

// The buffer side of a texel copy: bytes_per_row a multiple of 256 covering
// the row (may be 0 for one row), 4-byte offset, the last row inside.
static int32_t check_texel_buffer(const WgcTexelCopyBuffer &b, uint32_t w, uint32_t h, uint32_t &pitch)
{
    if (!b.buffer) return set_error(WGC_ERR_INVALID, "null buffer");
    uint64_t row = (uint64_t)w * 4u;
    pitch = b.bytes_per_row ? b.bytes_per_row : (uint32_t)row;
    if (h > 1 && b.bytes_per_row == 0) return set_error(WGC_ERR_INVALID, "bytes_per_row is required for more than one row");
    if (b.bytes_per_row % 256u) return set_error(WGC_ERR_INVALID, "bytes_per_row must be a multiple of 256");
    if (b.bytes_per_row && b.bytes_per_row < row) return set_error(WGC_ERR_INVALID, "bytes_per_row is shorter than a row");
    if (b.offset % 4u) return set_error(WGC_ERR_INVALID, "buffer offset must be a multiple of the texel size (4)");
    uint64_t need = h ? (uint64_t)pitch * (h - 1) + row : 0;
    if (b.offset > b.buffer->size || need > b.buffer->size - b.offset)
        return set_error(WGC_ERR_INVALID, "copy range outside the buffer");
    return WGC_OK;
}

static int32_t check_texel_texture(const WgcTexelCopyTexture &t, uint32_t w, uint32_t h)
{
    if (!t.texture) return set_error(WGC_ERR_INVALID, "null texture");
    if ((uint64_t)t.origin_x + w > t.texture->width || (uint64_t)t.origin_y + h > t.texture->height)
        return set_error(WGC_ERR_INVALID, "copy region outside the texture");
    return WGC_OK;
}

extern "C" int32_t wgc_command_encoder_copy_buffer_to_buffer(WgcCommandEncoder enc, WgcBuffer src, uint64_t src_off,
                                                             WgcBuffer dst, uint64_t dst_off, uint64_t size)
{
    if (!enc || !src || !dst) return set_error(WGC_ERR_INVALID, "null argument");
    if (src == dst) return set_error(WGC_ERR_INVALID, "source and destination are the same buffer");
    if (src_off % 4u || dst_off % 4u || size % 4u)
        return set_error(WGC_ERR_INVALID, "copy offsets and size must be multiples of 4");
    if (src_off > src->size || size > src->size - src_off || dst_off > dst->size || size > dst->size - dst_off)
        return set_error(WGC_ERR_INVALID, "copy range outside a buffer");
    PassCommand pc;
    pc.copy_kind = COPY_B2B;
    retain(src);
    retain(dst);
    pc.copy_src = src;
    pc.copy_dst = dst;
    pc.src_off = src_off;
    pc.dst_off = dst_off;
    pc.copy_bytes = size;
    return push_copy(enc, std::move(pc));
}

extern "C" int32_t wgc_command_encoder_copy_texture_to_buffer(WgcCommandEncoder enc, const WgcTexelCopyTexture *src,
                                                              const WgcTexelCopyBuffer *dst, uint32_t w, uint32_t h)
{
    if (!enc || !src || !dst) return set_error(WGC_ERR_INVALID, "null argument");
    uint32_t pitch = 0;
    int32_t r = check_texel_texture(*src, w, h);
    if (r == WGC_OK) r = check_texel_buffer(*dst, w, h, pitch);
    if (r != WGC_OK) return r;
    PassCommand pc;
    pc.copy_kind = COPY_T2B;
    retain(src->texture);
    retain(dst->buffer);
    pc.copy_src = src->texture;
    pc.copy_dst = dst->buffer;
    pc.sx = src->origin_x;
    pc.sy = src->origin_y;
    pc.dst_off = dst->offset;
    pc.dst_pitch = pitch;
    pc.cw = w;
    pc.ch = h;
    return push_copy(enc, std::move(pc));
}

extern "C" int32_t wgc_command_encoder_copy_buffer_to_texture(WgcCommandEncoder enc, const WgcTexelCopyBuffer *src,
                                                              const WgcTexelCopyTexture *dst, uint32_t w, uint32_t h)
{
    if (!enc || !src || !dst) return set_error(WGC_ERR_INVALID, "null argument");
    uint32_t pitch = 0;
    int32_t r = check_texel_texture(*dst, w, h);
    if (r == WGC_OK) r = check_texel_buffer(*src, w, h, pitch);
    if (r != WGC_OK) return r;
    PassCommand pc;
    pc.copy_kind = COPY_B2T;
    retain(src->buffer);
    retain(dst->texture);
    pc.copy_src = src->buffer;
    pc.copy_dst = dst->texture;
    pc.src_off = src->offset;
    pc.src_pitch = pitch;
    pc.dx = dst->origin_x;
    pc.dy = dst->origin_y;
    pc.cw = w;
    pc.ch = h;
    return push_copy(enc, std::move(pc));
}

static bool copy_compatible(uint32_t a, uint32_t b)
{
    auto base = [](uint32_t f) {
        return f == WGC_TEXTURE_FORMAT_RGBA8_UNORM_SRGB ? (uint32_t)WGC_TEXTURE_FORMAT_RGBA8_UNORM
             : f == WGC_TEXTURE_FORMAT_BGRA8_UNORM_SRGB ? (uint32_t)WGC_TEXTURE_FORMAT_BGRA8_UNORM : f;
    };
    return base(a) == base(b);
}

extern "C" int32_t wgc_command_encoder_copy_texture_to_texture(WgcCommandEncoder enc, const WgcTexelCopyTexture *src,
                                                               const WgcTexelCopyTexture *dst, uint32_t w, uint32_t h)
{
    if (!enc || !src || !dst) return set_error(WGC_ERR_INVALID, "null argument");
    int32_t r = check_texel_texture(*src, w, h);
    if (r == WGC_OK) r = check_texel_texture(*dst, w, h);
    if (r != WGC_OK) return r;
    if (src->texture == dst->texture) return set_error(WGC_ERR_INVALID, "source and destination are the same texture");
    if (!copy_compatible(src->texture->format, dst->texture->format))
        return set_error(WGC_ERR_INVALID, "texture formats are not copy-compatible");
    PassCommand pc;
    pc.copy_kind = COPY_T2T;
    retain(src->texture);
    retain(dst->texture);
    pc.copy_src = src->texture;
    pc.copy_dst = dst->texture;
    pc.sx = src->origin_x;
    pc.sy = src->origin_y;
    pc.dx = dst->origin_x;
    pc.dy = dst->origin_y;
    pc.cw = w;
    pc.ch = h;
    return push_copy(enc, std::move(pc));
}

extern "C" int32_t wgc_command_encoder_clear_buffer(WgcCommandEncoder enc, WgcBuffer buf, uint64_t offset, uint64_t size)
{
    if (!enc || !buf) return set_error(WGC_ERR_INVALID, "null argument");
    if (offset > buf->size) return set_error(WGC_ERR_INVALID, "clear range outside the buffer");
    if (size == 0) size = buf->size - offset;
    if (offset % 4u || size % 4u) return set_error(WGC_ERR_INVALID, "clear offset and size must be multiples of 4");
    if (size > buf->size - offset) return set_error(WGC_ERR_INVALID, "clear range outside the buffer");
    PassCommand pc;
    pc.copy_kind = CLEAR_BUFFER;
    retain(buf);
    pc.copy_dst = buf;
    pc.dst_off = offset;
    pc.copy_bytes = size;
    return push_copy(enc, std::move(pc));
}

extern "C" int32_t wgc_command_encoder_clear_texture(WgcCommandEncoder enc, WgcTexture tex)
{
    if (!enc || !tex) return set_error(WGC_ERR_INVALID, "null argument");
    PassCommand pc;
    pc.copy_kind = CLEAR_TEXTURE;
    retain(tex);
    pc.copy_dst = tex;
    pc.copy_bytes = tex->size;
    return push_copy(enc, std::move(pc));
}

extern "C" int32_t wgc_command_encoder_finish(WgcCommandEncoder enc, WgcCommandBuffer *out)
{
    if (!enc || !out) return set_error(WGC_ERR_INVALID, "null argument");
    if (enc->pass_open) return set_error(WGC_ERR_INVALID, "a render pass is still open");
    auto *cb = new WgcCommandBuffer_T();
    cb->passes = std::move(enc->passes);
    enc->passes.clear();
    *out = cb;
    return WGC_OK;
}

extern "C" void wgc_command_buffer_release(WgcCommandBuffer buffer) { release(buffer); }

// ------------------------------------------------------------------
// execution: planning (host) and launching (device stream)
// ------------------------------------------------------------------
//
// A command buffer is planned once per device state (arena generation,
// raster tile shape, batching switch): every draw becomes a DrawParams, and
// consecutive independent passes with one draw each are merged into launch
// groups that share all kernels (the batched form).  The plan is cached in
// the command buffer; a resubmission only launches it, as a hipGraph when
// graphs are enabled.

static uint8_t unorm8(float v)
{
    float x = v * 255.0f;
    if (x < 0.0f) x = 0.0f;
    if (x > 255.0f) x = 255.0f;
    if (!(x == x)) return 0;
    return (uint8_t)x;
}

static uint32_t pack_clear(const double c[4], uint32_t format)
{
    uint32_t r = unorm8((float)c[0]), g = unorm8((float)c[1]), b = unorm8((float)c[2]), a = unorm8((float)c[3]);
    bool bgra = format == WGC_TEXTURE_FORMAT_BGRA8_UNORM || format == WGC_TEXTURE_FORMAT_BGRA8_UNORM_SRGB;
    return bgra ? (b | (g << 8) | (r << 16) | (a << 24)) : (r | (g << 8) | (b << 16) | (a << 24));
}

static size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

struct PassState {
    WgcRenderPipeline_T *pipeline = nullptr;
    WgcBindGroup_T *groups[4] = {};
    BufferBinding vbs[kMaxVb];
    BufferBinding ib;
    float viewport[6];
    uint32_t scissor[4];
};

static uint32_t resolve_mode(const WgcRenderPipelineDescriptor &p, bool depth_test, bool blend)
{
    if (blend) return RES_ORDERED;
    if (!depth_test) return RES_LAST;
    bool w = p.depth_write_enabled != 0;
    switch (p.depth_compare) {
    case WGC_COMPARE_LESS: return w ? RES_MIN_FIRST : RES_PASS_LAST;
    case WGC_COMPARE_LESS_EQUAL: return w ? RES_MIN_LAST : RES_PASS_LAST;
    case WGC_COMPARE_GREATER: return w ? RES_MAX_FIRST : RES_PASS_LAST;
    case WGC_COMPARE_GREATER_EQUAL: return w ? RES_MAX_LAST : RES_PASS_LAST;
    case WGC_COMPARE_NOT_EQUAL: return w ? RES_ORDERED : RES_PASS_LAST;
    default: return RES_PASS_LAST;
    }
}

// Builds the kernel parameters of one draw from the pass state
// (DrawCall::execute, state.rs:239-480); scratch pointers come later.
static int32_t build_draw(WgcDevice_T *dev, const PassCommand &pc, const PassState &st, const SubCommand &c,
                          uint32_t w, uint32_t h, uint32_t tile_shift, DrawParams &d)
{
    memset(&d, 0, sizeof d);
    const WgcRenderPipeline_T *pl = st.pipeline;
    if (!pl) return set_error(WGC_ERR_INVALID, "No pipeline bound (state.rs:243)");
    const WgcRenderPipelineDescriptor &pd = pl->desc;
    bool extended = dev->semantics == WGC_SEMANTICS_EXTENDED;
    d.program = pl->program;
    d.nvar = pl->nvar;
    d.srgb_lut = dev->d_srgb;   // every draw: the tile kernels copy it to LDS
    d.interp = pl->interp;
    d.front_face_cw = pd.front_face;
    d.cull_mode = pd.cull_mode;
    d.depth_test = (pc.depth && pd.has_depth_stencil) ? 1u : 0u;
    d.depth_compare = pd.depth_compare;
    d.depth_write = pd.depth_write_enabled;
    d.blend = (extended && pd.blend == WGC_BLEND_ALPHA && pc.color) ? 1u : 0u;
    d.resolve = resolve_mode(pd, d.depth_test != 0, d.blend != 0);
    // every fragment must run its shader: a discard, an early depth test
    // (whose write a later discard keeps) or a shader-written depth makes the
    // outcome depend on each fragment in API order
    if (pl->user && (pl->user->discards || pl->user->early_depth || pl->user->frag_depth_tested))
        d.resolve = RES_ORDERED;
    // vertex buffers (vertex.rs:261-272)
    uint32_t n_vb = (uint32_t)pl->vb_stride.size();
    for (uint32_t b = 0; b < n_vb; ++b)
        if (!st.vbs[b].buffer)
            return set_error(WGC_ERR_INVALID, "Buffer " + std::to_string(b) + " not bound (vertex.rs:268)");
    d.n_attr = pl->n_attr;
    for (uint32_t a = 0; a < pl->n_attr; ++a) {
        const AttrDev &at = pl->attr[a];
        if (!at.ncomp) continue;  // location not provided by any buffer
        const BufferBinding &vb = st.vbs[at.buffer];
        d.attr[a] = AttrBind{vb.buffer->ptr + vb.offset, vb.size, pl->vb_stride[at.buffer],
                             pl->vb_step[at.buffer], at.offset, at.ncomp};
    }
    d.indexed = c.kind == CmdKind::DrawIndexed;
    if (d.indexed) {
        if (!st.ib.buffer) return set_error(WGC_ERR_INVALID, "No index buffer bound (state.rs:396)");
        d.index_data = st.ib.buffer->ptr + st.ib.offset;
        d.index_size = st.ib.size;
        d.index_format = st.ib.format;
        d.base_vertex = c.base_vertex;
    }
    d.first = c.first;
    d.count = c.count;
    d.first_instance = c.first_instance;
    d.instance_count = c.instance_count;
    // primitives per instance (primitive.rs:206-487); with restart an upper
    // bound, the primitives that span a separator are skipped on the device
    d.topology = pd.topology;
    d.strip_restart = pd.strip_index_format ? 1u : 0u;
    if (d.indexed && d.strip_restart && pd.strip_index_format != st.ib.format)
        return set_error(WGC_ERR_INVALID, "strip_index_format differs from the index format (state.rs:311)");
    switch (d.topology) {
    case TOPO_POINT_LIST: d.nprim = c.count; break;
    case TOPO_LINE_LIST: d.nprim = c.count / 2u; break;
    case TOPO_LINE_STRIP: d.nprim = c.count >= 2u ? c.count - 1u : 0u; break;
    case TOPO_TRIANGLE_STRIP: d.nprim = c.count >= 3u ? c.count - 2u : 0u; break;
    default: d.nprim = c.count / 3u; break;
    }
    if (d.topology <= TOPO_LINE_STRIP && d.topology != TOPO_POINT_LIST &&
        (uint64_t)d.nprim * c.instance_count > (1ull << 19))
        return set_error(WGC_ERR_UNSUPPORTED, "more than 2^19 lines in one draw");
    uint64_t total = (uint64_t)d.nprim * c.instance_count;
    if (total >= (1ull << 25)) return set_error(WGC_ERR_UNSUPPORTED, "more than 2^25 primitives in one draw");
    d.total_prims = (uint32_t)total;
    // bind groups of the built-in programs
    if (d.program == PROG_MVP_COLOR || d.program == PROG_MVP_TEXTURE) {
        const BindingSlot *u = st.groups[0] ? st.groups[0]->find(0) : nullptr;
        if (!u || u->kind != WGC_BINDING_BUFFER || u->size < 64)
            return set_error(WGC_ERR_INVALID, "@group(0) @binding(0) uniform mat4x4f not bound");
        d.uniform = (const float *)(u->buffer->ptr + u->offset);
    }
    if (d.program == PROG_MVP_TEXTURE) {
        const BindingSlot *t = st.groups[1] ? st.groups[1]->find(0) : nullptr;
        const BindingSlot *s = st.groups[1] ? st.groups[1]->find(1) : nullptr;
        if (!t || t->kind != WGC_BINDING_TEXTURE || !s || s->kind != WGC_BINDING_SAMPLER)
            return set_error(WGC_ERR_INVALID, "@group(1) texture_2d / sampler not bound");
        const WgcTexture_T *tex = t->texture;
        if (tex->format == WGC_TEXTURE_FORMAT_DEPTH32_FLOAT || tex->format == WGC_TEXTURE_FORMAT_BGRA8_UNORM ||
            tex->format == WGC_TEXTURE_FORMAT_BGRA8_UNORM_SRGB)
            return set_error(WGC_ERR_UNSUPPORTED, "texture format (texture.rs:186 todo!())");
        d.tex = tex->ptr;
        d.tex_w = tex->width;
        d.tex_h = tex->height;
        d.tex_srgb = tex->format == WGC_TEXTURE_FORMAT_RGBA8_UNORM_SRGB;
        d.addr_u = s->sampler->desc.address_mode_u;
        d.addr_v = s->sampler->desc.address_mode_v;
        d.bilinear = (extended && (s->sampler->desc.mag_filter == WGC_FILTER_LINEAR ||
                                   s->sampler->desc.min_filter == WGC_FILTER_LINEAR)) ? 1u : 0u;
        d.srgb = extended ? 1u : 0u;
        d.srgb_lut = dev->d_srgb;
    }
    if (pl->user) {
        // resource slots of the WGSL program (binding.rs:16-164)
        const UserProgram &up = *pl->user;
        d.user = &up.kernels;
        d.var_flat_mask = up.flat_mask;
        d.var_persp_mask = up.persp_mask;
        d.fs_early = up.early_depth;
        d.fs_depth_tested = up.frag_depth_tested ? 1u : 0u;
        d.srgb = extended ? 1u : 0u;
        d.srgb_lut = dev->d_srgb;
        for (size_t k = 0; k < up.slots.size(); ++k) {
            const UserProgram::Slot &sl = up.slots[k];
            std::string where = "@group(" + std::to_string(sl.group) + ") @binding(" + std::to_string(sl.binding) + ")";
            const BindingSlot *b = sl.group < 4 && st.groups[sl.group] ? st.groups[sl.group]->find(sl.binding) : nullptr;
            ResBind &r = d.res[k];
            switch (sl.kind) {
            case wgsl::ResKind::Uniform:
            case wgsl::ResKind::Storage:
                if (!b || b->kind != WGC_BINDING_BUFFER) return set_error(WGC_ERR_INVALID, where + ": buffer not bound");
                if (b->size < sl.min_size)
                    return set_error(WGC_ERR_INVALID, where + ": the bound buffer range is smaller than the variable");
                r.ptr = b->buffer->ptr + b->offset;
                r.size = b->size;
                break;
            case wgsl::ResKind::Texture: {
                if (!b || b->kind != WGC_BINDING_TEXTURE) return set_error(WGC_ERR_INVALID, where + ": texture not bound");
                const WgcTexture_T *tex = b->texture;
                if (tex->format != WGC_TEXTURE_FORMAT_RGBA8_UNORM && tex->format != WGC_TEXTURE_FORMAT_RGBA8_UNORM_SRGB)
                    return set_error(WGC_ERR_UNSUPPORTED, "texture format (texture.rs:186 todo!())");
                r.ptr = tex->ptr;
                r.w = tex->width;
                r.h = tex->height;
                r.srgb = tex->format == WGC_TEXTURE_FORMAT_RGBA8_UNORM_SRGB;
                break;
            }
            case wgsl::ResKind::Sampler: {
                if (!b || b->kind != WGC_BINDING_SAMPLER) return set_error(WGC_ERR_INVALID, where + ": sampler not bound");
                const WgcSamplerDescriptor &sd = b->sampler->desc;
                r.addr_u = sd.address_mode_u;
                r.addr_v = sd.address_mode_v;
                r.bilinear = (extended && (sd.mag_filter == WGC_FILTER_LINEAR || sd.min_filter == WGC_FILTER_LINEAR)) ? 1u : 0u;
                break;
            }
            }
        }
    }
    memcpy(d.viewport, st.viewport, sizeof d.viewport);
    memcpy(d.scissor, st.scissor, sizeof d.scissor);
    if ((uint64_t)d.scissor[0] + d.scissor[2] > w || (uint64_t)d.scissor[1] + d.scissor[3] > h)
        return set_error(WGC_ERR_INVALID, "scissor rect outside the framebuffer");
    d.width = w;
    d.height = h;
    // 16x16 tiles for the ordered (sorted, sequential) raster, whatever the
    // raster tile option: of the tile shapes it is the fastest for it (200k
    // blended triangles over 2560x2440: 0.63 ms, the others 0.74-3.0 ms) and
    // holds the most keys per pixel; WGSL pipelines' kernels are compiled for
    // 16x16 tiles only
    d.tile_shift = (d.resolve == RES_ORDERED || pl->user) ? 4u : tile_shift;
    // points and lines: unsigned pixel indices (their keys and visibility buffer)
    if (d.topology <= TOPO_LINE_STRIP && (uint64_t)w * h > 0xFFFFFFFFull)
        return set_error(WGC_ERR_UNSUPPORTED, "target too large for points and lines");
    tile_shift = d.tile_shift;
    d.tiles_x = (w + (1u << tile_shift) - 1) >> tile_shift;
    d.tiles_y = (h + (1u << tile_shift) - 1) >> tile_shift;
    d.color = pc.color ? pc.color->ptr : nullptr;
    d.color_bgra = pc.color && (pc.color->format == WGC_TEXTURE_FORMAT_BGRA8_UNORM ||
                                pc.color->format == WGC_TEXTURE_FORMAT_BGRA8_UNORM_SRGB);
    d.depth = pc.depth ? (float *)pc.depth->ptr : nullptr;
    return WGC_OK;
}

// Persistent raster grids: every workgroup resident at once (CUs x the
// kernel's occupancy), never more than there are work items, multiples of 8
// (the kernels split their work over the 8 XCDs).
static uint32_t resident_grid(WgcDevice_T *dev, uint32_t per_cu, uint64_t items)
{
    uint64_t cap = (uint64_t)dev->cu_count * per_cu;
    cap = std::max<uint64_t>(8u, cap / 8 * 8);
    uint64_t g = align_up(std::max<uint64_t>(1u, std::min<uint64_t>(items, cap)), 8);
    return (uint32_t)std::min(g, cap);
}

// Lays a group's scratch out at the start of the arena (groups run one after
// another on the stream and reuse it).  Returns the bytes it needs; with a
// base pointer, fills the pointers.
//
// Group-wide: tile counts and the active list (one word per tile of every
// draw), the bins (bin_cap keys per tile), the spill list, the clip list,
// the header and per-draw counters.  Per draw: triangle records by
// primitive, clipped sub-triangle records and their clip data, the
// big-triangle list and the first clip slot per primitive.
// A triangle strip drawn with restart separators: needs the separator
// summary pre-pass and is never merged into a batched group.
static bool restart_strip(const DrawParams &d)
{
    return d.topology == TOPO_TRIANGLE_STRIP && d.strip_restart && d.indexed;
}

static bool is_point_or_line(const DrawParams &d) { return d.topology <= TOPO_LINE_STRIP; }

// Scratch of a point / line draw: the per-pixel visibility values, or for
// an ordered draw the set-up primitives.
static size_t layout_lines(Step &st, uint8_t *base, uint32_t *frag_px, size_t frag_px_size, WgcDevice_T *dev)
{
    DrawParams &d = st.draws[0];
    bool ordered = d.resolve == RES_ORDERED;
    size_t bytes = ordered ? align_up(std::max<uint64_t>((uint64_t)kLpPrimBytes * d.total_prims, 256), 256)
                           : align_up(8ull * d.width * d.height, 256);
    if (!base) return bytes;
    st.gl.vis = ordered ? nullptr : (unsigned long long *)base;
    st.gl.lp = ordered ? (void *)base : nullptr;
    d.s.frags = dev->d_frags;
    d.s.error = dev->d_error;
    d.s.frag_px = (frag_px && frag_px_size == (size_t)d.width * d.height * 4u) ? frag_px : nullptr;
    return bytes;
}

static uint64_t group_tiles(const Step &st)
{
    uint64_t tiles = 0;
    for (const DrawParams &d : st.draws) tiles += (uint64_t)d.tiles_x * d.tiles_y;
    return tiles;
}

// A group's words in the zero region: its header, per-draw counters, tile
// counts and split-tile tickets.
static size_t zone_bytes(const Step &st)
{
    uint64_t tiles = group_tiles(st);
    return align_up(4ull * H_NUM, 256) + align_up(4ull * C_NUM * st.draws.size(), 256) + 2 * align_up(4ull * tiles, 256);
}

// A group's bytes in the kept-zero region: the merged winners of split
// tiles (order-independent groups only), 8 bytes per tile pixel.
static size_t keep_bytes(const Step &st)
{
    if (st.gl.ordered || st.multi || st.draws.empty()) return 0;
    uint64_t tile_px = 1ull << (2 * st.draws[0].tile_shift);
    return align_up(8ull * tile_px * group_tiles(st), 256);
}

// Bin entries per work item of a tile: one visibility batch of the group's
// tile shape, so that a tile of two batches or more spreads over several
// workgroups (a single pass has about one tile per resident workgroup, and
// its kernel lasts as long as its slowest tile) while ordinary ones stay
// whole.  Single order-independent draws only: ordered tiles are never
// split, and a batched group has many tiles per workgroup anyway.
static uint32_t split_for(const Step &st)
{
    if (st.gl.ordered || st.multi) return 0;
    uint32_t cfg = st.gl.user ? (uint32_t)RASTER_16x16_256 : st.gl.raster_cfg;
    return vis_batch_size(1u << raster_tile_shift(cfg), raster_threads(cfg));
}

static size_t layout_group(Step &st, uint8_t *base, uint8_t *zone, uint8_t *keep, uint32_t *frag_px,
                           size_t frag_px_size, WgcDevice_T *dev)
{
    uint32_t n = (uint32_t)st.draws.size();
    uint64_t total_tiles = 0, clip_rec_cap = 0, prims = 0, prims_padded = 0;
    bool any_clear = false;
    std::vector<uint64_t> clip_cap(n);
    for (uint32_t i = 0; i < n; ++i) {
        const DrawParams &d = st.draws[i];
        total_tiles += (uint64_t)d.tiles_x * d.tiles_y;
        clip_cap[i] = 2ull * d.total_prims + 1024ull;
        clip_rec_cap += clip_cap[i];
        prims += d.total_prims;
        prims_padded += st.multi ? align_up(d.total_prims, 64) : d.total_prims;
        any_clear = any_clear || d.init_clear;
    }
    // bins, per draw: slots for twice the tile's pixels, or four times the
    // draw's primitives per tile when that is more (a bin rarely fills; the
    // entries past it go to the spill list, which holds every entry at most)
    uint64_t tile_px = st.draws.empty() ? 0 : (1ull << (2 * st.draws[0].tile_shift));
    std::vector<uint64_t> bin_cap(n);
    uint64_t bins_total = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const DrawParams &d = st.draws[i];
        uint64_t tiles = std::max<uint64_t>((uint64_t)d.tiles_x * d.tiles_y, 1);
        uint64_t per_tile = (4ull * d.total_prims + tiles - 1) / tiles;
        uint64_t cap = 2 * tile_px;
        while (cap < per_tile) cap <<= 1;
        cap = std::min<uint64_t>(cap, d.total_prims + clip_cap[i]);   // no tile holds more
        bin_cap[i] = std::max<uint64_t>(std::min<uint64_t>(cap, 0xFFFFFFFFull / tiles), 1);
        bins_total += bin_cap[i] * tiles;
    }
    uint64_t spill_cap = (prims_padded + clip_rec_cap) * (uint64_t)kMaxBinTiles;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off = align_up(off + std::max<size_t>(bytes, 4), 256);
        return o;
    };
    size_t o_al = take(4 * total_tiles * kMaxItems);
    size_t o_bins = take(4 * bins_total), o_sp = take(8 * spill_cap);
    size_t o_cl = take(8 * prims);
    std::vector<size_t> o_big(n), o_tri(n), o_ctri(n), o_ct(n), o_cb(n), o_ss(n);
    for (uint32_t i = 0; i < n; ++i) {
        const DrawParams &d = st.draws[i];
        uint64_t rc = d.total_prims + clip_cap[i];
        o_ss[i] = take(restart_strip(d) ? 4ull * ((d.count + kStripChunk - 1) / kStripChunk) : 0);
        o_big[i] = take(16 * rc);
        o_tri[i] = take(4ull * kRecWords * d.total_prims);
        o_ctri[i] = take(4ull * kRecWords * clip_cap[i]);
        o_ct[i] = take(4ull * kClipFloats * clip_cap[i]);
        o_cb[i] = take(4ull * d.total_prims);
    }
    if (!base) return off;
    GroupLaunch &gl = st.gl;
    BatchArgs &b = st.batch;
    gl.total_tiles = (uint32_t)total_tiles;
    gl.n_counters = C_NUM * n;
    // zero region: header, counters, tile counts (zone_bytes)
    gl.header = (uint32_t *)zone;
    gl.counters = (uint32_t *)(zone + align_up(4ull * H_NUM, 256));
    gl.tile_count = (uint32_t *)(zone + align_up(4ull * H_NUM, 256) + align_up(4ull * C_NUM * n, 256));
    uint32_t *tile_done = gl.tile_count + align_up(4ull * total_tiles, 256) / 4;
    uint32_t split = split_for(st);
    gl.clip_list = (uint2 *)(base + o_cl);
    gl.clip_cap = (uint32_t)prims;
    // a single triangle draw whose setup workgroups are all resident at
    // once clips in setup (no separate clip launch)
    gl.inline_clip = 0;
    if (!st.multi && !is_point_or_line(st.draws[0])) {
        uint64_t blocks = (prims + 255) / 256;
        uint32_t per_cu = gl.user ? gl.user->occ_setup_clip : setup_clip_blocks_per_cu(gl.program);
        gl.inline_clip = blocks <= (uint64_t)dev->cu_count * per_cu ? 1u : 0u;
    }
    gl.vis = nullptr;
    gl.lp = nullptr;
    b.n = n;
    b.total_tiles = (uint32_t)total_tiles;
    b.n_chunks = (uint32_t)((total_tiles + kClearChunk - 1) / kClearChunk);
    b.header = gl.header;
    b.tile_count = gl.tile_count;
    b.active_list = (uint32_t *)(base + o_al);
    b.counters = gl.counters;
    b.n_counters = gl.n_counters;
    // a single draw clears its whole target in setup when setup has a thread
    // for every 32 pixels (then the stores drain while setup waits on its
    // loads and atomics); the tile kernel then skips the untouched tiles and
    // stores only the pixels the draw writes.  (A batched group leaves its
    // clears to the tile kernel, where they hide behind latency-bound work:
    // stored separately they cost more than they save.)
    bool preclear = false;
    if (!st.multi && any_clear && prims > 0) {
        uint64_t threads = align_up(prims, 256), px = (uint64_t)st.draws[0].width * st.draws[0].height;
        preclear = threads * 32 >= px;
    }
    b.preclear = preclear ? 1u : 0u;
    b.any_clear = (any_clear && !b.preclear) ? 1u : 0u;
    b.pad = 0;
    if (!st.multi) {
        b.draws = nullptr;
        b.prim_base = nullptr;
        b.tile_base = nullptr;
        b.wave_draw = nullptr;
        b.tile_draw = nullptr;
        b.total_prims = st.draws[0].total_prims;
    }
    uint32_t occ = 1;
    if (gl.user) {
        int m = st.multi ? 1 : 0;
        occ = gl.ordered ? gl.user->occ_raster[m] : gl.user->occ_tiles[m];
    } else {
        occ = raster_blocks_per_cu(gl.program, gl.ordered != 0, st.multi, gl.raster_cfg);
    }
    gl.raster_grid = resident_grid(dev, occ, total_tiles + b.n_chunks);
    uint64_t tile_base = 0, bin_base = 0;
    for (uint32_t i = 0; i < n; ++i) {
        DrawParams &d = st.draws[i];
        Scratch &s = d.s;
        d.preclear = (preclear && d.init_clear) ? 1u : 0u;
        s.tile_count = gl.tile_count + tile_base;
        s.tile_done = tile_done + tile_base;
        s.gvis = keep_bytes(st) ? (unsigned long long *)keep + tile_base * tile_px : nullptr;
        s.split = split;
        s.split_inv = split ? 1.0f / (float)split : 0.0f;
        s.bins = (uint32_t *)(base + o_bins) + bin_base;
        bin_base += bin_cap[i] * std::max<uint64_t>((uint64_t)d.tiles_x * d.tiles_y, 1);
        s.spill = (uint2 *)(base + o_sp);
        s.active_list = b.active_list;
        s.header = gl.header;
        s.clip_list = gl.clip_list;
        s.big = (uint4 *)(base + o_big[i]);
        s.tri = (uint32_t *)(base + o_tri[i]);
        s.clip_tri = (uint32_t *)(base + o_ctri[i]);
        s.clip_tris = (float *)(base + o_ct[i]);
        s.clip_base = (uint32_t *)(base + o_cb[i]);
        s.strip_sep = restart_strip(d) ? (int32_t *)(base + o_ss[i]) : nullptr;
        s.counters = gl.counters + C_NUM * i;
        s.frags = dev->d_frags;
        s.error = dev->d_error;
        // debug counts are only written when they were sized for this pass
        s.frag_px = (frag_px && frag_px_size == (size_t)d.width * d.height * 4u) ? frag_px : nullptr;
        s.tile0 = (uint32_t)tile_base;
        s.bin_cap = (uint32_t)bin_cap[i];
        s.spill_cap = (uint32_t)std::min<uint64_t>(spill_cap, 0xFFFFFFFFull);
        s.active_cap = (uint32_t)(total_tiles * kMaxItems);
        s.big_cap = (uint32_t)(d.total_prims + clip_cap[i]);
        s.clip_list_cap = (uint32_t)prims;
        s.clip_tri_cap = (uint32_t)clip_cap[i];
        tile_base += (uint64_t)d.tiles_x * d.tiles_y;
    }
    return off;
}

// Passes whose attachments and texture bindings do not meet may share a
// launch group.
static void pass_textures(const PassCommand &pc, std::vector<const void *> &writes, std::vector<const void *> &reads)
{
    if (pc.color) writes.push_back(pc.color);
    if (pc.depth) writes.push_back(pc.depth);
    for (const SubCommand &c : pc.cmds)
        if (c.kind == CmdKind::SetBindGroup && c.obj)
            for (const BindingSlot &e : ((WgcBindGroup_T *)c.obj)->entries)
                if (e.texture) reads.push_back(e.texture);
}

static int32_t collect_pass(WgcDevice_T *dev, const PassCommand &pc, uint32_t tile_shift, std::vector<DrawWork> &out,
                            bool &clear_only, Step &clear)
{
    uint32_t w = 0, h = 0;
    if (pc.color) { w = pc.color->width; h = pc.color->height; }
    else if (pc.depth) { w = pc.depth->width; h = pc.depth->height; }
    PassState st;
    st.viewport[0] = 0.0f; st.viewport[1] = 0.0f;
    st.viewport[2] = (float)w; st.viewport[3] = (float)h;
    st.viewport[4] = 0.0f; st.viewport[5] = 1.0f;
    st.scissor[0] = 0; st.scissor[1] = 0; st.scissor[2] = w; st.scissor[3] = h;
    bool clear_color = pc.color && pc.color_load == WGC_LOAD_OP_CLEAR;
    bool clear_depth = pc.depth && pc.has_depth_ops && pc.depth_load == WGC_LOAD_OP_CLEAR;
    bool clear_pending = clear_color || clear_depth;
    uint32_t clear_rgba = pc.color ? pack_clear(pc.clear_color, pc.color->format) : 0u;
    for (const SubCommand &c : pc.cmds) {
        switch (c.kind) {
        case CmdKind::SetPipeline: st.pipeline = (WgcRenderPipeline_T *)c.obj; break;
        case CmdKind::SetBindGroup:
            if (c.slot >= 4) return set_error(WGC_ERR_UNSUPPORTED, "bind group index >= 4");
            st.groups[c.slot] = (WgcBindGroup_T *)c.obj;
            break;
        case CmdKind::SetVertexBuffer: st.vbs[c.slot] = BufferBinding{(WgcBuffer_T *)c.obj, c.offset, c.size, 0}; break;
        case CmdKind::SetIndexBuffer: st.ib = BufferBinding{(WgcBuffer_T *)c.obj, c.offset, c.size, c.format}; break;
        case CmdKind::SetViewport: memcpy(st.viewport, c.viewport, sizeof st.viewport); break;
        case CmdKind::SetScissor: memcpy(st.scissor, c.rect, sizeof st.scissor); break;
        case CmdKind::SetBlendConstant:
        case CmdKind::SetStencilReference: break;
        case CmdKind::Draw:
        case CmdKind::DrawIndexed: {
            if (!st.pipeline) return set_error(WGC_ERR_INVALID, "No pipeline bound (state.rs:243)");
            if (!st.pipeline->has_fragment || w == 0) break;  // vertex shading only: no visible effect
            DrawWork dw;
            int32_t r = build_draw(dev, pc, st, c, w, h, tile_shift, dw.d);
            if (r != WGC_OK) return r;
            dw.pass = &pc;
            dw.d.init_clear = clear_pending ? 1u : 0u;
            dw.d.clear_color_on = clear_color;
            dw.d.clear_depth_on = clear_depth;
            dw.d.clear_rgba = clear_rgba;
            dw.d.clear_depth = pc.clear_depth;
            clear_pending = false;
            out.push_back(dw);
            break;
        }
        }
    }
    clear_only = clear_pending;
    if (clear_pending) {
        clear.kind = StepKind::Clear;
        clear.color = clear_color ? (uint32_t *)pc.color->ptr : nullptr;
        clear.rgba = clear_rgba;
        clear.depth = clear_depth ? (float *)pc.depth->ptr : nullptr;
        clear.dv = pc.clear_depth;
        clear.n = (uint64_t)w * h;
    }
    return WGC_OK;
}

static void add_group(std::vector<Step> &steps, std::vector<DrawParams> draws, uint32_t cfg, uint32_t stage)
{
    Step s;
    s.stage = stage;
    s.kind = StepKind::Group;
    s.multi = draws.size() > 1;
    s.gl.program = draws[0].program;
    s.gl.ordered = draws[0].resolve == RES_ORDERED;
    s.gl.user = (const UserKernels *)draws[0].user;
    s.gl.raster_cfg = s.gl.user ? (uint32_t)RASTER_16x16_256 : cfg;
    uint64_t prims = 0;
    for (auto &d : draws) prims += s.multi ? align_up(d.total_prims, 64) : d.total_prims;
    s.gl.total_prims = (uint32_t)prims;
    s.draws = std::move(draws);
    steps.push_back(std::move(s));
}

// The step of a copy or clear: rows of bytes between device addresses.
static void copy_step(const PassCommand &pc, Step &s)
{
    s.kind = StepKind::Copy;
    switch (pc.copy_kind) {
    case COPY_B2B: {
        auto *a = (WgcBuffer_T *)pc.copy_src, *b = (WgcBuffer_T *)pc.copy_dst;
        s.copy_src = a->ptr + pc.src_off;
        s.copy_dst = b->ptr + pc.dst_off;
        s.row_bytes = pc.copy_bytes;
        s.rows = pc.copy_bytes ? 1u : 0u;
        break;
    }
    case COPY_T2B: {
        auto *t = (WgcTexture_T *)pc.copy_src;
        auto *b = (WgcBuffer_T *)pc.copy_dst;
        s.copy_src = t->ptr + ((uint64_t)pc.sy * t->width + pc.sx) * 4u;
        s.src_pitch = (uint64_t)t->width * 4u;
        s.copy_dst = b->ptr + pc.dst_off;
        s.dst_pitch = pc.dst_pitch;
        s.row_bytes = (uint64_t)pc.cw * 4u;
        s.rows = pc.ch;
        break;
    }
    case COPY_B2T: {
        auto *b = (WgcBuffer_T *)pc.copy_src;
        auto *t = (WgcTexture_T *)pc.copy_dst;
        s.copy_src = b->ptr + pc.src_off;
        s.src_pitch = pc.src_pitch;
        s.copy_dst = t->ptr + ((uint64_t)pc.dy * t->width + pc.dx) * 4u;
        s.dst_pitch = (uint64_t)t->width * 4u;
        s.row_bytes = (uint64_t)pc.cw * 4u;
        s.rows = pc.ch;
        break;
    }
    case COPY_T2T: {
        auto *a = (WgcTexture_T *)pc.copy_src, *b = (WgcTexture_T *)pc.copy_dst;
        s.copy_src = a->ptr + ((uint64_t)pc.sy * a->width + pc.sx) * 4u;
        s.src_pitch = (uint64_t)a->width * 4u;
        s.copy_dst = b->ptr + ((uint64_t)pc.dy * b->width + pc.dx) * 4u;
        s.dst_pitch = (uint64_t)b->width * 4u;
        s.row_bytes = (uint64_t)pc.cw * 4u;
        s.rows = pc.ch;
        break;
    }
    case CLEAR_BUFFER:
        s.copy_dst = ((WgcBuffer_T *)pc.copy_dst)->ptr + pc.dst_off;
        s.row_bytes = pc.copy_bytes;
        break;
    default:   // CLEAR_TEXTURE
        s.copy_dst = ((WgcTexture_T *)pc.copy_dst)->ptr;
        s.row_bytes = pc.copy_bytes;
        break;
    }
}

static int32_t build_plan(WgcDevice_T *dev, WgcCommandBuffer_T *cb, Plan &plan, bool noalloc)
{
    plan.reset();
    plan.noalloc = noalloc;
    plan.hip_device = dev->hip_device;
    uint32_t cfg = dev->raster_cfg;
    uint32_t shift = raster_tile_shift(cfg);
    // 1. draws per pass, in order
    struct PassWork {
        std::vector<DrawWork> draws;
        bool clear_only = false;
        Step clear;
        bool copy = false;
        Step copy_step;
        std::vector<const void *> writes, reads;
    };
    std::vector<PassWork> passes(cb->passes.size());
    for (size_t i = 0; i < cb->passes.size(); ++i) {
        if (cb->passes[i].copy_kind != COPY_NONE) {
            passes[i].copy = true;
            copy_step(cb->passes[i], passes[i].copy_step);
            continue;
        }
        int32_t r = collect_pass(dev, cb->passes[i], shift, passes[i].draws, passes[i].clear_only, passes[i].clear);
        if (r != WGC_OK) return r;
        pass_textures(cb->passes[i], passes[i].writes, passes[i].reads);
    }
    // 2. steps: runs of independent single-draw passes become launch groups
    //    (one per program and resolve kind) of one stage; everything else
    //    runs in order, one stage per step
    size_t i = 0;
    uint32_t stage = 0;
    while (i < passes.size()) {
        PassWork &pw = passes[i];
        auto mergeable = [](const PassWork &q) {
            return q.draws.size() == 1 && !is_point_or_line(q.draws[0].d) && !restart_strip(q.draws[0].d);
        };
        // the allocation-free form has no draw tables, so no batched groups
        bool batchable = dev->batching && !noalloc && mergeable(pw);
        if (!batchable) {
            if (pw.copy) {
                pw.copy_step.stage = stage++;
                plan.steps.push_back(pw.copy_step);
            }
            for (auto &dw : pw.draws) {
                if (!is_point_or_line(dw.d)) {
                    add_group(plan.steps, {dw.d}, cfg, stage++);
                    continue;
                }
                // points and lines: the pass's clear (if still pending) first
                if (dw.d.init_clear) {
                    Step c;
                    c.kind = StepKind::Clear;
                    c.color = dw.d.clear_color_on ? (uint32_t *)dw.d.color : nullptr;
                    c.rgba = dw.d.clear_rgba;
                    c.depth = dw.d.clear_depth_on ? dw.d.depth : nullptr;
                    c.dv = dw.d.clear_depth;
                    c.n = (uint64_t)dw.d.width * dw.d.height;
                    c.stage = stage++;
                    plan.steps.push_back(c);
                    dw.d.init_clear = 0;
                }
                Step l;
                l.kind = StepKind::Lines;
                l.draws = {dw.d};
                l.gl.program = dw.d.program;
                l.gl.ordered = dw.d.resolve == RES_ORDERED;
                l.gl.user = (const UserKernels *)dw.d.user;
                l.stage = stage++;
                plan.steps.push_back(std::move(l));
            }
            if (pw.clear_only) {
                pw.clear.stage = stage++;
                plan.steps.push_back(pw.clear);
            }
            ++i;
            continue;
        }
        std::vector<const void *> written, read;
        size_t j = i;
        while (j < passes.size() && mergeable(passes[j])) {
            PassWork &q = passes[j];
            bool clash = false;
            for (auto *t : q.writes) {
                for (auto *u : written) clash = clash || t == u;
                for (auto *u : read) clash = clash || t == u;
            }
            for (auto *t : q.reads)
                for (auto *u : written) clash = clash || t == u;
            if (clash) break;
            written.insert(written.end(), q.writes.begin(), q.writes.end());
            read.insert(read.end(), q.reads.begin(), q.reads.end());
            ++j;
        }
        if (j - i == 1) {
            add_group(plan.steps, {pw.draws[0].d}, cfg, stage++);
        } else {
            // group by (program, ordered); ordered groups use the 16x16 raster
            std::map<std::tuple<uint32_t, bool, const void *>, std::vector<DrawParams>> groups;
            for (size_t k = i; k < j; ++k) {
                const DrawParams &d = passes[k].draws[0].d;
                groups[{d.program, d.resolve == RES_ORDERED, d.user}].push_back(d);
            }
            // the groups of one run of independent passes share a stage
            for (auto &kv : groups) add_group(plan.steps, std::move(kv.second), cfg, stage);
            ++stage;
        }
        i = j;
    }
    // 3. scratch: the groups of a stage side by side, stages reusing the
    //    arena (they run one after another)
    size_t need = 0, zneed = 0, kneed = 0;
    size_t tables = 0;
    size_t stage_off = 0;
    uint32_t cur_stage = UINT32_MAX;
    for (Step &s : plan.steps) {
        if (s.kind != StepKind::Group && s.kind != StepKind::Lines) continue;
        if (s.stage != cur_stage) {
            cur_stage = s.stage;
            stage_off = 0;
        }
        s.arena_off = stage_off;
        s.zone_off = zneed;
        s.keep_off = kneed;
        size_t bytes = s.kind == StepKind::Lines ? layout_lines(s, nullptr, nullptr, 0, dev)
                                                  : layout_group(s, nullptr, nullptr, nullptr, nullptr, 0, dev);
        stage_off = align_up(stage_off + bytes, 256);
        need = std::max(need, stage_off);
        if (s.kind == StepKind::Group) {
            // active-list entries hold the global tile in kItemShift bits
            if (group_tiles(s) >= (1ull << kItemShift))
                return set_error(WGC_ERR_INVALID, "too many tiles in one launch group");
            zneed += zone_bytes(s);
            kneed += keep_bytes(s);
        }
        if (s.multi) {
            // draws, primitive and tile prefixes, the draw of every wave of
            // primitives and of every tile
            s.table_off = tables;
            size_t n = s.draws.size(), waves = 0, tiles = 0;
            for (const DrawParams &d : s.draws) {
                waves += align_up(d.total_prims, 64) / 64;
                tiles += (size_t)d.tiles_x * d.tiles_y;
            }
            tables = align_up(tables + n * sizeof(DrawParams) + 8 * (n + 1) + 4 * (waves + tiles), 256);
        }
    }
    uint8_t *zone_base, *keep_base, *arena_base;
    if (noalloc) {
        // the static scratch: zero region, kept-zero region, then the stages
        size_t zone = align_up(zneed, 256), keep = align_up(kneed, 256);
        if (tables) return set_error(WGC_ERR_INVALID, "batched launch group in an allocation-free plan");
        if (zone + keep + need > kScratchBytes)
            return set_error(WGC_ERR_OUT_OF_MEMORY,
                             "the command buffer needs " + std::to_string(zone + keep + need) +
                                 " bytes of scratch; the allocation-free submission has " +
                                 std::to_string(kScratchBytes));
        if (!dev->scratch) return set_error(WGC_ERR_DEVICE, "no scratch address for this device");
        zone_base = dev->scratch;
        keep_base = dev->scratch + zone;
        arena_base = dev->scratch + zone + keep;
    } else if (zneed > dev->zone_size || kneed > dev->keep_size ||
        dev->zone_size + dev->keep_size + need > dev->arena_size) {
        HIP_TRY(hipStreamSynchronize(dev->stream), "arena");
        if (dev->arena) HIP_TRY(hipFree(dev->arena), "arena");
        dev->arena = nullptr;
        dev->arena_size = 0;
        size_t zone = align_up(std::max(zneed, dev->zone_size), 4096);
        size_t keep = align_up(std::max(kneed, dev->keep_size), 4096);
        size_t grow = zone + keep + std::max(need, (size_t)64 << 20);
        if (hipMalloc(&dev->arena, grow) != hipSuccess) {
            dev->zone_size = dev->keep_size = 0;
            return set_error(WGC_ERR_OUT_OF_MEMORY, "scratch arena");
        }
        HIP_TRY(hipMemsetAsync(dev->arena, 0, zone + keep, dev->stream), "arena zero regions");
        dev->arena_size = grow;
        dev->zone_size = zone;
        dev->keep_size = keep;
        dev->arena_gen++;
    }
    if (!noalloc) {
        zone_base = dev->arena;
        keep_base = dev->arena + dev->zone_size;
        arena_base = dev->arena + dev->zone_size + dev->keep_size;
    }
    // draw tables: pooled device memory, uploaded from pooled pinned staging
    // without a host wait (the upload is ordered before the plan's launches)
    if (tables) {
        HIP_TRY(dev->table_pool.alloc(tables, plan.d_tables), "draw tables");
        hipError_t e = dev->staging_pool.alloc(tables, plan.h_tables);
        if (e != hipSuccess) {
            dev->table_pool.release(plan.d_tables, tables);
            plan.d_tables = nullptr;
            return hip_fail(e, "draw table staging");
        }
        plan.tables_size = tables;
        plan.dev = dev;
        retain(dev);
    }
    for (Step &s : plan.steps) {
        if (s.kind == StepKind::Lines) {
            layout_lines(s, arena_base + s.arena_off, dev->frag_px, dev->frag_px_size, dev);
            plan.launches += s.draws[0].total_prims == 0 ? 0 : (s.gl.ordered ? 2 : 3);
            plan.prims += s.draws[0].total_prims;
            continue;
        }
        if (s.kind != StepKind::Group) {
            plan.launches += 1;
            continue;
        }
        layout_group(s, arena_base + s.arena_off, zone_base + s.zone_off, keep_base + s.keep_off, dev->frag_px,
                     dev->frag_px_size, dev);
        plan.launches += (s.gl.total_prims > 0 ? (s.gl.inline_clip ? 2 : 3) : 1) + (!s.multi && restart_strip(s.draws[0]) ? 2 : 0);
        for (auto &d : s.draws) plan.prims += d.total_prims;
        if (!s.multi) continue;
        uint32_t n = (uint32_t)s.draws.size();
        uint8_t *h = plan.h_tables + s.table_off;
        uint8_t *dptr = plan.d_tables + s.table_off;
        memcpy(h, s.draws.data(), n * sizeof(DrawParams));
        uint32_t *pb = (uint32_t *)(h + n * sizeof(DrawParams));
        uint32_t *tb = pb + (n + 1);
        uint32_t acc_p = 0, acc_t = 0;
        for (uint32_t k = 0; k < n; ++k) {
            pb[k] = acc_p;
            tb[k] = acc_t;
            acc_p += (uint32_t)align_up(s.draws[k].total_prims, 64);
            acc_t += s.draws[k].tiles_x * s.draws[k].tiles_y;
        }
        pb[n] = acc_p;
        tb[n] = acc_t;
        uint32_t *wd = tb + (n + 1), *td = wd + acc_p / 64;
        for (uint32_t k = 0; k < n; ++k) {
            for (uint32_t w = pb[k] / 64; w < pb[k + 1] / 64; ++w) wd[w] = k;
            for (uint32_t t = tb[k]; t < tb[k + 1]; ++t) td[t] = k;
        }
        s.batch.draws = (const DrawParams *)dptr;
        s.batch.prim_base = (const uint32_t *)(dptr + n * sizeof(DrawParams));
        s.batch.tile_base = s.batch.prim_base + (n + 1);
        s.batch.wave_draw = s.batch.tile_base + (n + 1);
        s.batch.tile_draw = s.batch.wave_draw + acc_p / 64;
        s.batch.n = n;
        s.batch.total_prims = acc_p;
        s.batch.total_tiles = acc_t;
    }
    if (tables)
        HIP_TRY(hipMemcpyAsync(plan.d_tables, plan.h_tables, tables, hipMemcpyHostToDevice, dev->stream),
                "draw tables");
    plan.arena_gen = dev->arena_gen;
    plan.zone_used = zneed;
    plan.keep_used = kneed;
    plan.raster_cfg = cfg;
    plan.batching = dev->batching;
    return WGC_OK;
}

static hipError_t launch_step(WgcDevice_T *dev, const Step &s, hipStream_t st, KernelTimer *kt)
{
    if (s.kind == StepKind::Clear) return launch_clear(s.color, s.rgba, s.depth, s.dv, s.n, st, kt);
    if (s.kind == StepKind::Copy) {
        if (!s.copy_src) return launch_fill(s.copy_dst, s.row_bytes, 0u, st, kt);
        return launch_copy_rows(s.copy_src, s.src_pitch, s.copy_dst, s.dst_pitch, s.row_bytes, s.rows, st, kt);
    }
    if (s.kind == StepKind::Lines) return launch_lines(s.draws[0], s.gl, st, kt);
    if (s.gl.user) return launch_user(*s.gl.user, s.draws[0], s.batch, s.gl, s.multi, st, kt);
    if (s.multi) return launch_batch(s.draws[0], s.batch, s.gl, st, kt);
    return launch_draw(s.draws[0], s.batch, s.gl, st, kt);
}

// Steps run in order on the device stream, except that the steps of one
// stage (independent launch groups) alternate between the device stream and
// a second stream, forked from and joined back into the device stream with
// events (the same under hipGraph capture).
static hipError_t launch_plan(WgcDevice_T *dev, const Plan &plan, KernelTimer *kt)
{
    hipError_t e = hipSuccess;
    size_t i = 0;
    while (i < plan.steps.size()) {
        size_t j = i + 1;
        while (j < plan.steps.size() && plan.steps[j].stage == plan.steps[i].stage) ++j;
        if (j - i == 1) {
            e = launch_step(dev, plan.steps[i], dev->stream, kt);
            if (e != hipSuccess) return e;
            i = j;
            continue;
        }
        e = hipEventRecord(dev->fork, dev->stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(dev->side, dev->fork, 0);
        for (size_t k = i; k < j && e == hipSuccess; ++k)
            e = launch_step(dev, plan.steps[k], ((k - i) & 1u) ? dev->side : dev->stream, kt);
        if (e == hipSuccess) e = hipEventRecord(dev->join, dev->side);
        if (e == hipSuccess) e = hipStreamWaitEvent(dev->stream, dev->join, 0);
        if (e != hipSuccess) return e;
        i = j;
    }
    return hipSuccess;
}

// The first submission of a plan launches directly; a command buffer that
// is submitted again is captured into a hipGraph then and replayed.
static int32_t submit_plan(WgcDevice_T *dev, Plan &plan)
{
    KernelTimer *kt = dev->timing ? &dev->timer : nullptr;
    bool graph = dev->use_graphs && !kt && (plan.exec || plan.submits > 0);
    plan.submits++;
    if (!graph) {
        HIP_TRY(launch_plan(dev, plan, kt), "launch");
        return WGC_OK;
    }
    if (!plan.exec) {
        HIP_TRY(hipStreamBeginCapture(dev->stream, hipStreamCaptureModeThreadLocal), "graph capture");
        hipError_t e = launch_plan(dev, plan, nullptr);
        hipGraph_t g = nullptr;
        hipError_t e2 = hipStreamEndCapture(dev->stream, &g);
        if (e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return hip_fail(e, "graph capture launch");
        }
        if (e2 != hipSuccess) return hip_fail(e2, "graph capture");
        plan.graph = g;
        if (!plan.dev) {   // its graph is destroyed through the device's deferred list
            plan.dev = dev;
            retain(dev);
        }
        HIP_TRY(hipGraphInstantiate(&plan.exec, g, nullptr, nullptr, 0), "graph instantiate");
    }
    HIP_TRY(hipGraphLaunch(plan.exec, dev->stream), "graph launch");
    return WGC_OK;
}

static bool is_mapped(const WgcBuffer_T *b) { return b && (b->map.mode || b->map.pending); }

// wgpu validation: a submission must not use a mapped buffer.
static bool uses_mapped_buffer(const WgcCommandBuffer_T *cb)
{
    for (const PassCommand &pc : cb->passes) {
        if (pc.copy_kind == COPY_B2B || pc.copy_kind == COPY_B2T) {
            if (is_mapped((const WgcBuffer_T *)pc.copy_src)) return true;
        }
        if (pc.copy_kind == COPY_B2B || pc.copy_kind == COPY_T2B || pc.copy_kind == CLEAR_BUFFER) {
            if (is_mapped((const WgcBuffer_T *)pc.copy_dst)) return true;
        }
        for (const SubCommand &c : pc.cmds) {
            if (c.kind == CmdKind::SetVertexBuffer || c.kind == CmdKind::SetIndexBuffer) {
                if (is_mapped((const WgcBuffer_T *)c.obj)) return true;
            } else if (c.kind == CmdKind::SetBindGroup && c.obj) {
                for (const BindingSlot &e : ((const WgcBindGroup_T *)c.obj)->entries)
                    if (is_mapped(e.buffer)) return true;
            }
        }
    }
    return false;
}

// A free pinned counter slot and completion event for a submission.  The
// allocation-free form never creates either: with none free it waits for
// the oldest submission in flight to retire and give its own back.
static int32_t take_slot_and_event(WgcDevice_T *d, bool noalloc, HostCounters *&slot, hipEvent_t &done)
{
    if (noalloc) {
        while ((d->slot_pool.empty() || d->event_pool.empty()) && !d->inflight.empty()) {
            int32_t r = retire(d, d->inflight.front().index, true, 0);
            if (r != WGC_OK) return r;
        }
        if (d->slot_pool.empty() || d->event_pool.empty())
            return set_error(WGC_ERR_OUT_OF_MEMORY, "no free counter slot or event");
    } else if (d->event_pool.empty()) {
        hipEvent_t e = nullptr;
        HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
        d->event_pool.push_back(e);
    }
    HIP_TRY(take_slot(d, slot), "pinned counter slot");
    done = d->event_pool.back();
    d->event_pool.pop_back();
    return WGC_OK;
}

static int32_t submit_plans(WgcDevice_T *device, uint32_t count, const WgcCommandBuffer *buffers, bool noalloc,
                            Inflight &f);

// Both submission forms.  `noalloc`: plans laid out in the static scratch,
// launched directly (no graph), no device or pinned allocation.
static int32_t submit(WgcDevice_T *device, uint32_t count, const WgcCommandBuffer *buffers,
                      uint64_t *submission_index, bool noalloc)
{
    if (!device || (count && !buffers)) return set_error(WGC_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(device->mu);
    HIP_TRY(hipSetDevice(device->hip_device), "hipSetDevice");
    for (uint32_t i = 0; i < count; ++i) {
        if (!buffers[i]) return set_error(WGC_ERR_INVALID, "null command buffer");
        if (uses_mapped_buffer(buffers[i]))
            return set_error(WGC_ERR_INVALID, "a buffer used by the command buffer is mapped");
    }
    Inflight f;
    f.prims = f.launches = 0;
    int32_t r0 = take_slot_and_event(device, noalloc, f.slot, f.done);
    if (r0 != WGC_OK) return r0;
    auto give_back = [&]() {
        device->slot_pool.push_back(f.slot);
        device->event_pool.push_back(f.done);
    };
    // the counters are zero here: every submission takes (and zeroes) them
    // at its end, or is followed by this reset when it stopped before that
    if (device->counters_dirty) {
        // a submission that stopped early may also have left launch-group
        // words in the arena's zero region
        hipError_t e0 = launch_reset(device->d_frags, device->d_error, device->stream);
        if (e0 == hipSuccess && device->arena && device->zone_size + device->keep_size)
            e0 = launch_fill(device->arena, device->zone_size + device->keep_size, 0u, device->stream, nullptr);
        if (e0 != hipSuccess) {
            give_back();
            return hip_fail(e0, "counter reset");
        }
        device->counters_dirty = false;
    }
    f.index = device->next_index++;
    device->counters_dirty = true;
    int32_t r = submit_plans(device, count, buffers, noalloc, f);
    if (r != WGC_OK) {
        // the counters stay dirty; the next submission resets them.  Kernels
        // of the plans launched before the failure may still run on the
        // command buffers' resources: they finish before those are let go.
        (void)hipStreamSynchronize(device->stream);
        for (auto *o : f.keep) release(o);
        give_back();
        return r;
    }
    device->counters_dirty = false;
    if (submission_index) *submission_index = f.index;
    device->inflight.push_back(std::move(f));
    // keep the queue short: retire what is already done
    return retire(device, UINT64_MAX, false, 0);
}

// Launches the plans of a submission and its end-of-submission kernel, and
// records its completion event.
static int32_t submit_plans(WgcDevice_T *device, uint32_t count, const WgcCommandBuffer *buffers, bool noalloc,
                            Inflight &f)
{
    KernelTimer *kt = (device->timing && !noalloc) ? &device->timer : nullptr;
    size_t zone_dirty = 0;   // bytes of the zero region the plans launched so far have used
    uint8_t *zone = noalloc ? device->scratch : device->arena;
    for (uint32_t i = 0; i < count; ++i) {
        WgcCommandBuffer_T *cb = buffers[i];
        Plan &plan = noalloc ? *cb->plan_noalloc : *cb->plan;
        if (plan.arena_gen != device->arena_gen || plan.raster_cfg != device->raster_cfg ||
            (!noalloc && plan.batching != device->batching) || plan.hip_device != device->hip_device) {
            int32_t r = build_plan(device, cb, plan, noalloc);
            if (r != WGC_OK) {
                plan.reset();
                return r;
            }
            // growing the arena may have invalidated plans built before
            if (plan.arena_gen != device->arena_gen) {
                r = build_plan(device, cb, plan, noalloc);
                if (r != WGC_OK) return r;
            }
            zone = noalloc ? device->scratch : device->arena;
        }
        if (noalloc) {
            // the scratch keeps whatever the last user left in it (another
            // device object on this GPU, a submission that stopped early):
            // its zero and kept-zero regions are cleared before each plan
            size_t z = align_up(plan.zone_used, 256) + align_up(plan.keep_used, 256);
            HIP_TRY(launch_fill(device->scratch, z, 0u, device->stream, nullptr), "scratch zero regions");
            zone_dirty = 0;
        } else if (zone_dirty && plan.zone_used) {
            // the plans of one submission share the zero region
            HIP_TRY(launch_fill(device->arena, zone_dirty, 0u, device->stream, nullptr), "zero region");
            zone_dirty = 0;
        }
        if (noalloc) {
            HIP_TRY(launch_plan(device, plan, kt), "launch");
        } else {
            int32_t r = submit_plan(device, plan);
            if (r != WGC_OK) return r;
        }
        zone_dirty = std::max(zone_dirty, plan.zone_used);
        f.prims += plan.prims;
        f.launches += plan.launches;
        retain(cb);
        f.keep.push_back(cb);
    }
    // the submission's counters, read by the host once its event has passed
    // (the event is the end kernel's own completion: one packet less than a
    // separate record, about 1 us per submission)
    HIP_TRY(launch_end_submission(device->d_frags, device->d_error, f.slot, zone, zone_dirty, device->stream, f.done),
            "counter readback");
    return WGC_OK;
}

extern "C" int32_t wgc_queue_submit(WgcDevice device, uint32_t count, const WgcCommandBuffer *buffers,
                                    uint64_t *submission_index)
{
    return submit(device, count, buffers, submission_index, false);
}

extern "C" int32_t wgc_queue_submit_noalloc(WgcDevice device, uint32_t count, const WgcCommandBuffer *buffers,
                                            uint64_t *submission_index)
{
    return submit(device, count, buffers, submission_index, true);
}

extern "C" int32_t wgc_device_set_option(WgcDevice device, uint32_t option, uint32_t value)
{
    if (!device) return set_error(WGC_ERR_INVALID, "null device");
    std::lock_guard<std::mutex> lk(device->mu);
    switch (option) {
    case WGC_OPTION_RASTER_TILE:
        if (value >= RASTER_NUM_CFG) return set_error(WGC_ERR_INVALID, "bad raster tile shape");
        device->raster_cfg = value;
        return WGC_OK;
    case WGC_OPTION_BATCHING: device->batching = value != 0; return WGC_OK;
    case WGC_OPTION_GRAPHS: device->use_graphs = value != 0; return WGC_OK;
    default: return set_error(WGC_ERR_INVALID, "unknown option");
    }
}

extern "C" int32_t wgc_device_get_option(WgcDevice device, uint32_t option, uint32_t *value)
{
    if (!device || !value) return set_error(WGC_ERR_INVALID, "null argument");
    switch (option) {
    case WGC_OPTION_RASTER_TILE: *value = device->raster_cfg; return WGC_OK;
    case WGC_OPTION_BATCHING: *value = device->batching ? 1u : 0u; return WGC_OK;
    case WGC_OPTION_GRAPHS: *value = device->use_graphs ? 1u : 0u; return WGC_OK;
    default: return set_error(WGC_ERR_INVALID, "unknown option");
    }
}

// ------------------------------------------------------------------
// measurement
// ------------------------------------------------------------------

extern "C" int32_t wgc_device_timestamp(WgcDevice device, uint32_t slot)
{
    if (!device || slot >= 64) return set_error(WGC_ERR_INVALID, "bad slot");
    std::lock_guard<std::mutex> lk(device->mu);
    HIP_TRY(hipSetDevice(device->hip_device), "hipSetDevice");
    if (!device->ts[slot]) HIP_TRY(hipEventCreate(&device->ts[slot]), "hipEventCreate");
    HIP_TRY(hipEventRecord(device->ts[slot], device->stream), "hipEventRecord");
    device->ts_recorded[slot] = true;
    return WGC_OK;
}

extern "C" int32_t wgc_device_elapsed_ms(WgcDevice device, uint32_t a, uint32_t b, float *ms)
{
    if (!device || a >= 64 || b >= 64 || !ms) return set_error(WGC_ERR_INVALID, "bad slot");
    if (!device->ts_recorded[a] || !device->ts_recorded[b]) return set_error(WGC_ERR_INVALID, "slot not recorded");
    HIP_TRY(hipEventSynchronize(device->ts[b]), "hipEventSynchronize");
    HIP_TRY(hipEventElapsedTime(ms, device->ts[a], device->ts[b]), "hipEventElapsedTime");
    return WGC_OK;
}

extern "C" int32_t wgc_device_set_kernel_timing(WgcDevice device, uint32_t enabled)
{
    if (!device) return set_error(WGC_ERR_INVALID, "null device");
    std::lock_guard<std::mutex> lk(device->mu);
    device->timing = enabled != 0;
    device->ktimes.clear();
    return WGC_OK;
}

extern "C" int32_t wgc_device_kernel_time(WgcDevice device, const char *name, double *total_ms, uint64_t *launches)
{
    if (!device || !name || !total_ms || !launches) return set_error(WGC_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(device->mu);
    HIP_TRY(hipSetDevice(device->hip_device), "hipSetDevice");
    int32_t r = drain(device);
    if (r != WGC_OK) return r;
    r = resolve_timers(device);
    if (r != WGC_OK) return r;
    *total_ms = 0.0;
    *launches = 0;
    for (auto &kv : device->ktimes)
        if (kv.first.find(name) != std::string::npos) {
            *total_ms += kv.second.first;
            *launches += kv.second.second;
        }
    return WGC_OK;
}
