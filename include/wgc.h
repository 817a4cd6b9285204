/*
 * wgc.h -- C ABI of libwgc, an MI355X (gfx950) render-pass executor that
 * drops in where wgpu-cpu's CPU executor sits.
 *
 * The reference boundary is the set of `wgpu::custom::*Interface` trait
 * objects that wgpu-cpu implements; each entry point below names the trait
 * method it replaces (paths relative to the jgraef/wgpu-cpu checkout).  A
 * Rust `wgpu::custom` adapter (or any FFI) binds these functions one to one.
 *
 * Conventions
 *  - Opaque handles.  Every object is reference counted like the reference's
 *    Arc clones: recorded commands keep the resources they use alive until
 *    the submission that uses them has executed.  `*_release` drops the
 *    caller's reference.
 *  - Return codes: WGC_OK (0) or a negative WgcStatus.  Where the reference
 *    panics (unwrap / expect / todo!), libwgc returns an error and keeps a
 *    thread-local message readable with wgc_last_error().
 *  - One HIP stream per device executes submissions in order, as the
 *    reference's single engine thread does (wgpu-cpu/src/engine.rs:26-46).
 *    Queue writes are ordered with respect to submissions.
 *  - Semantics: WGC_SEMANTICS_REFERENCE reproduces the reference bit for bit
 *    (integer scanline coverage, corner sampling, nearest texel rounding,
 *    no sRGB decode, no blending).  WGC_SEMANTICS_EXTENDED additionally
 *    honours what the reference stores but ignores or leaves todo!():
 *    perspective-correct varyings, linear filtering, sRGB decode and
 *    src-over blending.
 */
#ifndef WGC_H
#define WGC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WGC_ABI_VERSION 1u

typedef enum WgcStatus {
    WGC_OK = 0,
    WGC_ERR_INVALID = -1,      /* invalid argument or state (reference: panic) */
    WGC_ERR_UNSUPPORTED = -2,  /* the reference has todo!() here */
    WGC_ERR_OUT_OF_MEMORY = -3,
    WGC_ERR_DEVICE = -4,       /* HIP runtime error */
    WGC_ERR_TIMEOUT = -5,      /* poll timeout (reference: PollError::Timeout) */
    WGC_ERR_EXECUTION = -6,    /* a submitted pass failed on the device */
} WgcStatus;

typedef struct WgcDevice_T *WgcDevice;
typedef struct WgcBuffer_T *WgcBuffer;
typedef struct WgcTexture_T *WgcTexture;
typedef struct WgcSampler_T *WgcSampler;
typedef struct WgcBindGroup_T *WgcBindGroup;
typedef struct WgcShaderModule_T *WgcShaderModule;
typedef struct WgcRenderPipeline_T *WgcRenderPipeline;
typedef struct WgcCommandEncoder_T *WgcCommandEncoder;
typedef struct WgcRenderPass_T *WgcRenderPass;
typedef struct WgcCommandBuffer_T *WgcCommandBuffer;
typedef struct WgcHeap_T *WgcHeap;

/* ---- enums (numbering follows wgpu-types where it has one) ---- */

typedef enum WgcSemantics {
    WGC_SEMANTICS_REFERENCE = 0,
    WGC_SEMANTICS_EXTENDED = 1,
} WgcSemantics;

typedef enum WgcTextureFormat {
    WGC_TEXTURE_FORMAT_RGBA8_UNORM = 0,
    WGC_TEXTURE_FORMAT_RGBA8_UNORM_SRGB = 1,
    WGC_TEXTURE_FORMAT_BGRA8_UNORM = 2,
    WGC_TEXTURE_FORMAT_BGRA8_UNORM_SRGB = 3,
    WGC_TEXTURE_FORMAT_DEPTH32_FLOAT = 4,
} WgcTextureFormat;

typedef enum WgcVertexFormat {
    WGC_VERTEX_FORMAT_FLOAT32 = 1,
    WGC_VERTEX_FORMAT_FLOAT32X2 = 2,
    WGC_VERTEX_FORMAT_FLOAT32X3 = 3,
    WGC_VERTEX_FORMAT_FLOAT32X4 = 4,
} WgcVertexFormat;

typedef enum WgcVertexStepMode { WGC_STEP_VERTEX = 0, WGC_STEP_INSTANCE = 1 } WgcVertexStepMode;
typedef enum WgcIndexFormat { WGC_INDEX_UINT16 = 2, WGC_INDEX_UINT32 = 4 } WgcIndexFormat;
/* wgpu::PrimitiveTopology (state.rs:434-455) */
typedef enum WgcTopology {
    WGC_TOPOLOGY_POINT_LIST = 0,
    WGC_TOPOLOGY_LINE_LIST = 1,
    WGC_TOPOLOGY_LINE_STRIP = 2,
    WGC_TOPOLOGY_TRIANGLE_LIST = 3,
    WGC_TOPOLOGY_TRIANGLE_STRIP = 4
} WgcTopology;
typedef enum WgcFrontFace { WGC_FRONT_FACE_CCW = 0, WGC_FRONT_FACE_CW = 1 } WgcFrontFace;
typedef enum WgcCullMode { WGC_CULL_NONE = 0, WGC_CULL_FRONT = 1, WGC_CULL_BACK = 2 } WgcCullMode;
typedef enum WgcCompare {
    WGC_COMPARE_NEVER = 1, WGC_COMPARE_LESS = 2, WGC_COMPARE_EQUAL = 3,
    WGC_COMPARE_LESS_EQUAL = 4, WGC_COMPARE_GREATER = 5, WGC_COMPARE_NOT_EQUAL = 6,
    WGC_COMPARE_GREATER_EQUAL = 7, WGC_COMPARE_ALWAYS = 8,
} WgcCompare;
typedef enum WgcAddressMode {
    WGC_ADDRESS_CLAMP_TO_EDGE = 0, WGC_ADDRESS_REPEAT = 1, WGC_ADDRESS_MIRROR_REPEAT = 2,
} WgcAddressMode;
typedef enum WgcFilterMode { WGC_FILTER_NEAREST = 0, WGC_FILTER_LINEAR = 1 } WgcFilterMode;
typedef enum WgcLoadOp { WGC_LOAD_OP_LOAD = 0, WGC_LOAD_OP_CLEAR = 1 } WgcLoadOp;
typedef enum WgcBlend { WGC_BLEND_REPLACE = 0, WGC_BLEND_ALPHA = 1 } WgcBlend;
typedef enum WgcInterpolation {
    WGC_INTERPOLATE_LINEAR = 0, WGC_INTERPOLATE_PERSPECTIVE = 1, WGC_INTERPOLATE_FLAT = 2,
} WgcInterpolation;

/* Built-in WGSL programs (their WGSL text ships in the package's shaders/). */
typedef enum WgcBuiltinProgram {
    WGC_PROGRAM_MVP_COLOR = 1,        /* camera.matrix * position, colour varying */
    WGC_PROGRAM_MVP_TEXTURE = 2,      /* camera.matrix * position, uv, textureSample */
    WGC_PROGRAM_COLORED_TRIANGLE = 3, /* vertex_index only (wgpu-cpu-tests colored_triangle.wgsl) */
    WGC_PROGRAM_CLIP_COLOR = 4,       /* clip-space position and colour pass-through */
} WgcBuiltinProgram;

/* ---- descriptors ---- */

typedef struct WgcDeviceDescriptor {
    int32_t hip_device;      /* HIP ordinal */
    uint32_t semantics;      /* WgcSemantics */
} WgcDeviceDescriptor;

typedef struct WgcSamplerDescriptor {
    uint32_t address_mode_u, address_mode_v;  /* WgcAddressMode */
    uint32_t mag_filter, min_filter;          /* WgcFilterMode */
} WgcSamplerDescriptor;

typedef enum WgcBindingKind { WGC_BINDING_BUFFER = 1, WGC_BINDING_TEXTURE = 2, WGC_BINDING_SAMPLER = 3 } WgcBindingKind;

typedef struct WgcBindGroupEntry {
    uint32_t binding;
    uint32_t kind;           /* WgcBindingKind */
    WgcBuffer buffer;
    uint64_t offset, size;   /* size 0 = to the end */
    WgcTexture texture;
    WgcSampler sampler;
} WgcBindGroupEntry;

typedef struct WgcVertexAttribute {
    uint32_t format;         /* WgcVertexFormat */
    uint32_t offset;
    uint32_t shader_location;
} WgcVertexAttribute;

typedef struct WgcVertexBufferLayout {
    uint32_t array_stride;
    uint32_t step_mode;      /* WgcVertexStepMode */
    uint32_t attribute_count;
    const WgcVertexAttribute *attributes;
} WgcVertexBufferLayout;

typedef struct WgcRenderPipelineDescriptor {
    WgcShaderModule vertex_module;
    uint32_t vertex_buffer_count;
    const WgcVertexBufferLayout *vertex_buffers;
    uint32_t topology;       /* WgcTopology */
    uint32_t front_face;     /* WgcFrontFace */
    uint32_t cull_mode;      /* WgcCullMode */
    uint32_t has_depth_stencil;
    uint32_t depth_format;   /* WGC_TEXTURE_FORMAT_DEPTH32_FLOAT */
    uint32_t depth_write_enabled;
    uint32_t depth_compare;  /* WgcCompare */
    WgcShaderModule fragment_module; /* NULL: no fragment stage */
    uint32_t target_format;  /* WgcTextureFormat */
    uint32_t blend;          /* WgcBlend */
    /* Entry points of WGSL modules (wgpu::VertexState::entry_point,
     * FragmentState::entry_point); NULL: the module's only entry point of
     * that stage (state.rs via CompiledModule::find_entry_point). */
    const char *vertex_entry;
    const char *fragment_entry;
    /* PrimitiveState::strip_index_format: 0 none, else the WgcIndexFormat
     * whose maximum value restarts strips (index.rs:72-106) */
    uint32_t strip_index_format;
} WgcRenderPipelineDescriptor;

typedef struct WgcColorAttachment {
    WgcTexture view;
    uint32_t load_op;        /* WgcLoadOp */
    double clear_value[4];   /* wgpu::Color (f64) */
} WgcColorAttachment;

typedef struct WgcDepthStencilAttachment {
    WgcTexture view;
    uint32_t has_depth_ops;
    uint32_t depth_load_op;  /* WgcLoadOp */
    float depth_clear_value;
} WgcDepthStencilAttachment;

typedef struct WgcRenderPassDescriptor {
    uint32_t color_attachment_count;  /* 0 or 1 */
    const WgcColorAttachment *color_attachments;
    const WgcDepthStencilAttachment *depth_stencil_attachment; /* may be NULL */
} WgcRenderPassDescriptor;

typedef struct WgcPassStats {
    uint64_t fragments;      /* emitted by the scanline rule inside the scissor */
    uint64_t primitives;     /* assembled, before culling */
    uint64_t kernel_launches;
} WgcPassStats;

/* ---- library ---- */

uint32_t wgc_abi_version(void);
/* Thread-local message of the last failing call on this thread. */
const char *wgc_last_error(void);

/* ---- instance / adapter / device (wgpu-cpu/src/lib.rs:25-27,
 *      instance.rs:72-102, adapter.rs:24-42, device.rs:40-74) ---- */
int32_t wgc_device_create(const WgcDeviceDescriptor *desc, WgcDevice *out);
void wgc_device_release(WgcDevice device);

/* ---- buffers (DeviceInterface::create_buffer device.rs:157-163,
 *      QueueInterface::write_buffer device.rs:332-344) ---- */
int32_t wgc_buffer_create(WgcDevice device, uint64_t size, WgcBuffer *out);
int32_t wgc_queue_write_buffer(WgcDevice device, WgcBuffer buffer, uint64_t offset,
                               const void *data, uint64_t size);
/* Readback after the queue drained (BufferInterface map_async +
 * get_mapped_range, buffer.rs:112-167). */
int32_t wgc_buffer_read(WgcBuffer buffer, uint64_t offset, void *dst, uint64_t size);
void wgc_buffer_release(WgcBuffer buffer);

/* Mapping (BufferInterface::map_async / get_mapped_range / unmap,
 * buffer.rs:112-167).  map_async copies [offset, offset + size) into host
 * memory after the work submitted so far; the callback runs from
 * wgc_device_poll once the range is mapped (status WGC_OK), or with an error
 * status if the buffer is unmapped or released first.  A write mapping is
 * copied back to the buffer by wgc_buffer_unmap.  Mapping a mapped buffer is
 * WGC_ERR_INVALID (reference: panic "Buffer already mapped"); offset must be
 * a multiple of 8 and size of 4 (wgpu's MAP_ALIGNMENT / COPY_BUFFER_ALIGNMENT). */
typedef enum WgcMapMode { WGC_MAP_READ = 1, WGC_MAP_WRITE = 2 } WgcMapMode;
typedef void (*WgcBufferMapCallback)(int32_t status, void *userdata);
int32_t wgc_buffer_map_async(WgcBuffer buffer, uint32_t mode, uint64_t offset, uint64_t size,
                             WgcBufferMapCallback callback, void *userdata);
/* Host pointer to [offset, offset + size) of the mapped range. */
int32_t wgc_buffer_get_mapped_range(WgcBuffer buffer, uint64_t offset, uint64_t size, void **data);
int32_t wgc_buffer_unmap(WgcBuffer buffer);

/* ---- textures (DeviceInterface::create_texture device.rs:165-176,
 *      QueueInterface::write_texture device.rs:371-434) ---- */
int32_t wgc_texture_create(WgcDevice device, uint32_t width, uint32_t height,
                           uint32_t format, WgcTexture *out);
/* Copies rows of bytes_per_row bytes one after another from data + offset,
 * as the reference does (device.rs:388-434). */
int32_t wgc_queue_write_texture(WgcDevice device, WgcTexture texture, const void *data,
                                uint64_t data_size, uint64_t offset, uint32_t bytes_per_row);
/* The readback wgpu_cpu::image::rgba_texture_image provides (lib.rs:158-173):
 * waits for the queue, then copies the whole texture. */
int32_t wgc_texture_read(WgcTexture texture, void *dst, uint64_t size);
void wgc_texture_release(WgcTexture texture);

/* ---- heaps (no reference counterpart: the reference allocates every
 *      resource separately on the host, buffer.rs:24-52, texture.rs:35-50).
 *      For callers that create and drop resources at a high rate (a serving
 *      loop's per-request targets): one device allocation up front, resources
 *      sub-allocated from it (first fit, 256-byte granules), zero-filled on
 *      the device stream without a host wait, and returned to the heap when
 *      released (after every submission using them has executed). A full heap
 *      returns WGC_ERR_OUT_OF_MEMORY, which a caller can use for admission
 *      control. ---- */
int32_t wgc_heap_create(WgcDevice device, uint64_t size, WgcHeap *out);
void wgc_heap_release(WgcHeap heap);
int32_t wgc_heap_stats(WgcHeap heap, uint64_t *used, uint64_t *capacity, uint64_t *largest_free);
int32_t wgc_buffer_create_in_heap(WgcHeap heap, uint64_t size, WgcBuffer *out);
int32_t wgc_texture_create_in_heap(WgcHeap heap, uint32_t width, uint32_t height, uint32_t format,
                                   WgcTexture *out);

/* ---- samplers, bind groups (device.rs:114-127, sampler.rs:3-29,
 *      bind_group.rs:16-117) ---- */
int32_t wgc_sampler_create(WgcDevice device, const WgcSamplerDescriptor *desc, WgcSampler *out);
void wgc_sampler_release(WgcSampler sampler);
int32_t wgc_bind_group_create(WgcDevice device, uint32_t entry_count,
                              const WgcBindGroupEntry *entries, WgcBindGroup *out);
void wgc_bind_group_release(WgcBindGroup group);

/* ---- shaders and pipelines (device.rs:91-99,129-134, shader.rs:31-74,
 *      pipeline.rs:57-76) ---- */
/* A built-in WGSL program; `interpolation` is the @interpolate attribute of
 * its user varying. */
int32_t wgc_shader_module_create_builtin(WgcDevice device, uint32_t program,
                                         uint32_t interpolation, WgcShaderModule *out);
void wgc_shader_module_release(WgcShaderModule module);

/* A shader module from WGSL source (wgpu::Device::create_shader_module with
 * ShaderSource::Wgsl; the reference's ShaderModule::new, shader.rs, runs
 * naga's front end and validator).  The module is translated and checked
 * here; pipelines compile their kernels from it.  Constructs the reference's
 * shader compiler leaves as todo!() return WGC_ERR_UNSUPPORTED with the
 * construct named in wgc_last_error(); invalid WGSL returns
 * WGC_ERR_INVALID with line:column. */
int32_t wgc_shader_module_create_wgsl(WgcDevice device, const char *source, size_t length,
                                      WgcShaderModule *out_module);

/* Runs one entry point of a WGSL module for `count` invocations on the
 * device and returns their outputs (naga-cranelift EntryPoint::run,
 * product.rs:282-300, the harness of its tests.rs).  Invocation i has
 * vertex_index first_index + i (vertex) or position (first_index + i, 0, 0,
 * 1) (fragment); location inputs read as zero; resources come from
 * `groups[g]` for @group(g).  `results` receives WGC_RUN_WORDS words per
 * invocation: the abort code (0, or WGC_ABORT_*), the clip position
 * (vertex) or frag depth (fragment), then 16 locations x 4 output words. */
#define WGC_RUN_WORDS 69
enum { WGC_ABORT_KILL = 3, WGC_ABORT_OUT_OF_BOUNDS = 4, WGC_ABORT_DIVISION_BY_ZERO = 5, WGC_ABORT_OVERFLOW = 6 };
int32_t wgc_shader_module_run(WgcDevice device, WgcShaderModule module, const char *entry_point,
                              uint32_t first_index, uint32_t count, const WgcBindGroup *groups,
                              uint32_t group_count, uint32_t *results, size_t results_words);

/* Translation only, without a device (for tools and tests): what = 0 the
 * emitted C++ functions, what = 1 the module's interface as JSON.  Writes at
 * most `capacity` bytes (NUL-terminated) and the full size + 1 to *needed.
 * No counterpart in the reference (its compiler output is machine code). */
int32_t wgc_wgsl_translate(const char *source, size_t length, uint32_t what, char *out, size_t capacity,
                           size_t *needed);

/* Compiles the kernels of a pipeline made of the module's vertex and (if it
 * has one) fragment entry point for gfx950, without a device: an offline
 * check of the stage interface and of the generated code.  `semantics` is a
 * WgcSemantics.  No counterpart in the reference. */
int32_t wgc_wgsl_check_pipeline(const char *source, size_t length, const char *vertex_entry,
                                const char *fragment_entry, uint32_t semantics);
int32_t wgc_render_pipeline_create(WgcDevice device, const WgcRenderPipelineDescriptor *desc,
                                   WgcRenderPipeline *out);
void wgc_render_pipeline_release(WgcRenderPipeline pipeline);

/* ---- command recording (command.rs:82-98, render_pass/mod.rs:53-329) ---- */
int32_t wgc_command_encoder_create(WgcDevice device, WgcCommandEncoder *out);
int32_t wgc_command_encoder_begin_render_pass(WgcCommandEncoder encoder,
                                              const WgcRenderPassDescriptor *desc,
                                              WgcRenderPass *out);
int32_t wgc_render_pass_set_pipeline(WgcRenderPass pass, WgcRenderPipeline pipeline);
int32_t wgc_render_pass_set_bind_group(WgcRenderPass pass, uint32_t index, WgcBindGroup group);
int32_t wgc_render_pass_set_vertex_buffer(WgcRenderPass pass, uint32_t slot, WgcBuffer buffer,
                                          uint64_t offset, uint64_t size);
int32_t wgc_render_pass_set_index_buffer(WgcRenderPass pass, WgcBuffer buffer, uint32_t format,
                                         uint64_t offset, uint64_t size);
int32_t wgc_render_pass_set_viewport(WgcRenderPass pass, float x, float y, float width,
                                     float height, float min_depth, float max_depth);
int32_t wgc_render_pass_set_scissor_rect(WgcRenderPass pass, uint32_t x, uint32_t y,
                                         uint32_t width, uint32_t height);
int32_t wgc_render_pass_set_blend_constant(WgcRenderPass pass, const double color[4]);
int32_t wgc_render_pass_set_stencil_reference(WgcRenderPass pass, uint32_t reference);
int32_t wgc_render_pass_draw(WgcRenderPass pass, uint32_t first_vertex, uint32_t vertex_count,
                             uint32_t first_instance, uint32_t instance_count);
int32_t wgc_render_pass_draw_indexed(WgcRenderPass pass, uint32_t first_index,
                                     uint32_t index_count, int32_t base_vertex,
                                     uint32_t first_instance, uint32_t instance_count);
/* RenderPassInterface::end (render_pass/mod.rs:309-321); releases `pass`. */
int32_t wgc_render_pass_end(WgcRenderPass pass);
/* Copies and clears recorded between passes, executed in order with them
 * (CommandEncoderInterface::copy_* and clear_*, command.rs:37-73 and
 * 102-116, all todo!() in the reference; wgpu's validation rules apply:
 * 4-byte buffer offsets and sizes, bytes_per_row a multiple of 256 for
 * buffer <-> texture copies, regions inside the resources, no copy within
 * one resource).  Textures are 2D with 4-byte texels; texel origins and the
 * extent are in texels. */
typedef struct WgcTexelCopyBuffer {
    WgcBuffer buffer;
    uint64_t offset;
    uint32_t bytes_per_row;   /* may be 0 when height == 1 */
} WgcTexelCopyBuffer;
typedef struct WgcTexelCopyTexture {
    WgcTexture texture;
    uint32_t origin_x, origin_y;
} WgcTexelCopyTexture;
int32_t wgc_command_encoder_copy_buffer_to_buffer(WgcCommandEncoder encoder, WgcBuffer source, uint64_t source_offset,
                                                  WgcBuffer destination, uint64_t destination_offset, uint64_t size);
int32_t wgc_command_encoder_copy_texture_to_buffer(WgcCommandEncoder encoder, const WgcTexelCopyTexture *source,
                                                   const WgcTexelCopyBuffer *destination, uint32_t width,
                                                   uint32_t height);
int32_t wgc_command_encoder_copy_buffer_to_texture(WgcCommandEncoder encoder, const WgcTexelCopyBuffer *source,
                                                   const WgcTexelCopyTexture *destination, uint32_t width,
                                                   uint32_t height);
int32_t wgc_command_encoder_copy_texture_to_texture(WgcCommandEncoder encoder, const WgcTexelCopyTexture *source,
                                                    const WgcTexelCopyTexture *destination, uint32_t width,
                                                    uint32_t height);
/* size 0: to the end of the buffer */
int32_t wgc_command_encoder_clear_buffer(WgcCommandEncoder encoder, WgcBuffer buffer, uint64_t offset, uint64_t size);
int32_t wgc_command_encoder_clear_texture(WgcCommandEncoder encoder, WgcTexture texture);
int32_t wgc_command_encoder_finish(WgcCommandEncoder encoder, WgcCommandBuffer *out);
void wgc_command_encoder_release(WgcCommandEncoder encoder);
void wgc_command_buffer_release(WgcCommandBuffer buffer);

/* ---- queue (QueueInterface::submit device.rs:436-462,
 *      DeviceInterface::poll device.rs:237-295) ---- */
/* Enqueues the command buffers on the device stream; they stay valid and may
 * be submitted again. */
int32_t wgc_queue_submit(WgcDevice device, uint32_t count, const WgcCommandBuffer *buffers,
                         uint64_t *submission_index);
/* The allocation-free variant of wgc_queue_submit.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * running while a HIP graph is captured, tight serving loops, harnesses that
 * forbid hipMalloc).  It calls no hipMalloc, hipFree, hipHostMalloc or other
 * allocating routine and creates no HIP objects: the launch groups' scratch
 * (bins, triangle records, lists, counters) lives in a static device array of
 * 128 MiB per GPU (held from the library's first use of that GPU until the
 * process ends), completion events and counter slots come from pools filled
 * at wgc_device_create (when all are in flight it waits for the oldest
 * submission), and plans run without batching tables and without graphs.
 * Results equal wgc_queue_submit's.  A command buffer whose scratch exceeds
 * the array returns WGC_ERR_OUT_OF_MEMORY before anything is launched.  The
 * array is shared by every device object on one GPU and keeps the last
 * submission's contents (it is cleared where it must start at zero):
 * allocation-free submissions of different device objects that overlap on
 * one GPU race, those of one device object run in stream order. */
int32_t wgc_queue_submit_noalloc(WgcDevice device, uint32_t count, const WgcCommandBuffer *buffers,
                                 uint64_t *submission_index);
/* Waits until `submission_index` has executed (UINT64_MAX: everything).
 * timeout_ns 0 = no timeout; a timeout of 1 ns checks without waiting
 * (WGC_ERR_TIMEOUT while it is still running). Returns WGC_ERR_EXECUTION if a
 * pass of a submission retired by this call failed. Waiting for one
 * submission does not wait for later ones. */
int32_t wgc_device_poll(WgcDevice device, uint64_t submission_index, uint64_t timeout_ns);
/* Counters of the last submission (waits for the queue): fragments of all its
 * command buffers' passes, primitives assembled, kernel launches. */
int32_t wgc_device_last_pass_stats(WgcDevice device, WgcPassStats *out);
/* Fragments and submission count summed over every submission that has
 * executed so far; does not wait (no reference counterpart). */
int32_t wgc_device_totals(WgcDevice device, uint64_t *fragments, uint64_t *submissions);

/* ---- execution options (no reference counterpart: the reference runs one
 *      pass at a time on one thread, engine.rs:26-46) ---- */
typedef enum WgcOption {
    /* raster tile shape: 0 = 16x16 px / 256 lanes, 1 = 8x8 / 64, 2 = 16x16 / 64,
     * 3 = 32x32 / 256, 4 = 16x16 / 128, 5 = 32x32 / 128, 6 = 32x32 / 512,
     * 7 = 32x32 / 64 (the autotuner's search space: tiles of 8, 16 and 32
     * pixels with 1, 2, 4 or 8 waves) */
    WGC_OPTION_RASTER_TILE = 1,
    /* 1 (default): consecutive independent passes with one draw each share
     * their launches (the batched form); 0: one launch set per draw */
    WGC_OPTION_BATCHING = 2,
    /* 1: a command buffer's launches are captured once into a hipGraph and
     * replayed on every later submission; 0 (default): launched directly
     * (cheaper on the host and on the GPU with this ROCm) */
    WGC_OPTION_GRAPHS = 3,
} WgcOption;
int32_t wgc_device_set_option(WgcDevice device, uint32_t option, uint32_t value);
int32_t wgc_device_get_option(WgcDevice device, uint32_t option, uint32_t *value);

/* Debug aid with no reference counterpart: with host_out == NULL, (re)allocates
 * a zeroed width*height u32 buffer to which every later pass adds, per pixel,
 * the number of fragments the scanline rule emitted; with host_out, waits for
 * the queue and copies that buffer out (width and height must match). */
int32_t wgc_device_debug_fragment_counts(WgcDevice device, uint32_t width, uint32_t height,
                                         uint32_t *host_out);

/* ---- measurement (no reference counterpart; render_pass/mod.rs:346,392 only
 *      logs the pass time) ---- */
/* Records an event on the device stream into slot (0..63). */
int32_t wgc_device_timestamp(WgcDevice device, uint32_t slot);
/* Milliseconds between two recorded slots (waits for the later one). */
int32_t wgc_device_elapsed_ms(WgcDevice device, uint32_t slot_begin, uint32_t slot_end,
                              float *ms);
/* Per-kernel timing with HIP events around every launch on the device
 * stream; off by default. */
int32_t wgc_device_set_kernel_timing(WgcDevice device, uint32_t enabled);
/* Sum of event-timed milliseconds and launch count of kernels whose name
 * contains `name`, since timing was enabled. */
int32_t wgc_device_kernel_time(WgcDevice device, const char *name, double *total_ms,
                               uint64_t *launches);

#ifdef __cplusplus
}
#endif
#endif /* WGC_H */
