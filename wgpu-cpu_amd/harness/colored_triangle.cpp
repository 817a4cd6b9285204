// colored_triangle.cpp -- drives the C ABI from C++ alone, the way the
// reference's wgpu-cpu-tests drive wgpu (wgpu-cpu-tests/src/tests/
// colored_triangle.rs:11-138): a WGSL module, a pipeline without vertex
// buffers, a 512x512 Rgba8UnormSrgb target with Depth32Float, one pass, a
// copy of the target into a buffer with 256-byte aligned rows, and a mapped
// read of it.  Writes the target's rows (2048 bytes each) to the output file.
//
// usage: colored_triangle <shader.wgsl> <out.rgba> [topology vertices front_face cull_mode]
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "wgc.h"

namespace {

int check(int32_t rc, const char *what)
{
    if (rc != WGC_OK) {
        std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, wgc_last_error());
        std::exit(1);
    }
    return rc;
}

void on_mapped(int32_t status, void *userdata) { *static_cast<int32_t *>(userdata) = status; }

}  // namespace

int main(int argc, char **argv)
{
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s shader.wgsl out.rgba [topology vertices front_face cull_mode]\n", argv[0]);
        return 2;
    }
    uint32_t topology = argc > 3 ? (uint32_t)std::atoi(argv[3]) : WGC_TOPOLOGY_TRIANGLE_LIST;
    uint32_t vertices = argc > 4 ? (uint32_t)std::atoi(argv[4]) : 3u;
    uint32_t front_face = argc > 5 ? (uint32_t)std::atoi(argv[5]) : 0u;
    uint32_t cull_mode = argc > 6 ? (uint32_t)std::atoi(argv[6]) : 0u;
    std::ifstream f(argv[1]);
    std::stringstream src;
    src << f.rdbuf();
    std::string wgsl = src.str();

    const uint32_t w = 512, h = 512, pitch = w * 4;
    WgcDeviceDescriptor dd{0, WGC_SEMANTICS_REFERENCE};
    WgcDevice dev = nullptr;
    check(wgc_device_create(&dd, &dev), "device");
    WgcShaderModule module = nullptr;
    check(wgc_shader_module_create_wgsl(dev, wgsl.data(), wgsl.size(), &module), "shader module");
    WgcTexture target = nullptr, depth = nullptr;
    check(wgc_texture_create(dev, w, h, WGC_TEXTURE_FORMAT_RGBA8_UNORM_SRGB, &target), "target");
    check(wgc_texture_create(dev, w, h, WGC_TEXTURE_FORMAT_DEPTH32_FLOAT, &depth), "depth");

    WgcRenderPipelineDescriptor pd{};
    pd.vertex_module = module;
    pd.topology = topology;
    pd.front_face = front_face;
    pd.cull_mode = cull_mode;
    pd.has_depth_stencil = 1;
    pd.depth_format = WGC_TEXTURE_FORMAT_DEPTH32_FLOAT;
    pd.depth_write_enabled = 1;
    pd.depth_compare = WGC_COMPARE_LESS;
    pd.fragment_module = module;
    pd.target_format = WGC_TEXTURE_FORMAT_RGBA8_UNORM_SRGB;
    WgcRenderPipeline pipeline = nullptr;
    check(wgc_render_pipeline_create(dev, &pd, &pipeline), "pipeline");

    WgcCommandEncoder enc = nullptr;
    check(wgc_command_encoder_create(dev, &enc), "encoder");
    WgcColorAttachment ca{};
    ca.view = target;
    ca.load_op = WGC_LOAD_OP_CLEAR;
    ca.clear_value[3] = 1.0;
    WgcDepthStencilAttachment da{};
    da.view = depth;
    da.has_depth_ops = 1;
    da.depth_load_op = WGC_LOAD_OP_CLEAR;
    da.depth_clear_value = 1.0f;
    WgcRenderPassDescriptor rpd{};
    rpd.color_attachment_count = 1;
    rpd.color_attachments = &ca;
    rpd.depth_stencil_attachment = &da;
    WgcRenderPass pass = nullptr;
    check(wgc_command_encoder_begin_render_pass(enc, &rpd, &pass), "begin pass");
    check(wgc_render_pass_set_pipeline(pass, pipeline), "set pipeline");
    check(wgc_render_pass_draw(pass, 0, vertices, 0, 1), "draw");
    check(wgc_render_pass_end(pass), "end pass");

    WgcBuffer readback = nullptr;
    check(wgc_buffer_create(dev, (uint64_t)pitch * h, &readback), "readback buffer");
    WgcTexelCopyTexture src_t{target, 0, 0};
    WgcTexelCopyBuffer dst_b{readback, 0, pitch};
    check(wgc_command_encoder_copy_texture_to_buffer(enc, &src_t, &dst_b, w, h), "copy");
    WgcCommandBuffer cb = nullptr;
    check(wgc_command_encoder_finish(enc, &cb), "finish");
    uint64_t index = 0;
    check(wgc_queue_submit(dev, 1, &cb, &index), "submit");

    int32_t mapped = 1;
    check(wgc_buffer_map_async(readback, WGC_MAP_READ, 0, (uint64_t)pitch * h, on_mapped, &mapped), "map_async");
    check(wgc_device_poll(dev, UINT64_MAX, UINT64_MAX), "poll");
    if (mapped != WGC_OK) {
        std::fprintf(stderr, "mapping failed: %d\n", mapped);
        return 1;
    }
    void *data = nullptr;
    check(wgc_buffer_get_mapped_range(readback, 0, (uint64_t)pitch * h, &data), "get_mapped_range");
    std::ofstream out(argv[2], std::ios::binary);
    out.write(static_cast<const char *>(data), (std::streamsize)pitch * h);
    out.close();
    check(wgc_buffer_unmap(readback), "unmap");

    wgc_command_buffer_release(cb);
    wgc_command_encoder_release(enc);
    wgc_buffer_release(readback);
    wgc_render_pipeline_release(pipeline);
    wgc_texture_release(depth);
    wgc_texture_release(target);
    wgc_shader_module_release(module);
    wgc_device_release(dev);
    std::printf("wrote %s\n", argv[2]);
    return 0;
}
