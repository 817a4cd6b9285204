// WGC_PROGRAM_MVP_TEXTURE: camera matrix times a vec4 position; a vec2 uv
// varying; the fragment stage returns textureSample of group 1.
struct Camera { matrix: mat4x4f, }
@group(0) @binding(0) var<uniform> camera: Camera;
@group(1) @binding(0) var albedo: texture_2d<f32>;
@group(1) @binding(1) var albedo_sampler: sampler;

struct VsIn {
    @location(0) position: vec4f,
    @location(1) uv: vec2f,
}
struct VsOut {
    @builtin(position) position: vec4f,
    @location(0) @interpolate(linear) uv: vec2f,
}

@vertex
fn vs_main(v: VsIn) -> VsOut {
    return VsOut(camera.matrix * v.position, v.uv);
}

@fragment
fn fs_main(f: VsOut) -> @location(0) vec4f {
    return textureSample(albedo, albedo_sampler, f.uv);
}
