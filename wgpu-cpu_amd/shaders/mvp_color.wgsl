// WGC_PROGRAM_MVP_COLOR: camera matrix times a vec4 position; a vec3 or vec4
// colour varying (the @interpolate attribute is chosen when the module is
// created); the fragment stage writes the colour (alpha 1 for vec3).
struct Camera { matrix: mat4x4f, }
@group(0) @binding(0) var<uniform> camera: Camera;

struct VsIn {
    @location(0) position: vec4f,
    @location(1) color: vec3f,
}
struct VsOut {
    @builtin(position) position: vec4f,
    @location(0) @interpolate(linear) color: vec3f,
}

@vertex
fn vs_main(v: VsIn) -> VsOut {
    return VsOut(camera.matrix * v.position, v.color);
}

@fragment
fn fs_main(f: VsOut) -> @location(0) vec4f {
    return vec4f(f.color.x, f.color.y, f.color.z, 1.0);
}
