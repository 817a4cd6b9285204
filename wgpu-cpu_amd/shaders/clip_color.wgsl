// WGC_PROGRAM_CLIP_COLOR: the vertex buffer holds clip-space positions; a
// vec3 or vec4 colour passes through (alpha 1 for vec3).
struct VsIn {
    @location(0) position: vec4f,
    @location(1) color: vec4f,
}
struct VsOut {
    @builtin(position) position: vec4f,
    @location(0) @interpolate(linear) color: vec4f,
}

@vertex
fn vs_main(v: VsIn) -> VsOut {
    return VsOut(v.position, v.color);
}

@fragment
fn fs_main(f: VsOut) -> @location(0) vec4f {
    return f.color;
}
