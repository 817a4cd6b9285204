// WGC_PROGRAM_COLORED_TRIANGLE: no vertex buffers; the position and a
// one-hot colour come from vertex_index % 3.
struct VsOut {
    @builtin(position) position: vec4f,
    @location(0) @interpolate(linear) color: vec4f,
}

@vertex
fn vs_main(@builtin(vertex_index) index: u32) -> VsOut {
    let k = index % 3u;
    let x = f32(i32(k) - 1);
    let y = f32(i32(k & 1u) * 2 - 1);
    let r = f32(k == 0u);
    let g = f32(k == 1u);
    let b = f32(k == 2u);
    return VsOut(vec4f(x, y, 0.0, 1.0), vec4f(r, g, b, 1.0));
}

@fragment
fn fs_main(f: VsOut) -> @location(0) vec4f {
    return f.color;
}
