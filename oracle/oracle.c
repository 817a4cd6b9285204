/*
 * oracle.c -- CPU restatement of wgpu-cpu's render-pass executor.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Built with -ffp-contract=off and
 * SSE arithmetic so that every f32 operation rounds once, as the reference's
 * Rust code and Cranelift JIT do on x86-64.
 *
 * The structure follows the reference, not the GPU design: one thread, the
 * vertex shader once per index (state.rs:519-535), the clipper's plane-by-
 * plane queue (clipper.rs:756-775), the scanline iterator row by row
 * (scanline.rs:107-166) and the fragment shader once per emitted pixel.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct { float x, y, z, w; } v4;

/* ------------------------------------------------------------------ */
/* Scalar helpers with Rust semantics                                 */
/* ------------------------------------------------------------------ */

/* `f as u32` in Rust: NaN -> 0, saturating, truncating. */
static uint32_t f32_as_u32(float f)
{
    if (!(f == f)) return 0;
    if (f <= 0.0f) return 0;
    if (f >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)f;
}

/* texture.rs:376-379: (value * 255).clamp(0, 255) as u8 */
uint8_t oracle_f32_to_unorm8(float v)
{
    float x = v * 255.0f;
    if (x < 0.0f) x = 0.0f;     /* f32::clamp keeps NaN */
    if (x > 255.0f) x = 255.0f;
    if (!(x == x)) return 0;    /* `as u8` maps NaN to 0 */
    return (uint8_t)x;
}

/* f32::rem_euclid (std): r = self % rhs; if r < 0 { r + |rhs| } */
static float rem_euclid(float x, float m)
{
    float r = fmodf(x, m);
    if (r < 0.0f) r = r + fabsf(m);
    return r;
}

/* binding.rs:150-164 */
uint32_t oracle_texel_coordinate(float x, uint32_t address_mode, uint32_t size)
{
    switch (address_mode) {
    case OR_ADDR_CLAMP:
        if (x < 0.0f) x = 0.0f;
        if (x > 1.0f) x = 1.0f;
        break;
    case OR_ADDR_REPEAT:
        x = rem_euclid(x, 1.0f);
        break;
    case OR_ADDR_MIRROR: {
        float r = rem_euclid(x, 2.0f);
        x = (r <= 1.0f) ? r : 2.0f - r;
        break;
    }
    default:
        break;
    }
    /* (x * (size - 1) as f32).round() as u32; round = half away from zero */
    return f32_as_u32(roundf(x * (float)(size - 1u)));
}

static float srgb_table[256];
static int srgb_ready = 0;

const float *oracle_srgb_table(void)
{
    if (!srgb_ready) {
        for (int i = 0; i < 256; ++i) {
            double c = (double)i / 255.0;
            double l = (c <= 0.04045) ? c / 12.92 : pow((c + 0.055) / 1.055, 2.4);
            srgb_table[i] = (float)l;
        }
        srgb_ready = 1;
    }
    return srgb_table;
}

/* mod.rs:434-452 */
static int compare(uint32_t fn, float value, float reference)
{
    switch (fn) {
    case OR_CMP_NEVER: return 0;
    case OR_CMP_LESS: return value < reference;
    case OR_CMP_EQUAL: return value == reference;
    case OR_CMP_LESS_EQUAL: return value <= reference;
    case OR_CMP_GREATER: return value > reference;
    case OR_CMP_NOT_EQUAL: return value != reference;
    case OR_CMP_GREATER_EQUAL: return value >= reference;
    case OR_CMP_ALWAYS: return 1;
    default: return 0;
    }
}

/* ------------------------------------------------------------------ */
/* util/sort.rs and util/scanline.rs                                  */
/* ------------------------------------------------------------------ */

typedef struct { int64_t x, y; } ip2;

/* sort.rs:3-22: compare_and_swap(0,1), (0,2), (1,2); swap only on Greater */
static void bubblesort3_by_y(ip2 v[3])
{
    static const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    for (int k = 0; k < 3; ++k) {
        int l = pairs[k][0], r = pairs[k][1];
        if (v[l].y > v[r].y) { ip2 t = v[l]; v[l] = v[r]; v[r] = t; }
    }
}

static void bubblesort3_by_x(ip2 v[3])
{
    static const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    for (int k = 0; k < 3; ++k) {
        int l = pairs[k][0], r = pairs[k][1];
        if (v[l].x > v[r].x) { ip2 t = v[l]; v[l] = v[r]; v[r] = t; }
    }
}

void oracle_bubblesort3_i64(int64_t v[3])
{
    static const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    for (int k = 0; k < 3; ++k) {
        int l = pairs[k][0], r = pairs[k][1];
        if (v[l] > v[r]) { int64_t t = v[l]; v[l] = v[r]; v[r] = t; }
    }
}

/* One emitted scanline: row y, inclusive span [x1, x2]. */
typedef struct { int64_t y, x1, x2; } scanline_t;

/* Scanlines iterator state (scanline.rs:14-56, 77-171). */
typedef struct {
    int state; /* 0 degenerate line, 1 first half, 2 second half, 3 done */
    int64_t ac_x, ac_y, ab_x, ab_y, bc_x, bc_y;
    int64_t a_x, b_x, b_y, c_y, y, i, i2;
    int64_t line_y, line_x1, line_x2;
} scanlines_t;

static void scanlines_init(scanlines_t *s, const uint32_t fb[3][2])
{
    ip2 tri[3];
    for (int k = 0; k < 3; ++k) { tri[k].x = fb[k][0]; tri[k].y = fb[k][1]; }
    bubblesort3_by_y(tri);
    ip2 a = tri[0], b = tri[1], c = tri[2];
    s->ac_x = c.x - a.x; s->ac_y = c.y - a.y;
    s->ab_x = b.x - a.x; s->ab_y = b.y - a.y;
    s->bc_x = c.x - b.x; s->bc_y = c.y - b.y;
    if (s->ac_y == 0) {
        bubblesort3_by_x(tri);
        s->state = 0;
        s->line_y = a.y;
        s->line_x1 = tri[0].x;
        s->line_x2 = tri[2].x;
        return;
    }
    s->a_x = a.x; s->b_x = b.x; s->b_y = b.y; s->c_y = c.y;
    s->y = a.y; s->i = 0; s->i2 = 0;
    s->state = (s->ab_y == 0) ? 2 : 1;
}

/* Scanline::new_from_isize: order the span; a negative value panics. */
static int make_line(scanline_t *out, int64_t y, int64_t x1, int64_t x2)
{
    if (x1 > x2) { int64_t t = x1; x1 = x2; x2 = t; }
    if (y < 0 || x1 < 0 || x2 < 0 || y > 0xFFFFFFFFll || x2 > 0xFFFFFFFFll) return -1;
    out->y = y; out->x1 = x1; out->x2 = x2;
    return 0;
}

/* Returns 1 with a line, 0 when done, -1 where the reference panics.
 * Rust's isize `/` truncates toward zero, as C's `/` does. */
static int scanlines_next(scanlines_t *s, scanline_t *out)
{
    for (;;) {
        switch (s->state) {
        case 0:
            s->state = 3;
            return make_line(out, s->line_y, s->line_x1, s->line_x2) ? -1 : 1;
        case 3:
            return 0;
        case 1:
            if (s->y < s->b_y) {
                int64_t x1 = s->a_x + (s->ac_x * s->i) / s->ac_y;
                int64_t x2 = s->a_x + (s->ab_x * s->i) / s->ab_y;
                int64_t y = s->y;
                s->y += 1; s->i += 1;
                return make_line(out, y, x1, x2) ? -1 : 1;
            }
            if (s->bc_y == 0) s->state = 3;
            else { s->state = 2; s->i2 = 0; }
            break;
        case 2:
            if (s->y <= s->c_y) {
                int64_t x1 = s->a_x + (s->ac_x * s->i) / s->ac_y;
                int64_t x2 = s->b_x + (s->bc_x * s->i2) / s->bc_y;
                int64_t y = s->y;
                s->i2 += 1;
                s->y += 1; s->i += 1;
                return make_line(out, y, x1, x2) ? -1 : 1;
            }
            s->state = 3;
            break;
        default:
            return 0;
        }
    }
}

int64_t oracle_scanline_pixels(const uint32_t tri[6], uint32_t *out_xy, int64_t cap)
{
    uint32_t fb[3][2] = {{tri[0], tri[1]}, {tri[2], tri[3]}, {tri[4], tri[5]}};
    scanlines_t s;
    scanline_t line;
    int64_t n = 0;
    scanlines_init(&s, fb);
    for (;;) {
        int r = scanlines_next(&s, &line);
        if (r < 0) return -1;
        if (r == 0) break;
        for (int64_t x = line.x1; x <= line.x2; ++x) {
            if (n < cap && out_xy) {
                out_xy[2 * n] = (uint32_t)x;
                out_xy[2 * n + 1] = (uint32_t)line.y;
            }
            ++n;
        }
    }
    return n;
}

/* ------------------------------------------------------------------ */
/* Clipper (clipper.rs:150-179, 625-802)                              */
/* ------------------------------------------------------------------ */

typedef struct { float p[4]; float b[3]; } cvert; /* position + barycentric */
typedef struct { cvert v[3]; } ctri;

/* ClipVolume::WEBGPU (clipper.rs:153-163): L, R, B, T, near (z >= 0), far (z <= w) */
static const float planes[6][4] = {
    { 1.0f, 0.0f, 0.0f, 1.0f},
    {-1.0f, 0.0f, 0.0f, 1.0f},
    { 0.0f, 1.0f, 0.0f, 1.0f},
    { 0.0f, -1.0f, 0.0f, 1.0f},
    { 0.0f, 0.0f, 1.0f, 0.0f},
    { 0.0f, 0.0f, -1.0f, 1.0f},
};

/* nalgebra Vector4::dot: (p0*q0 + p2*q2) + (p1*q1 + p3*q3).  With the plane
 * coefficients in {0, +-1} every grouping gives the same value up to the
 * sign of an exact zero. */
static float clip_distance(const float *pl, const float *p)
{
    float a = p[0] * pl[0];
    float b = p[1] * pl[1];
    float c = p[2] * pl[2];
    float d = p[3] * pl[3];
    a = a + c;
    b = b + d;
    return a + b;
}

/* interpolation.rs:84-89: lerp(x0, x1, t) = x0 * (1 - t) + x1 * t, on
 * position and barycentric together (clipper.rs:602-622). */
static cvert clip_lerp(const cvert *x0, const cvert *x1, float t)
{
    cvert r;
    float s = 1.0f - t;
    for (int k = 0; k < 4; ++k) r.p[k] = x0->p[k] * s + x1->p[k] * t;
    for (int k = 0; k < 3; ++k) r.b[k] = x0->b[k] * s + x1->b[k] * t;
    return r;
}

/* clipper.rs:625-728; returns the number of output triangles (0..2). */
static int clip_against_plane(const ctri *in, const float *pl, ctri out[2])
{
    float d[3];
    int o[3];
    for (int k = 0; k < 3; ++k) {
        d[k] = clip_distance(pl, in->v[k].p);
        o[k] = d[k] < 0.0f;
    }
#define CLIP(i, j) clip_lerp(&in->v[i], &in->v[j], d[i] / (d[i] - d[j]))
    const cvert *A = &in->v[0], *B = &in->v[1], *C = &in->v[2];
    if (o[0] && o[1] && o[2]) return 0;
    if (!o[0] && !o[1] && !o[2]) { out[0] = *in; return 1; }
    if (o[0] && o[1] && !o[2]) {            /* clip(2,0), clip(1,2) */
        cvert an = CLIP(2, 0), bn = CLIP(1, 2);
        out[0].v[0] = an; out[0].v[1] = bn; out[0].v[2] = *C;
        return 1;
    }
    if (!o[0] && o[1] && o[2]) {            /* clip(0,1), clip(2,0) */
        cvert bn = CLIP(0, 1), cn = CLIP(2, 0);
        out[0].v[0] = *A; out[0].v[1] = bn; out[0].v[2] = cn;
        return 1;
    }
    if (o[0] && !o[1] && o[2]) {            /* clip(0,1), clip(1,2) */
        cvert an = CLIP(0, 1), cn = CLIP(1, 2);
        out[0].v[0] = an; out[0].v[1] = *B; out[0].v[2] = cn;
        return 1;
    }
    if (o[0] && !o[1] && !o[2]) {           /* clip(0,1), clip(0,2) */
        cvert bn = CLIP(0, 1), cn = CLIP(0, 2);
        out[0].v[0] = bn; out[0].v[1] = *B; out[0].v[2] = cn;
        out[1].v[0] = *B; out[1].v[1] = *C; out[1].v[2] = cn;
        return 2;
    }
    if (!o[0] && o[1] && !o[2]) {           /* clip(0,1), clip(1,2) */
        cvert an = CLIP(0, 1), cn = CLIP(1, 2);
        out[0].v[0] = *A; out[0].v[1] = an; out[0].v[2] = cn;
        out[1].v[0] = *A; out[1].v[1] = cn; out[1].v[2] = *C;
        return 2;
    }
    /* !o[0] && !o[1] && o[2]: clip(1,2), clip(0,2) */
    {
        cvert bn = CLIP(1, 2), an = CLIP(0, 2);
        out[0].v[0] = *A; out[0].v[1] = *B; out[0].v[2] = bn;
        out[1].v[0] = *A; out[1].v[1] = bn; out[1].v[2] = an;
        return 2;
    }
#undef CLIP
}

/* clipper.rs:756-775: a FIFO queue, processed plane by plane.  At most
 * 2^6 triangles can come out. */
#define CLIP_CAP 64
static int clip_triangle(const float pos[3][4], ctri *out)
{
    ctri qa[CLIP_CAP], qb[CLIP_CAP];
    ctri *cur = qa, *nxt = qb;
    int n = 1;
    for (int k = 0; k < 3; ++k) {
        for (int c = 0; c < 4; ++c) cur[0].v[k].p[c] = pos[k][c];
        for (int c = 0; c < 3; ++c) cur[0].v[k].b[c] = (c == k) ? 1.0f : 0.0f;
    }
    for (int pl = 0; pl < 6; ++pl) {
        int m = 0;
        for (int t = 0; t < n; ++t) {
            ctri res[2];
            int r = clip_against_plane(&cur[t], planes[pl], res);
            for (int q = 0; q < r; ++q) nxt[m++] = res[q];
        }
        ctri *tmp = cur; cur = nxt; nxt = tmp;
        n = m;
    }
    for (int t = 0; t < n; ++t) out[t] = cur[t];
    return n;
}

int oracle_clip_triangle(const float in_xyzw[12], float *out, int cap)
{
    float pos[3][4];
    ctri res[CLIP_CAP];
    for (int k = 0; k < 3; ++k)
        for (int c = 0; c < 4; ++c) pos[k][c] = in_xyzw[4 * k + c];
    int n = clip_triangle(pos, res);
    for (int t = 0; t < n && t < cap; ++t)
        for (int k = 0; k < 3; ++k) {
            float *o = out + (t * 3 + k) * 7;
            for (int c = 0; c < 4; ++c) o[c] = res[t].v[k].p[c];
            for (int c = 0; c < 3; ++c) o[4 + c] = res[t].v[k].b[c];
        }
    return n;
}

/* ------------------------------------------------------------------ */
/* Raster (raster.rs:129-160, 280-362)                                */
/* ------------------------------------------------------------------ */

/* ToRaster::new + to_raster */
void oracle_to_raster(const float viewport[6], const float clip[4], uint32_t fb[2], float frag[4])
{
    float tx = viewport[0] + 0.5f * viewport[2];
    float ty = viewport[1] + 0.5f * viewport[3];
    float sx = 0.5f * viewport[2];
    float sy = 0.5f * viewport[3];
    sy = sy * -1.0f;
    /* Point3::from_homogeneous: coords / w */
    float nx = clip[0] / clip[3];
    float ny = clip[1] / clip[3];
    float nz = clip[2] / clip[3];
    float pd = 1.0f / clip[3];
    float vx = nx * sx + tx;
    float vy = ny * sy + ty;
    vx = fmaxf(vx, 0.0f);
    vy = fmaxf(vy, 0.0f);
    fb[0] = f32_as_u32(vx);
    fb[1] = f32_as_u32(vy);
    frag[0] = vx; frag[1] = vy; frag[2] = nz; frag[3] = pd;
}

/* shoelace(a, b, c) (raster.rs:294-298), evaluated left to right */
static float shoelace(float ax, float ay, float bx, float by, float cx, float cy)
{
    float t1 = (by - ay) * (bx + ax);
    float t2 = (cy - by) * (cx + bx);
    float t3 = (ay - cy) * (ax + cx);
    return (t1 + t2) + t3;
}

/* ------------------------------------------------------------------ */
/* Vertex stage: attribute fetch + the built-in programs              */
/* ------------------------------------------------------------------ */

typedef struct {
    float clip[4];
    float var[OR_MAX_VAR];
    int nvar;
} vout;

/* vertex.rs:191-195 + :143-147: raw bytes of the attribute; the reference
 * panics on an out-of-range slice. */
static int fetch_attr(const OrDraw *d, uint32_t loc, uint32_t instance, uint32_t vertex,
                      float *dst, int *ncomp)
{
    if (loc >= d->n_attr) return -1;
    const OrAttribute *at = &d->attr[loc];
    if (at->buffer >= d->n_vb) return -1;
    const OrVertexBuffer *vb = &d->vb[at->buffer];
    uint64_t start = vb->step_instance ? (uint64_t)instance * vb->stride
                                       : (uint64_t)vertex * vb->stride;
    uint64_t n = at->format;
    if (start + at->offset + 4 * n > vb->size) return -1;
    memcpy(dst, vb->data + start + at->offset, 4 * n);
    *ncomp = (int)n;
    return 0;
}

/* naga-cranelift binary.rs:299-321: sum_i column_i * v_i, separate fmul/fadd */
static void mat4_mul_vec4(const float *m, const float *v, float *r)
{
    for (int k = 0; k < 4; ++k) {
        float s = m[0 * 4 + k] * v[0];
        s = s + m[1 * 4 + k] * v[1];
        s = s + m[2 * 4 + k] * v[2];
        s = s + m[3 * 4 + k] * v[3];
        r[k] = s;
    }
}

void oracle_mat4_mul_vec4(const float m[16], const float v[4], float r[4])
{
    mat4_mul_vec4(m, v, r);
}

static int run_vs(const OrDraw *d, uint32_t instance, uint32_t vertex, vout *o)
{
    int nc;
    float pos[4];
    o->nvar = 0;
    switch (d->program) {
    case OR_PROG_MVP_COLOR:
    case OR_PROG_MVP_TEXTURE:
        if (fetch_attr(d, 0, instance, vertex, pos, &nc) || nc != 4) return -1;
        if (!d->uniform) return -1;
        mat4_mul_vec4(d->uniform, pos, o->clip);
        if (fetch_attr(d, 1, instance, vertex, o->var, &nc)) return -1;
        o->nvar = nc;
        return 0;
    case OR_PROG_CLIP_COLOR:
        if (fetch_attr(d, 0, instance, vertex, pos, &nc) || nc != 4) return -1;
        memcpy(o->clip, pos, sizeof pos);
        if (fetch_attr(d, 1, instance, vertex, o->var, &nc)) return -1;
        o->nvar = nc;
        return 0;
    case OR_PROG_COLORED_TRIANGLE: {
        /* colored_triangle.wgsl: vi = vertex_index % 3;
         * x = f32(i32(vi) - 1); y = f32(i32(vi & 1u) * 2 - 1);
         * colour = (f32(vi == 0), f32(vi == 1), f32(vi == 2), 1.0) */
        uint32_t vi = vertex % 3u;
        int32_t ix = (int32_t)vi - 1;
        int32_t iy = (int32_t)(vi & 1u) * 2 - 1;
        o->clip[0] = (float)ix;
        o->clip[1] = (float)iy;
        o->clip[2] = 0.0f;
        o->clip[3] = 1.0f;
        o->var[0] = (vi == 0) ? 1.0f : 0.0f;
        o->var[1] = (vi == 1) ? 1.0f : 0.0f;
        o->var[2] = (vi == 2) ? 1.0f : 0.0f;
        o->var[3] = 1.0f;
        o->nvar = 4;
        return 0;
    }
    default:
        return -1;
    }
}

/* ------------------------------------------------------------------ */
/* Fragment stage                                                     */
/* ------------------------------------------------------------------ */

static void fetch_texel(const OrDraw *d, uint32_t u, uint32_t v, float out[4])
{
    const uint8_t *t = d->texture + ((uint64_t)v * d->tex_w + u) * 4u;
    if (d->srgb_decode && d->tex_srgb) {
        const float *lut = oracle_srgb_table();
        for (int k = 0; k < 3; ++k) out[k] = lut[t[k]];
        out[3] = (float)t[3] / 255.0f;
    } else {
        for (int k = 0; k < 4; ++k) out[k] = (float)t[k] / 255.0f; /* texture.rs:173-175 */
    }
}

/* Extension: bilinear filtering with the WebGPU texel-centre convention,
 * u' = u*W - 0.5, on wrapped integer coordinates. */
static int32_t wrap_index(int32_t i, uint32_t mode, uint32_t size)
{
    int32_t n = (int32_t)size;
    switch (mode) {
    case OR_ADDR_REPEAT: { int32_t r = i % n; return r < 0 ? r + n : r; }
    case OR_ADDR_MIRROR: {
        int32_t p = 2 * n;
        int32_t r = i % p; if (r < 0) r += p;
        return r < n ? r : p - 1 - r;
    }
    default: return i < 0 ? 0 : (i >= n ? n - 1 : i);
    }
}

static void sample_bilinear(const OrDraw *d, float u, float v, float out[4])
{
    float fu = u * (float)d->tex_w - 0.5f;
    float fv = v * (float)d->tex_h - 0.5f;
    float u0f = floorf(fu), v0f = floorf(fv);
    float au = fu - u0f, av = fv - v0f;
    /* coordinates beyond +-2^30 are left to the clamp of the cast */
    int32_t iu = (int32_t)fmaxf(fminf(u0f, 1073741824.0f), -1073741824.0f);
    int32_t iv = (int32_t)fmaxf(fminf(v0f, 1073741824.0f), -1073741824.0f);
    uint32_t x0 = (uint32_t)wrap_index(iu, d->addr_u, d->tex_w);
    uint32_t x1 = (uint32_t)wrap_index(iu + 1, d->addr_u, d->tex_w);
    uint32_t y0 = (uint32_t)wrap_index(iv, d->addr_v, d->tex_h);
    uint32_t y1 = (uint32_t)wrap_index(iv + 1, d->addr_v, d->tex_h);
    float t00[4], t10[4], t01[4], t11[4];
    fetch_texel(d, x0, y0, t00);
    fetch_texel(d, x1, y0, t10);
    fetch_texel(d, x0, y1, t01);
    fetch_texel(d, x1, y1, t11);
    float su = 1.0f - au, sv = 1.0f - av;
    for (int k = 0; k < 4; ++k) {
        float top = t00[k] * su + t10[k] * au;
        float bot = t01[k] * su + t11[k] * au;
        out[k] = top * sv + bot * av;
    }
}

/* binding.rs:93-141 (reference path) */
static void texture_sample(const OrDraw *d, float u, float v, float out[4])
{
    if (d->bilinear) { sample_bilinear(d, u, v, out); return; }
    uint32_t x = oracle_texel_coordinate(u, d->addr_u, d->tex_w);
    uint32_t y = oracle_texel_coordinate(v, d->addr_v, d->tex_h);
    fetch_texel(d, x, y, out);
}

static void run_program_fs(const OrDraw *d, const float *var, int nvar, float out[4])
{
    switch (d->program) {
    case OR_PROG_MVP_TEXTURE:
        texture_sample(d, var[0], var[1], out);
        return;
    default:
        /* fs returns the colour varying; a vec3 becomes
         * vec4f(c.x, c.y, c.z, 1.0) (scalar-composed vector) */
        out[0] = var[0]; out[1] = var[1]; out[2] = var[2];
        out[3] = (nvar >= 4) ? var[3] : 1.0f;
        return;
    }
}

/* The fragment stage of a draw: the program's shader, then the test
 * program's steps in the order its WGSL states them.  Returns 0 when the
 * shader discards (EntryPointError::Killed, fragment.rs:202-213); *frag_depth
 * receives a written @builtin(frag_depth). */
static int run_fs(const OrDraw *d, const float *var, int nvar, uint32_t prim, float out[4], float *frag_depth)
{
    if (d->fs_flags & OR_FS_DISCARD) {
        if (var[3] < 0.5f) return 0;
    }
    run_program_fs(d, var, nvar, out);
    if (d->fs_flags & OR_FS_PRIM_RED) out[0] = (float)(prim % 8u) * 0.125f;
    if (d->fs_flags & OR_FS_DEPTH_ALPHA) *frag_depth = var[3];
    if (d->fs_flags & OR_FS_DEPTH_CONST) *frag_depth = 0.3125f;
    return 1;
}

/* texture.rs:213-218 + 373-411 */
static void put_pixel(OrPass *p, uint32_t x, uint32_t y, const float c[4])
{
    uint8_t *t = p->color + ((uint64_t)y * p->width + x) * 4u;
    uint8_t r = oracle_f32_to_unorm8(c[0]), g = oracle_f32_to_unorm8(c[1]);
    uint8_t b = oracle_f32_to_unorm8(c[2]), a = oracle_f32_to_unorm8(c[3]);
    if (p->color_format == OR_FMT_BGRA8UNORM || p->color_format == OR_FMT_BGRA8UNORM_SRGB) {
        t[0] = b; t[1] = g; t[2] = r; t[3] = a;
    } else {
        t[0] = r; t[1] = g; t[2] = b; t[3] = a;
    }
}

/* Extension: src-over blending, BlendState::ALPHA_BLENDING:
 * rgb = src*srcA + dst*(1-srcA); a = srcA + dstA*(1-srcA), the destination
 * read back as u8/255 and the result written with the truncating writer. */
static void blend_pixel(OrPass *p, uint32_t x, uint32_t y, const float s[4])
{
    uint8_t *t = p->color + ((uint64_t)y * p->width + x) * 4u;
    int bgra = (p->color_format == OR_FMT_BGRA8UNORM || p->color_format == OR_FMT_BGRA8UNORM_SRGB);
    float dst[4];
    dst[0] = (float)t[bgra ? 2 : 0] / 255.0f;
    dst[1] = (float)t[1] / 255.0f;
    dst[2] = (float)t[bgra ? 0 : 2] / 255.0f;
    dst[3] = (float)t[3] / 255.0f;
    float ia = 1.0f - s[3];
    float o[4];
    for (int k = 0; k < 3; ++k) o[k] = s[k] * s[3] + dst[k] * ia;
    o[3] = s[3] + dst[3] * ia;
    put_pixel(p, x, y, o);
}

/* ------------------------------------------------------------------ */
/* Draw (state.rs:490-593)                                            */
/* ------------------------------------------------------------------ */

static int fail(OrPass *p, const char *msg)
{
    p->error = 1;
    snprintf(p->message, sizeof p->message, "%s", msg);
    return -1;
}

/* index.rs:45-70: u32::from(index).strict_add_signed(base_vertex) */
static int resolve_index(const OrDraw *d, uint32_t i, uint32_t *vertex)
{
    if (!d->indexed) { *vertex = i; return 0; }
    uint64_t off = (uint64_t)i * d->index_format;
    if (off + d->index_format > d->index_size) return -1;
    uint32_t idx;
    if (d->index_format == 2) {
        uint16_t s; memcpy(&s, d->index_data + off, 2); idx = s;
    } else {
        memcpy(&idx, d->index_data + off, 4);
    }
    int64_t v = (int64_t)idx + d->base_vertex;
    if (v < 0 || v > 0xFFFFFFFFll) return -1;
    *vertex = (uint32_t)v;
    return 0;
}

/* ------------------------------------------------------------------ */
/* Primitive assembly (primitive.rs:206-487, index.rs:72-106)         */
/* ------------------------------------------------------------------ */

static int topo_vertices(uint32_t topology)
{
    switch (topology) {
    case OR_TOPO_POINT_LIST: return 1;
    case OR_TOPO_LINE_LIST:
    case OR_TOPO_LINE_STRIP: return 2;
    default: return 3;
    }
}

/* TriStripGenerator::push (primitive.rs:349-365) */
typedef struct {
    uint32_t buf[2];
    int odd;
} strip_gen;

static void strip_push(strip_gen *g, uint32_t v, uint32_t out[3])
{
    out[0] = g->buf[0];
    out[1] = g->buf[1];
    out[2] = v;
    g->buf[g->odd] = v;
    g->odd = !g->odd;
}

int64_t oracle_assemble(uint32_t topology, uint32_t restart, uint32_t index_format, const uint8_t *indices,
                        uint32_t count, uint32_t *out, int64_t cap)
{
    int n = topo_vertices(topology);
    int64_t np = 0;
#define EMIT(...)                                                         \
    do {                                                                  \
        uint32_t v_[3] = {__VA_ARGS__};                                   \
        if (np < cap)                                                     \
            for (int k_ = 0; k_ < n; ++k_) out[np * n + k_] = v_[k_];     \
        ++np;                                                             \
    } while (0)
    int strip = topology == OR_TOPO_LINE_STRIP || topology == OR_TOPO_TRIANGLE_STRIP;
    if (restart && !strip) return -1;   /* state.rs:306-310: assert!(is_strip) */
    if (!strip) {
        /* List: array_chunks_::<n>, a remainder is dropped */
        for (uint32_t p = 0; p < count / (uint32_t)n; ++p) EMIT(p * n, p * n + 1, p * n + 2);
        return np;
    }
    if (!restart || !indices) {
        if (topology == OR_TOPO_LINE_STRIP) {
            /* tuple_windows (primitive.rs:270-282) */
            for (uint32_t k = 0; k + 1 < count; ++k) EMIT(k, k + 1, 0);
        } else if (count >= 2) {
            /* TriStripIter (primitive.rs:367-398) */
            strip_gen g = {{0, 1}, 0};
            for (uint32_t k = 2; k < count; ++k) {
                uint32_t t[3];
                strip_push(&g, k, t);
                EMIT(t[0], t[1], t[2]);
            }
        }
        return np;
    }
    uint32_t sep = index_format == 2 ? 0xFFFFu : 0xFFFFFFFFu;
#define IS_SEP(i) (index_format == 2 ? ((uint32_t)indices[2 * (i)] | ((uint32_t)indices[2 * (i) + 1] << 8)) == sep \
                                     : ((uint32_t)indices[4 * (i)] | ((uint32_t)indices[4 * (i) + 1] << 8) |      \
                                        ((uint32_t)indices[4 * (i) + 2] << 16) |                                  \
                                        ((uint32_t)indices[4 * (i) + 3] << 24)) == sep)
    if (topology == OR_TOPO_LINE_STRIP) {
        /* Strip<2> with separators (primitive.rs:285-318) */
        uint32_t buf[2];
        int nb = 0;
        for (uint32_t k = 0; k < count; ++k) {
            if (IS_SEP(k)) {
                nb = 0;
                continue;
            }
            buf[nb++] = k;
            if (nb == 2) {
                EMIT(buf[0], buf[1], 0);
                buf[0] = buf[1];
                nb = 1;
            }
        }
        return np;
    }
    /* ResetTriStripIter (primitive.rs:400-487) */
    uint32_t k = 0;
    for (;;) {
        /* try_fill_buffer: two vertices, a separator empties the buffer */
        uint32_t buf[2];
        int nb = 0;
        while (k < count && nb < 2) {
            if (IS_SEP(k)) nb = 0;
            else buf[nb++] = k;
            ++k;
        }
        if (nb < 2) return np;
        strip_gen g = {{buf[0], buf[1]}, 0};
        int refill = 0;
        while (k < count) {
            uint32_t i = k++;
            if (IS_SEP(i)) {
                refill = 1;
                break;
            }
            uint32_t t[3];
            strip_push(&g, i, t);
            EMIT(t[0], t[1], t[2]);
        }
        if (!refill) return np;
    }
#undef IS_SEP
#undef EMIT
}

/* ------------------------------------------------------------------ */
/* Lines and points (clipper.rs:207-407, raster.rs:184-265,           */
/* util/bresenham.rs:78-153)                                          */
/* ------------------------------------------------------------------ */

typedef struct {
    int64_t x, y;
} p64;

/* Octant::from_points (bresenham.rs:20-44) */
static int octant_of(p64 s, p64 e)
{
    int64_t dx = e.x - s.x, dy = e.y - s.y;
    int o = 0;
    if (dy < 0) {
        dx = -dx;
        dy = -dy;
        o += 4;
    }
    if (dx < 0) {
        int64_t t = dx;
        dx = dy;
        dy = -t;
        o += 2;
    }
    if (dx < dy) o += 1;
    return o;
}

static p64 to_octant0(int o, p64 p)
{
    switch (o) {
    case 0: return (p64){p.x, p.y};
    case 1: return (p64){p.y, p.x};
    case 2: return (p64){p.y, -p.x};
    case 3: return (p64){-p.x, p.y};
    case 4: return (p64){-p.x, -p.y};
    case 5: return (p64){-p.y, -p.x};
    case 6: return (p64){-p.y, p.x};
    default: return (p64){p.x, -p.y};
    }
}

static p64 from_octant0(int o, p64 p)
{
    switch (o) {
    case 0: return (p64){p.x, p.y};
    case 1: return (p64){p.y, p.x};
    case 2: return (p64){-p.y, p.x};
    case 3: return (p64){-p.x, p.y};
    case 4: return (p64){-p.x, -p.y};
    case 5: return (p64){-p.y, -p.x};
    case 6: return (p64){p.y, -p.x};
    default: return (p64){p.x, -p.y};
    }
}

typedef struct {
    p64 x, d;
    int64_t x1, diff;
    int octant;
    float t, dt;
} bres_t;

static void bres_init(bres_t *b, const uint32_t start[2], const uint32_t end[2])
{
    p64 s = {(int64_t)start[0], (int64_t)start[1]}, e = {(int64_t)end[0], (int64_t)end[1]};
    b->octant = octant_of(s, e);
    s = to_octant0(b->octant, s);
    e = to_octant0(b->octant, e);
    b->d = (p64){e.x - s.x, e.y - s.y};
    b->dt = 1.0f / (float)b->d.x;
    b->x = s;
    b->x1 = e.x;
    b->diff = b->d.y - b->d.x;
    b->t = 0.0f;
}

/* Bresenham::next: 1 and the point and its t, or 0 at the end */
static int bres_next(bres_t *b, uint32_t xy[2], float *t)
{
    if (b->x.x >= b->x1) return 0;
    p64 p = b->x;
    if (b->diff >= 0) {
        b->x.y += 1;
        b->diff -= b->d.x;
    }
    b->diff += b->d.y;
    *t = b->t;
    b->t = b->t + b->dt;
    b->x.x += 1;
    p64 q = from_octant0(b->octant, p);
    xy[0] = (uint32_t)q.x;
    xy[1] = (uint32_t)q.y;
    return 1;
}

int64_t oracle_bresenham(const uint32_t start[2], const uint32_t end[2], uint32_t *out_xy, float *out_t,
                         int64_t cap)
{
    bres_t b;
    bres_init(&b, start, end);
    int64_t n = 0;
    uint32_t xy[2];
    float t;
    while (bres_next(&b, xy, &t)) {
        if (n < cap) {
            out_xy[2 * n] = xy[0];
            out_xy[2 * n + 1] = xy[1];
            out_t[n] = t;
        }
        ++n;
    }
    return n;
}

/* lerp (interpolation.rs:84-89) on a Vector4: x0 * (1 - t) + x1 * t */
static void lerp4(const float *x0, const float *x1, float t, float *out)
{
    float s = 1.0f - t;
    for (int c = 0; c < 4; ++c) {
        float a = x0[c] * s;
        float b = x1[c] * t;
        out[c] = a + b;
    }
}

int oracle_clip_line(const float a[4], const float b[4], float out[8], float alphas[2])
{
    float dist[2][6];
    unsigned oc[2] = {0, 0};
    for (int k = 0; k < 2; ++k)
        for (int pl = 0; pl < 6; ++pl) {
            dist[k][pl] = clip_distance(planes[pl], k ? b : a);
            if (dist[k][pl] < 0.0f) oc[k] |= 1u << pl;
        }
    unsigned any = oc[0] | oc[1], both = oc[0] & oc[1];
    if (!any) {
        memcpy(out, a, 16);
        memcpy(out + 4, b, 16);
        alphas[0] = 0.0f;
        alphas[1] = 1.0f;
        return 1;
    }
    if (both) return 0;
    float al[2] = {0.0f, 1.0f};
    for (int pl = 0; pl < 6; ++pl) {
        unsigned bit = 1u << pl;
        if (!(any & bit)) continue;
        float da = dist[0][pl], db = dist[1][pl];
        float alpha = da / (da - db);
        if (oc[0] & bit) {
            if (alpha > al[0]) al[0] = alpha;
        } else {
            if (alpha < al[1]) al[1] = alpha;
        }
        if (al[0] > al[1]) return 0;
    }
    /* update_point (clipper.rs:343-356): the second update reads the
     * already updated first point */
    float p[2][4];
    memcpy(p[0], a, 16);
    memcpy(p[1], b, 16);
    if (oc[0]) {
        float u[4];
        lerp4(p[0], p[1], al[0], u);
        memcpy(p[0], u, 16);
    }
    if (oc[1]) {
        float u[4];
        lerp4(p[0], p[1], al[1], u);
        memcpy(p[1], u, 16);
    }
    memcpy(out, p[0], 16);
    memcpy(out + 4, p[1], 16);
    alphas[0] = al[0];
    alphas[1] = al[1];
    return 1;
}

/* FragmentOutput::depth_test (fragment.rs:427-453): compares `z` with the
 * stored depth and writes it when the test passes and writes are on; true
 * without a depth attachment or depth state. */
static int depth_test(OrPass *p, const OrDraw *d, uint32_t x, uint32_t y, float z)
{
    if (!(p->depth && d->depth_enabled)) return 1;
    float *dp = p->depth + (uint64_t)y * p->width + x;
    if (!compare(d->depth_compare, z, *dp)) return 0;
    if (d->depth_write) *dp = z;
    return 1;
}

/* FragmentProcessingState::process (fragment.rs:127-214) with the output
 * side (fragment.rs:457-488): the early depth test of @early_depth_test
 * (force: early only; allow: early and late), the fragment shader, then its
 * outputs in member order -- a frag_depth member replaces the depth the
 * late test uses only when it comes before the @location member, where
 * the late test runs -- and the colour write (or src-over blend). */
static void write_fragment(OrPass *p, const OrDraw *d, uint32_t x, uint32_t y, float z, const float *var, int nvar,
                           uint32_t prim)
{
    int known = 0;
    if (d->fs_early) {
        if (!depth_test(p, d, x, y, z)) return;   /* early reject */
        known = d->fs_early == 1;                 /* force: the early result stands */
    }
    float color[4];
    float frag_depth = z;
    if (!run_fs(d, var, nvar, prim, color, &frag_depth)) return;
    float tz = (d->fs_flags & OR_FS_DEPTH_FIRST) ? frag_depth : z;
    if (!known && !depth_test(p, d, x, y, tz)) return;
    if (p->color) {
        if (d->blend) blend_pixel(p, x, y, color);
        else put_pixel(p, x, y, color);
    }
}

static int in_scissor(const OrDraw *d, uint32_t x, uint32_t y)
{
    uint64_t sx0 = d->scissor[0], sy0 = d->scissor[1];
    uint64_t sx1 = sx0 + d->scissor[2], sy1 = sy0 + d->scissor[3];
    return x >= sx0 && x < sx1 && y >= sy0 && y < sy1;
}

/* PointRasterizer (raster.rs:198-215): no clipping; the vertex's own
 * varyings (NoInterpolation) */
static int draw_point(OrPass *p, const OrDraw *d, const vout *v, uint32_t prim)
{
    if (v->clip[3] == 0.0f) return fail(p, "clip position w = 0");
    uint32_t fb[2];
    float fr[4];
    oracle_to_raster(d->viewport, v->clip, fb, fr);
    if (!in_scissor(d, fb[0], fb[1])) return 0;
    p->fragments++;
    if (p->frag_count) p->frag_count[(uint64_t)fb[1] * p->width + fb[0]]++;
    write_fragment(p, d, fb[0], fb[1], fr[2], v->var, v->nvar, prim);
    return 0;
}

/* CohenSutherland + LineRasterizer (raster.rs:230-265): Bresenham between
 * the clipped end points; L = lerp of the clip alphas by t; the fragment is
 * L applied to the clipped end points' (x, y, z, 1/w), the varyings L
 * applied to the unclipped vertices'. */
static int draw_line(OrPass *p, const OrDraw *d, const vout v[2], uint32_t prim)
{
    float cp[8], al[2];
    if (!oracle_clip_line(v[0].clip, v[1].clip, cp, al)) return 0;
    if (cp[3] == 0.0f || cp[7] == 0.0f) return fail(p, "clip position w = 0");
    uint32_t fb[2][2];
    float fr[2][4];
    oracle_to_raster(d->viewport, cp, fb[0], fr[0]);
    oracle_to_raster(d->viewport, cp + 4, fb[1], fr[1]);
    bres_t b;
    bres_init(&b, fb[0], fb[1]);
    uint32_t xy[2];
    float t;
    int nvar = v[0].nvar;
    while (bres_next(&b, xy, &t)) {
        if (!in_scissor(d, xy[0], xy[1])) continue;
        p->fragments++;
        if (p->frag_count) p->frag_count[(uint64_t)xy[1] * p->width + xy[0]]++;
        float s = 1.0f - t;
        float l0 = al[0] * s;
        float l1 = al[1] * t;
        float L = l0 + l1;
        float m = 1.0f - L;
        float za = fr[0][2] * m, zb = fr[1][2] * L;
        float z = za + zb;
        float var[OR_MAX_VAR];
        if (d->interp == OR_INTERP_FLAT) {
            for (int c = 0; c < nvar; ++c) var[c] = v[0].var[c];
        } else if (d->interp == OR_INTERP_PERSPECTIVE) {
            /* extension: perspective-correct weights q_i = w_i(L) * (1/w_i) */
            float q0 = m * (1.0f / v[0].clip[3]);
            float q1 = L * (1.0f / v[1].clip[3]);
            float qs = q0 + q1;
            for (int c = 0; c < nvar; ++c) {
                float acc = v[0].var[c] * q0;
                acc = acc + v[1].var[c] * q1;
                var[c] = acc / qs;
            }
        } else {
            for (int c = 0; c < nvar; ++c) {
                float a = v[0].var[c] * m;
                float bb = v[1].var[c] * L;
                var[c] = a + bb;
            }
        }
        write_fragment(p, d, xy[0], xy[1], z, var, nvar, prim);
    }
    return 0;
}

static int draw_triangle(OrPass *p, const OrDraw *d, const vout vo[3], uint32_t prim)
{
    uint64_t sx0 = d->scissor[0], sy0 = d->scissor[1];
    uint64_t sx1 = sx0 + d->scissor[2], sy1 = sy0 + d->scissor[3];
    /* TriFace::new (primitive.rs:169-184) */
    float nd[3][3];
    for (int k = 0; k < 3; ++k) {
        if (vo[k].clip[3] == 0.0f) return fail(p, "clip position w = 0");
        for (int c = 0; c < 3; ++c) nd[k][c] = vo[k].clip[c] / vo[k].clip[3];
    }
    float abx = nd[1][0] - nd[0][0], aby = nd[1][1] - nd[0][1];
    float acx = nd[2][0] - nd[0][0], acy = nd[2][1] - nd[0][1];
    float area_z = abx * acy - aby * acx;
    int ccw = area_z > 0.0f;
    int front_facing = ccw ? (d->front_face_cw == 0) : (d->front_face_cw == 1);
    int cull = (d->cull_mode == OR_CULL_FRONT) ? front_facing
             : (d->cull_mode == OR_CULL_BACK) ? !front_facing : 0;
    if (cull) return 0;
    (void)front_facing; /* the built-in programs do not read front_facing */

    float pos[3][4];
    for (int k = 0; k < 3; ++k) memcpy(pos[k], vo[k].clip, sizeof pos[k]);
    ctri sub[CLIP_CAP];
    int nsub = clip_triangle(pos, sub);
    for (int t = 0; t < nsub; ++t) {
        uint32_t fb[3][2];
        float fr[3][4];
        for (int k = 0; k < 3; ++k) oracle_to_raster(d->viewport, sub[t].v[k].p, fb[k], fr[k]);
        float total = shoelace(fr[0][0], fr[0][1], fr[1][0], fr[1][1], fr[2][0], fr[2][1]);
        if (total == 0.0f) continue;
        scanlines_t sl;
        scanline_t line;
        scanlines_init(&sl, fb);
        for (;;) {
            int r = scanlines_next(&sl, &line);
            if (r < 0) return fail(p, "scanline produced a negative coordinate");
            if (r == 0) break;
            for (int64_t xx = line.x1; xx <= line.x2; ++xx) {
                uint64_t x = (uint64_t)xx, y = (uint64_t)line.y;
                if (!(x >= sx0 && x < sx1 && y >= sy0 && y < sy1)) continue;
                p->fragments++;
                if (p->frag_count) p->frag_count[y * p->width + x]++;
                /* barycentric at the pixel corner (raster.rs:313-335) */
                float px = (float)x, py = (float)y;
                float l0 = shoelace(px, py, fr[1][0], fr[1][1], fr[2][0], fr[2][1]) / total;
                float l1 = shoelace(fr[0][0], fr[0][1], px, py, fr[2][0], fr[2][1]) / total;
                float l2 = shoelace(fr[0][0], fr[0][1], fr[1][0], fr[1][1], px, py) / total;
                /* compose with the clip barycentrics (raster.rs:338-343) */
                float lc[3];
                for (int c = 0; c < 3; ++c) {
                    float acc = sub[t].v[0].b[c] * l0;
                    acc = acc + sub[t].v[1].b[c] * l1;
                    acc = acc + sub[t].v[2].b[c] * l2;
                    lc[c] = acc;
                }
                /* fragment = lc . [a, b, c].fragment (raster.rs:345-346) */
                float frag_z = fr[0][2] * lc[0];
                frag_z = frag_z + fr[1][2] * lc[1];
                frag_z = frag_z + fr[2][2] * lc[2];
                /* varyings of the unclipped vertices (fragment.rs:150-152, 346-366) */
                float var[OR_MAX_VAR];
                int nvar = vo[0].nvar;
                if (d->interp == OR_INTERP_FLAT) {
                    for (int c = 0; c < nvar; ++c) var[c] = vo[0].var[c];
                } else if (d->interp == OR_INTERP_PERSPECTIVE) {
                    /* extension: perspective-correct weights q_i = lc_i * (1/w_i) */
                    float q0 = lc[0] * (1.0f / vo[0].clip[3]);
                    float q1 = lc[1] * (1.0f / vo[1].clip[3]);
                    float q2 = lc[2] * (1.0f / vo[2].clip[3]);
                    float qs = (q0 + q1) + q2;
                    for (int c = 0; c < nvar; ++c) {
                        float acc = vo[0].var[c] * q0;
                        acc = acc + vo[1].var[c] * q1;
                        acc = acc + vo[2].var[c] * q2;
                        var[c] = acc / qs;
                    }
                } else {
                    for (int c = 0; c < nvar; ++c) {
                        float acc = vo[0].var[c] * lc[0];
                        acc = acc + vo[1].var[c] * lc[1];
                        acc = acc + vo[2].var[c] * lc[2];
                        var[c] = acc;
                    }
                }
                write_fragment(p, d, (uint32_t)x, (uint32_t)y, frag_z, var, nvar, prim);
            }
        }
    }
    return 0;
}

static int draw(OrPass *p, const OrDraw *d)
{
    if (d->program == OR_PROG_MVP_TEXTURE && !d->texture) return fail(p, "no texture bound");
    /* ToRaster scissor box (raster.rs:130-143) */
    uint64_t sx1 = (uint64_t)d->scissor[0] + d->scissor[2], sy1 = (uint64_t)d->scissor[1] + d->scissor[3];
    if (sx1 > p->width || sy1 > p->height) return fail(p, "scissor outside the framebuffer");
    int n = topo_vertices(d->topology);
    int strip = d->topology == OR_TOPO_LINE_STRIP || d->topology == OR_TOPO_TRIANGLE_STRIP;
    if (d->strip_restart && !strip) return fail(p, "strip_index_format set, but not a strip topology");
    /* the separators are raw index values (index.rs:72-106): read them in range */
    const uint8_t *raw = NULL;
    if (d->indexed && d->strip_restart) {
        uint64_t end = ((uint64_t)d->first + d->count) * d->index_format;
        if (end > d->index_size) return fail(p, "index out of range or vertex index overflow");
        raw = d->index_data + (uint64_t)d->first * d->index_format;
    }
    int64_t cap = (int64_t)d->count + 3;
    uint32_t *pos = (uint32_t *)malloc(sizeof(uint32_t) * 3 * (size_t)cap);
    if (!pos) return fail(p, "out of memory");
    int64_t nprim = oracle_assemble(d->topology, raw != NULL, d->index_format, raw, d->count, pos, cap);
    int rc = 0;
    for (uint32_t inst = d->first_instance; inst < d->first_instance + d->instance_count && !rc; ++inst) {
        for (int64_t prim = 0; prim < nprim && !rc; ++prim) {
            vout vo[3];
            for (int k = 0; k < n && !rc; ++k) {
                uint32_t vertex;
                if (resolve_index(d, d->first + pos[prim * n + k], &vertex))
                    rc = fail(p, "index out of range or vertex index overflow");
                else if (run_vs(d, inst, vertex, &vo[k]))
                    rc = fail(p, "vertex fetch out of range");
            }
            if (rc) break;
            /* primitive_index: the primitive's place in its instance's
             * assembly order, culled ones counted (state.rs:536-541) */
            if (n == 1) rc = draw_point(p, d, &vo[0], (uint32_t)prim);
            else if (n == 2) rc = draw_line(p, d, vo, (uint32_t)prim);
            else rc = draw_triangle(p, d, vo, (uint32_t)prim);
        }
    }
    free(pos);
    return rc;
}

/* RenderPassCommand::execute (render_pass/mod.rs:345-394) */
int oracle_execute_pass(OrPass *p)
{
    p->error = 0;
    p->message[0] = 0;
    p->fragments = 0;
    uint64_t npx = (uint64_t)p->width * p->height;
    /* State::load (state.rs:135-145) */
    if (p->color && p->color_clear) {
        float c[4];
        for (int k = 0; k < 4; ++k) c[k] = (float)p->clear_color[k]; /* f64 -> f32 cast */
        uint32_t w = p->width;
        for (uint64_t i = 0; i < npx; ++i) put_pixel(p, (uint32_t)(i % w), (uint32_t)(i / w), c);
    }
    if (p->depth && p->depth_clear)
        for (uint64_t i = 0; i < npx; ++i) p->depth[i] = p->clear_depth;
    for (uint32_t i = 0; i < p->n_draws; ++i)
        if (draw(p, &p->draws[i])) return -1;
    return 0;
}
