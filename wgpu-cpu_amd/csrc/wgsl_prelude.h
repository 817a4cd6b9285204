// wgsl_prelude.h -- value types and scalar operations that the WGSL
// translator's emitted code is written against.
//
// The emitted functions are compiled by hipRTC together with the render
// kernels (WGSL_FN = __device__ inline) and, for the translator's own
// known-answer tests, by a host C++ compiler (WGSL_FN = static inline).
// Every operation has the semantics of the reference's shader compiler
// (naga-cranelift/src/expression/*.rs): IEEE-754 binary32 arithmetic
// without fused multiply-add, wrapping integer arithmetic, shift amounts
// taken modulo 32, and checked integer division.
#pragma once

#ifndef WGSL_FN
#define WGSL_FN __device__ __forceinline__
#endif

namespace wgsl {

template <class T, int N>
struct vec {
    T v[N];
};

// Column-major float matrix: C columns of R rows.
template <int C, int R>
struct mat {
    vec<float, R> c[C];
};

template <class T, int N>
struct arr {
    T v[N];
};

// Abort codes of an invocation (naga-cranelift/src/function.rs:60-68).
enum : uint32_t {
    OK = 0u,
    ABORT_KILL = 3u,       // discard
    ABORT_OOB = 4u,        // index outside an array
    ABORT_DIV0 = 5u,       // integer division or remainder by zero
    ABORT_OVERFLOW = 6u,   // i32::MIN / -1, or a float -> int conversion out of range
};

// Wrapping i32 arithmetic (cranelift iadd / isub / imul / ineg).
WGSL_FN int32_t iadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
WGSL_FN int32_t isub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }
WGSL_FN int32_t imul(int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); }
WGSL_FN int32_t ineg(int32_t a) { return (int32_t)(0u - (uint32_t)a); }

// Shifts: the amount is taken modulo the bit width (cranelift ishl / ushr /
// sshr); the signed right shift is arithmetic.
WGSL_FN int32_t ishl(int32_t a, uint32_t b) { return (int32_t)((uint32_t)a << (b & 31u)); }
WGSL_FN uint32_t ushl(uint32_t a, uint32_t b) { return a << (b & 31u); }
WGSL_FN uint32_t ushr(uint32_t a, uint32_t b) { return a >> (b & 31u); }
WGSL_FN int32_t sshr(int32_t a, uint32_t b)
{
    uint32_t s = b & 31u;
    uint32_t u = (uint32_t)a >> s;
    if (a < 0 && s) u |= ~(0xFFFFFFFFu >> s);
    return (int32_t)u;
}

// Checked division (binary.rs:451-531): zero divisors abort; i32::MIN / -1
// aborts with overflow; the remainder i32::MIN % -1 is 0 (cranelift srem).
WGSL_FN uint32_t sdiv(int32_t a, int32_t b, int32_t &r)
{
    if (a == (int32_t)0x80000000 && b == -1) return ABORT_OVERFLOW;
    if (b == 0) return ABORT_DIV0;
    r = a / b;
    return OK;
}
WGSL_FN uint32_t srem(int32_t a, int32_t b, int32_t &r)
{
    if (b == 0) return ABORT_DIV0;
    r = (b == -1) ? 0 : a % b;
    return OK;
}
WGSL_FN uint32_t udiv(uint32_t a, uint32_t b, uint32_t &r)
{
    if (b == 0u) return ABORT_DIV0;
    r = a / b;
    return OK;
}
WGSL_FN uint32_t urem(uint32_t a, uint32_t b, uint32_t &r)
{
    if (b == 0u) return ABORT_DIV0;
    r = a % b;
    return OK;
}

// Float remainder e1 - e2 * trunc(e1 / e2) (binary.rs:436-447).
WGSL_FN float frem(float a, float b)
{
    float q = __builtin_truncf(a / b);
    float m = b * q;
    return a - m;
}

// Conversions (as.rs:61-172).  Float to integer traps in the reference
// (cranelift fcvt_to_sint / fcvt_to_uint) for NaN and out-of-range values;
// here that is the overflow abort.
WGSL_FN uint32_t f2i(float f, int32_t &r)
{
    if (!(f >= -2147483648.0f && f < 2147483648.0f)) return ABORT_OVERFLOW;
    r = (int32_t)f;
    return OK;
}
WGSL_FN uint32_t f2u(float f, uint32_t &r)
{
    if (!(f > -1.0f && f < 4294967296.0f)) return ABORT_OVERFLOW;
    r = (f <= 0.0f) ? 0u : (uint32_t)f;
    return OK;
}
// float -> bool compares against -0.0 with an unordered not-equal: NaN is true.
WGSL_FN bool f2b(float f) { return f != -0.0f; }

WGSL_FN uint32_t f32_bits(float f) { return __builtin_bit_cast(uint32_t, f); }
WGSL_FN float bits_f32(uint32_t u) { return __builtin_bit_cast(float, u); }

// Little-endian loads from host-shareable memory (uniform / storage
// buffers) at 4-byte aligned offsets.
WGSL_FN uint32_t ld_u32(const uint8_t *p) { return *reinterpret_cast<const uint32_t *>(p); }
WGSL_FN int32_t ld_i32(const uint8_t *p) { return (int32_t)ld_u32(p); }
WGSL_FN float ld_f32(const uint8_t *p) { return bits_f32(ld_u32(p)); }
WGSL_FN bool ld_bool(const uint8_t *p) { return ld_u32(p) != 0u; }

// Scalar <-> 32-bit word (inter-stage variables and vertex attributes are
// passed as raw words, vertex.rs:139-152).
WGSL_FN uint32_t to_word(float f) { return f32_bits(f); }
WGSL_FN uint32_t to_word(int32_t i) { return (uint32_t)i; }
WGSL_FN uint32_t to_word(uint32_t u) { return u; }
WGSL_FN uint32_t to_word(bool b) { return b ? 1u : 0u; }
WGSL_FN void from_word(uint32_t w, float &f) { f = bits_f32(w); }
WGSL_FN void from_word(uint32_t w, int32_t &i) { i = (int32_t)w; }
WGSL_FN void from_word(uint32_t w, uint32_t &u) { u = w; }
WGSL_FN void from_word(uint32_t w, bool &b) { b = w != 0u; }

}  // namespace wgsl
