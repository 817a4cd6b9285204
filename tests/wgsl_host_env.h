// wgsl_host_env.h -- host-side environment for running the WGSL front end's
// emitted functions on the CPU in the translator's known-answer tests (the
// render kernels provide the device-side environment).
#pragma once

#include <stdint.h>
#include <string.h>

#include <cmath>

#define WGSL_FN static inline
#include "wgsl_prelude.h"

struct HostEnv {
    uint32_t vi = 0, ii = 0, pi = 0;
    bool ff = true;
    float frag_pos[4] = {0.0f, 0.0f, 0.0f, 1.0f};
    uint32_t in_loc[16][4] = {};
    // outputs
    float pos[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    uint32_t loc[16][4] = {};
    uint32_t loc_n[16] = {};
    float depth = 0.0f;
    // buffers by resource slot
    const uint8_t *bufs[8] = {};
    uint64_t sizes[8] = {};

    uint32_t vertex_index() const { return vi; }
    uint32_t instance_index() const { return ii; }
    uint32_t primitive_index() const { return pi; }
    bool front_facing() const { return ff; }
    void position(float *p) const { memcpy(p, frag_pos, sizeof frag_pos); }
    void in_location(uint32_t l, uint32_t *w, uint32_t n) const { memcpy(w, in_loc[l], 4 * n); }
    void out_position(const float *p) { memcpy(pos, p, sizeof pos); }
    void out_location(uint32_t l, const uint32_t *w, uint32_t n)
    {
        memcpy(loc[l], w, 4 * n);
        loc_n[l] = n;
    }
    void out_frag_depth(float d) { depth = d; }
    const uint8_t *buffer(uint32_t s) const { return bufs[s]; }
    uint64_t buffer_size(uint32_t s) const { return sizes[s]; }
    void sample(uint32_t, uint32_t, float, float, float *out) const
    {
        for (int k = 0; k < 4; ++k) out[k] = 0.0f;
    }
};
