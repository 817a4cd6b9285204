// wgc_launch.h -- host-side interface between the runtime and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace wgc {

struct DrawParams;

// Brackets launches with HIP events on the launching stream; the runtime
// resolves the pairs once the stream has passed them.
struct KernelTimer {
    struct Pending {
        std::string name;
        hipEvent_t begin, end;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;

    hipEvent_t get()
    {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        (void)hipEventCreate(&e);
        return e;
    }
    void begin(const char *name, hipStream_t st)
    {
        Pending p{name, get(), nullptr};
        (void)hipEventRecord(p.begin, st);
        pending.push_back(p);
    }
    void end(hipStream_t st)
    {
        pending.back().end = get();
        (void)hipEventRecord(pending.back().end, st);
    }
};

hipError_t launch_draw(const DrawParams &d, hipStream_t st, KernelTimer *kt);
hipError_t launch_reset(unsigned long long *frags, hipStream_t st);
hipError_t launch_clear(uint32_t *color, uint32_t rgba, float *depth, float dv, uint64_t n,
                        hipStream_t st, KernelTimer *kt);

}  // namespace wgc
