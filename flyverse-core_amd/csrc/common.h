// Shared helpers for the flyhip kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

#define FH_EXPORT extern "C" __attribute__((visibility("default")))

// Return codes of every entry point: 0 = ok, a positive value is the hipError_t of the failing call,
// negative values are argument errors detected on the host before any launch.
enum { FH_ERR_ARGS = -1, FH_ERR_SIZE = -2 };

#define FH_LAUNCH_CHECK() do { hipError_t _e = hipGetLastError(); return (int)_e; } while (0)

static constexpr int kWave = 64;

// ------------------------------------------------------------------ Philox4x32-10 (Salmon et al., SC'11)
// Counter (neuron, step, fly, stream), key (seed low, seed high); the first output word gives the uniform.
__device__ __forceinline__ uint32_t fh_philox_first(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                                    uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    return c0;
}

// Uniform in [0, 1): top 24 bits times 2^-24, exact in fp32.
__device__ __forceinline__ float fh_uniform(uint64_t seed, uint32_t step, uint32_t neuron, uint32_t fly,
                                            uint32_t stream) {
    uint32_t x = fh_philox_first(neuron, step, fly, stream, (uint32_t)seed, (uint32_t)(seed >> 32));
    return (float)(x >> 8) * 5.9604644775390625e-08f;
}

__device__ __forceinline__ float fh_clip01(float x) { return x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x); }

__device__ __forceinline__ float fh_readlane(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

__device__ __forceinline__ int fh_lane() { return __lane_id(); }
