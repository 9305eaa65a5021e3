// Fused Shiu-style LIF update for the plain (B, N) fly-major state, one lane per (fly, neuron).
// Semantics: flyverse/brain.py:838-868 (operation order) and the flag word of flyverse/kernels/neural.cu:11-52.
// Built with -ffp-contract=off: every product and sum rounds on its own, like the reference's CPU torch ops.
#include "common.h"

namespace {

struct LifArgs {
    float *v, *g, *drive, *adapt, *refrac, *active, *prob, *rnd;
    float *res, *std_u, *spikes, *out, *rate, *counts;
    const float *p, *phase;
};

__device__ __forceinline__ bool lif_one(LifArgs s, int n, int64_t t, int i, int fly, int flags, uint64_t seed,
                                        const uint32_t* step_base, uint32_t step_offset);

// flags: 1 Poisson, 2 STD, 4 record counts, 8 adaptation, 16 phase coefficients (8, N), 32 frozen-at-rest skip,
// 64 draw the Poisson uniform in-kernel (Philox keyed by seed, step, neuron, fly) instead of reading rnd.
__global__ __launch_bounds__(256) void k_lif(LifArgs s, int n, int64_t total, int flags, uint64_t seed,
                                             const uint32_t* step_base, uint32_t step_offset, uint64_t* bits) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int fly = blockIdx.y;
    const int64_t t = (int64_t)fly * n + i;
    // packed spike bits (one 64-bit word per 64 cells of a fly): every lane of the wavefront takes part in the ballot
    bool fired_out = false;
    if (i < n) {
        fired_out = lif_one(s, n, t, i, fly, flags, seed, step_base, step_offset);
    }
    if (bits) {
        const uint64_t word = __builtin_amdgcn_ballot_w64(fired_out);
        if ((threadIdx.x & 63) == 0 && i < n) bits[(int64_t)fly * ((n + 63) / 64) + (i >> 6)] = word;
    }
}

__device__ __forceinline__ bool lif_one(LifArgs s, int n, int64_t t, int i, int fly, int flags, uint64_t seed,
                                        const uint32_t* step_base, uint32_t step_offset) {
    uint32_t step = step_offset + (step_base ? *step_base : 0u);
    const float* P = s.p;
    if ((flags & 32) && s.active[i] == 0.0f && (!(flags & 1) || s.prob[t] == 0.0f) && s.v[t] == P[0] &&
        s.g[t] == 0.0f && s.drive[t] == 0.0f && s.adapt[t] == 0.0f && s.refrac[t] == 0.0f && s.rate[t] == 0.0f &&
        s.spikes[t] == 0.0f && s.out[t] == 0.0f && s.res[t] == 1.0f)
        return false;
    float am = P[3], dt = P[4], aad = P[7], ast = P[8], ar = P[9], rg = P[10], u = 1.0f, pois = 1.0f;
    if (flags & 16) {
        const float* ph = s.phase;
        am = ph[i]; dt = ph[n + i]; aad = ph[2 * n + i]; ast = ph[3 * n + i];
        ar = ph[4 * n + i]; rg = ph[5 * n + i]; u = ph[6 * n + i]; pois = ph[7 * n + i];
    }
    float target = P[0] + s.g[t];
    target = target + s.drive[t];
    target = target - s.adapt[t];
    float dv = s.v[t] - target;
    dv = dv * am;
    float v = target + dv;
    float ref = s.refrac[t];
    if (ref > 0.0f) v = P[1];
    ref = ref - dt;
    if (ref < 0.0f) ref = 0.0f;
    float spike = (v >= P[2]) ? 1.0f : 0.0f;
    spike = spike * u;
    spike = spike * s.active[i];
    if (flags & 1) {
        float pp = s.prob[t] * pois;
        if (flags & 64) {
            // a Philox uniform lies in [0, 1), so it is never below a probability <= 0: draw only where it can fire
            if (pp > 0.0f && fh_uniform(seed, step, (uint32_t)i, (uint32_t)fly, 0u) < pp) spike = fmaxf(spike, 1.0f);
        } else if (s.rnd[t] < pp) {
            spike = fmaxf(spike, 1.0f);
        }
    }
    bool fired = spike > 0.0f;
    if (fired) { v = P[1]; ref = P[5]; }
    s.v[t] = v; s.refrac[t] = ref; s.spikes[t] = spike;
    if (flags & 8) {
        float a = s.adapt[t] * aad;
        float j = spike * P[6];
        s.adapt[t] = a + j;
    }
    if (flags & 2) {
        float r = s.res[t];
        s.out[t] = spike * r;
        if (fired) { float keep = 1.0f - s.std_u[i]; r = r * keep; }
        float d = 1.0f - r;
        d = d * ast;
        s.res[t] = 1.0f - d;
    } else {
        s.out[t] = spike;
    }
    if (flags & 4) s.counts[t] = s.counts[t] + spike;
    float rr = s.rate[t] * ar;
    float rs = spike * rg;
    s.rate[t] = rr + rs;
    return fired;
}

int lif_launch(void** p, int n, int batch, int flags, uint64_t seed, const uint32_t* step_base, uint32_t step,
               uint64_t* bits, void* stream) {
    if (n < 0 || batch < 0) return FH_ERR_ARGS;
    int64_t total = (int64_t)n * batch;
    if (total == 0) return 0;
    LifArgs s{(float*)p[0], (float*)p[1], (float*)p[2], (float*)p[3], (float*)p[4], (float*)p[5], (float*)p[6],
              (float*)p[7], (float*)p[8], (float*)p[9], (float*)p[10], (float*)p[11], (float*)p[12], (float*)p[13],
              (const float*)p[14], (const float*)p[15]};
    if (batch > 65535) return FH_ERR_SIZE;
    hipLaunchKernelGGL(k_lif, dim3((unsigned)((n + 255) / 256), (unsigned)batch), dim3(256), 0, (hipStream_t)stream, s, n,
                       total, flags, seed, step_base, step, bits);
    FH_LAUNCH_CHECK();
}

__global__ __launch_bounds__(256) void k_uniform(float* out, int n, int64_t total, uint64_t seed,
                                                 const uint32_t* step_base, uint32_t step_offset, uint32_t stream_tag) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    uint32_t step = step_offset + (step_base ? *step_base : 0u);
    out[t] = fh_uniform(seed, step, (uint32_t)(t % n), (uint32_t)(t / n), stream_tag);
}

__global__ __launch_bounds__(256) void k_scale(float* x, int64_t total, float a) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < total) x[t] = x[t] * a;
}

}  // namespace

// Same contract as flyverse/kernels/neural.cu:54 (launch_lif): 16 pointers
// [v, g, drive, adapt, refrac, active, prob, rnd, res, std_u, spikes, out, rate, counts, P, phase].
FH_EXPORT int launch_lif(void** p, int n, int batch, int flags, void* stream) {
    return lif_launch(p, n, batch, flags & 63, 0, nullptr, 0, nullptr, stream);
}

// launch_lif with the Poisson uniform drawn in-kernel from Philox(seed, step, neuron, fly); rnd is not read.
// p has a 17th pointer: a device uint32 added to `step` (null = 0), so a captured graph can advance the step.
FH_EXPORT int launch_lif_rng(void** p, int n, int batch, int flags, uint64_t seed, uint32_t step, void* stream) {
    return lif_launch(p, n, batch, (flags & 63) | 64, seed, (const uint32_t*)p[16], step, nullptr, stream);
}

// launch_lif_rng that also writes the transmitted spikes as packed bits: p[17] = uint64 words (batch,
// ceil(n / 64)), bit i % 64 of word i / 64 set when cell i fired.  For the plain pull over bits; valid when the
// transmitted value is the 0/1 spike (no short-term depression).
FH_EXPORT int launch_lif_bits(void** p, int n, int batch, int flags, uint64_t seed, uint32_t step, void* stream) {
    if (flags & 2) return FH_ERR_ARGS;
    if (!p[17]) return FH_ERR_ARGS;
    return lif_launch(p, n, batch, (flags & 63) | 64, seed, (const uint32_t*)p[16], step, (uint64_t*)p[17], stream);
}

// out[b, i] = Philox uniform (seed, *step_base + step, i, b, stream_tag): the draw torch.rand makes at
// flyverse/brain.py:877 (stream_tag 0); step_base may be null.
FH_EXPORT int launch_uniform(float* out, int n, int batch, uint64_t seed, const uint32_t* step_base, uint32_t step,
                             uint32_t stream_tag, void* stream) {
    if (n < 0 || batch < 0) return FH_ERR_ARGS;
    int64_t total = (int64_t)n * batch;
    if (total == 0) return 0;
    hipLaunchKernelGGL(k_uniform, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, (hipStream_t)stream, out, n,
                       total, seed, step_base, step, stream_tag);
    FH_LAUNCH_CHECK();
}

// x *= a over `count` floats (g.mul_(a_s), flyverse/brain.py:832).
FH_EXPORT int launch_scale(float* x, int64_t count, float a, void* stream) {
    if (count < 0) return FH_ERR_ARGS;
    if (count == 0) return 0;
    hipLaunchKernelGGL(k_scale, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, (hipStream_t)stream, x, count, a);
    FH_LAUNCH_CHECK();
}
