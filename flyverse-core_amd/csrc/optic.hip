// Elementwise optic-lobe kernels for the plain (B, n) fly-major layout, and the per-frame coupling kernels.
// Semantics: the substep of flyverse/optic.py:520-551 (native spelling flyverse/kernels/neural.cu:62-88), the
// photoreceptor stage optic.py:496-517, the feedback optic.py:572 and the output drive optic.py:599-602.
#include "common.h"

namespace {

__global__ __launch_bounds__(256) void k_optic_dr(const float* __restrict__ v, const float* __restrict__ b,
                                                  float* __restrict__ dr, int n, int64_t total) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    float base = b[t % n];
    float s = v[t] + base;
    dr[t] = fh_clip01(s) - base;
}

struct OpticArgs { float *v, *adapt, *dr; const float *y, *pr, *spk, *a, *b; };

__global__ __launch_bounds__(256) void k_optic_update(OpticArgs s, int n, int64_t total, float gr, float ga,
                                                      float aad) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    int i = (int)(t % n);
    float old = s.dr[t];
    float inp = gr * s.y[t];
    inp = inp + s.pr[t];
    float gad = ga * s.adapt[t];
    inp = inp - gad;
    inp = inp + s.spk[t];
    float d = s.v[t] - inp;
    d = d * s.a[i];
    float v = inp + d;
    s.v[t] = v;
    float da = s.adapt[t] - old;
    da = da * aad;
    s.adapt[t] = old + da;
    float q = v + s.b[i];
    s.dr[t] = fh_clip01(q) - s.b[i];
}

// Photoreceptor stage with one (column, family) cell per photoreceptor.  rad (B, n_pr, 4), sens (n_pr, 4),
// I_lp / I_mean (B, n_pr), fresh (B) uint8 flags, coefs = [a_lp, 1-a_lp, a_ad, 1-a_ad, eps, contrast_clip].
// fresh is cleared by the caller after the launch (launch_photoreceptor does it in a second tiny kernel).
__global__ __launch_bounds__(256) void k_photoreceptor(const float* __restrict__ rad, const float* __restrict__ sens,
                                                       float* __restrict__ I_lp, float* __restrict__ I_mean,
                                                       const uint8_t* __restrict__ fresh, const float* __restrict__ cf,
                                                       float* __restrict__ a_pr, int n_pr, int64_t total) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    int p = (int)(t % n_pr), b = (int)(t / n_pr);
    float I = 0.0f;
#pragma unroll
    for (int c = 0; c < 4; ++c) { float q = rad[t * 4 + c] * sens[p * 4 + c]; I = I + q; }
    bool fr = fresh[b] != 0;
    float lp = fr ? I : I_lp[t];
    float mean = fr ? I : I_mean[t];
    float l1 = cf[0] * lp, l2 = cf[1] * I;
    lp = l1 + l2;
    float m1 = cf[2] * mean, m2 = cf[3] * lp;
    mean = m1 + m2;
    I_lp[t] = lp; I_mean[t] = mean;
    float num = lp - mean, den = mean + cf[4];
    float c = num / den;
    c = c < -1.0f ? -1.0f : c;
    c = c > cf[5] ? cf[5] : c;
    a_pr[t] = c;
}

__global__ void k_clear_u8(uint8_t* f, int n) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) f[t] = 0;
}

// s[b, k] = clamp(rate[b, spk_idx[k]] / 100, 0, 3) (true division, as the CPU reference computes it).
__global__ __launch_bounds__(256) void k_feedback(const float* __restrict__ rate, const int* __restrict__ spk_idx,
                                                  float* __restrict__ s, int n, int n_spk, int64_t total) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    int k = (int)(t % n_spk), b = (int)(t / n_spk);
    float q = rate[(int64_t)b * n + spk_idx[k]] / 100.0f;
    q = q < 0.0f ? 0.0f : q;
    q = q > 3.0f ? 3.0f : q;
    s[t] = q;
}

// drive[b, spk_idx[k]] = clamp(gain_out * y_sr[b, k], -clip, clip); the caller zeroes drive's other entries once.
__global__ __launch_bounds__(256) void k_drive_out(const float* __restrict__ y_sr, const int* __restrict__ spk_idx,
                                                   float* __restrict__ drive, int n, int n_spk, float gain_out,
                                                   float clip, int64_t total) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    int k = (int)(t % n_spk), b = (int)(t / n_spk);
    float d = gain_out * y_sr[t];
    d = d < -clip ? -clip : d;
    d = d > clip ? clip : d;
    drive[(int64_t)b * n + spk_idx[k]] = d;
}

__global__ __launch_bounds__(128) void k_means(const int* __restrict__ ptr, const int* __restrict__ idx,
                                               const float* __restrict__ rates, double* __restrict__ out, int n,
                                               int groups, int batch, int stride) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= groups * batch) return;
    int b = t / groups, k = t % groups;
    double sum = 0.0;
    for (int j = ptr[k]; j < ptr[k + 1]; ++j) sum += (double)rates[(int64_t)b * n + idx[j]];
    int cnt = ptr[k + 1] - ptr[k];
    out[(int64_t)b * stride + k] = cnt ? sum / cnt : 0.0;
}

inline unsigned nblk(int64_t total) { return (unsigned)((total + 255) / 256); }

}  // namespace

// Same contract as flyverse/kernels/neural.cu:80 (launch_optic_dr): p = [v, b, dr].
FH_EXPORT int launch_optic_dr(void** p, int n, int batch, void* stream) {
    if (n < 0 || batch < 0) return FH_ERR_ARGS;
    int64_t total = (int64_t)n * batch;
    if (!total) return 0;
    hipLaunchKernelGGL(k_optic_dr, dim3(nblk(total)), dim3(256), 0, (hipStream_t)stream, (const float*)p[0],
                       (const float*)p[1], (float*)p[2], n, total);
    FH_LAUNCH_CHECK();
}

// Same contract as flyverse/kernels/neural.cu:84 (launch_optic): p = [v, adapt, dr, y, pr, spk, a, b].
FH_EXPORT int launch_optic(void** p, int n, int batch, float gr, float ga, float aad, void* stream) {
    if (n < 0 || batch < 0) return FH_ERR_ARGS;
    int64_t total = (int64_t)n * batch;
    if (!total) return 0;
    OpticArgs s{(float*)p[0], (float*)p[1], (float*)p[2], (const float*)p[3], (const float*)p[4],
                (const float*)p[5], (const float*)p[6], (const float*)p[7]};
    hipLaunchKernelGGL(k_optic_update, dim3(nblk(total)), dim3(256), 0, (hipStream_t)stream, s, n, total, gr, ga, aad);
    FH_LAUNCH_CHECK();
}

// Photoreceptor stage: p = [rad (B, n_pr, 4), sens (n_pr, 4), I_lp, I_mean, fresh uint8 (B), coefs (6), a_pr].
FH_EXPORT int launch_photoreceptor(void** p, int n_pr, int batch, void* stream) {
    if (n_pr < 0 || batch < 0) return FH_ERR_ARGS;
    int64_t total = (int64_t)n_pr * batch;
    if (!total) return 0;
    hipLaunchKernelGGL(k_photoreceptor, dim3(nblk(total)), dim3(256), 0, (hipStream_t)stream, (const float*)p[0],
                       (const float*)p[1], (float*)p[2], (float*)p[3], (const uint8_t*)p[4], (const float*)p[5],
                       (float*)p[6], n_pr, total);
    hipLaunchKernelGGL(k_clear_u8, dim3((batch + 255) / 256), dim3(256), 0, (hipStream_t)stream, (uint8_t*)p[4], batch);
    FH_LAUNCH_CHECK();
}

// Spiking feedback into the optic lobe: p = [rate (B, n), spk_idx int32 (n_spk), s (B, n_spk)].
FH_EXPORT int launch_feedback(void** p, int n, int n_spk, int batch, void* stream) {
    if (n < 0 || n_spk < 0 || batch < 0) return FH_ERR_ARGS;
    int64_t total = (int64_t)n_spk * batch;
    if (!total) return 0;
    hipLaunchKernelGGL(k_feedback, dim3(nblk(total)), dim3(256), 0, (hipStream_t)stream, (const float*)p[0],
                       (const int*)p[1], (float*)p[2], n, n_spk, total);
    FH_LAUNCH_CHECK();
}

// Optic output into the LIF drive: p = [y_sr (B, n_spk), spk_idx int32 (n_spk), drive (B, n)].
FH_EXPORT int launch_drive_out(void** p, int n, int n_spk, int batch, float gain_out, float clip, void* stream) {
    if (n < 0 || n_spk < 0 || batch < 0) return FH_ERR_ARGS;
    int64_t total = (int64_t)n_spk * batch;
    if (!total) return 0;
    hipLaunchKernelGGL(k_drive_out, dim3(nblk(total)), dim3(256), 0, (hipStream_t)stream, (const float*)p[0],
                       (const int*)p[1], (float*)p[2], n, n_spk, gain_out, clip, total);
    FH_LAUNCH_CHECK();
}

// Same contract as flyverse/kernels/neural.cu:141 (launch_means): p = [ptr, idx, rates (B, n), out double (B, stride)].
FH_EXPORT int launch_means(void** p, int n, int groups, int batch, int stride, void* stream) {
    if (n < 0 || groups < 0 || batch < 0) return FH_ERR_ARGS;
    if (!(groups * batch)) return 0;
    hipLaunchKernelGGL(k_means, dim3((groups * batch + 127) / 128), dim3(128), 0, (hipStream_t)stream,
                       (const int*)p[0], (const int*)p[1], (const float*)p[2], (double*)p[3], n, groups, batch, stride);
    FH_LAUNCH_CHECK();
}
