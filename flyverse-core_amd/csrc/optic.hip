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

// Photoreceptor stage over (column, family) cells (flyverse/optic.py:496-517), one thread per (fly b, cell k),
// cell = column * 5 + family.  The cell's input is its row of the averaging matrix (optic.py:320-325: weights
// cell_w = 1/count, photoreceptors ascending) applied to the intensities I_p = sum_c rad[b, column(p), c] * sens[p, c]
// (channel order, from +0), in the CSR products' order: chunks of 16 folded from +0, chunk sums folded from +0.
// Then the low-pass and adaptation state of the cell, its contrast (0 for a cell without photoreceptors,
// optic.py:513-514), and that contrast as the activity of each of its photoreceptors (optic.py:517).
// rad (B, n_col, 4); I_lp, I_mean, contrast (B, n_cells); a_pr (B, n_pr); fresh (B) uint8 flags, cleared by the
// launch's second kernel; coefs = [a_lp, 1-a_lp, a_ad, 1-a_ad, eps, contrast_clip].
struct PrCells {
    const float* rad;
    const float* sens;
    const int* pr_col;
    const int* cell_ptr;
    const int* cell_pr;
    const float* cell_w;
};

__global__ __launch_bounds__(256) void k_photoreceptor(PrCells pc, float* __restrict__ I_lp,
                                                       float* __restrict__ I_mean, const uint8_t* __restrict__ fresh,
                                                       const float* __restrict__ cf, float* __restrict__ contrast,
                                                       float* __restrict__ a_pr, int n_col, int n_pr,
                                                       int64_t total) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int n_cells = n_col * 5;
    const int k = (int)(t % n_cells), b = (int)(t / n_cells);
    const int e0 = pc.cell_ptr[k], e1 = pc.cell_ptr[k + 1];
    float I = 0.0f;
    for (int cs = e0; cs < e1; cs += 16) {
        const int ce = min(cs + 16, e1);
        float s = 0.0f;
        for (int e = cs; e < ce; ++e) {
            const int p = pc.cell_pr[e];
            const float* r = pc.rad + ((int64_t)b * n_col + pc.pr_col[p]) * 4;
            float Ip = 0.0f;
#pragma unroll
            for (int c = 0; c < 4; ++c) { float q = r[c] * pc.sens[p * 4 + c]; Ip = Ip + q; }
            float q = pc.cell_w[e] * Ip;
            s = s + q;
        }
        I = I + s;
    }
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
    if (e1 == e0) c = 0.0f;
    contrast[t] = c;
    for (int e = e0; e < e1; ++e) a_pr[(int64_t)b * n_pr + pc.cell_pr[e]] = c;
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

// Photoreceptor stage: p = [rad (B, n_col, 4), sens (n_pr, 4), pr_col int32 (n_pr), cell_ptr int32 (5 n_col + 1),
// cell_pr int32 (n_pr), cell_w (n_pr), I_lp (B, 5 n_col), I_mean (B, 5 n_col), fresh uint8 (B), coefs (6),
// contrast (B, 5 n_col), a_pr (B, n_pr)].
FH_EXPORT int launch_photoreceptor(void** p, int n_col, int n_pr, int batch, void* stream) {
    if (n_col < 0 || n_pr < 0 || batch < 0) return FH_ERR_ARGS;
    int64_t total = (int64_t)n_col * 5 * batch;
    if (!total) return 0;
    PrCells pc{(const float*)p[0], (const float*)p[1], (const int*)p[2], (const int*)p[3], (const int*)p[4],
               (const float*)p[5]};
    hipLaunchKernelGGL(k_photoreceptor, dim3(nblk(total)), dim3(256), 0, (hipStream_t)stream, pc, (float*)p[6],
                       (float*)p[7], (const uint8_t*)p[8], (const float*)p[9], (float*)p[10], (float*)p[11], n_col,
                       n_pr, total);
    hipLaunchKernelGGL(k_clear_u8, dim3((batch + 255) / 256), dim3(256), 0, (hipStream_t)stream, (uint8_t*)p[8], batch);
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
