// Elementwise optic-lobe kernels for the plain (B, n) fly-major layout, and the per-frame coupling kernels.
// Semantics: the substep of flyverse/optic.py:520-551 (native spelling flyverse/kernels/neural.cu:62-88), the
// photoreceptor stage optic.py:496-517, the feedback optic.py:572 and the output drive optic.py:599-602.
#include "common.h"
#include "groups.h"
#include "panel.h"

namespace {

// The feedback value clamp(rate / 100, 0, 3) (true division, as the CPU reference computes it).
__device__ __forceinline__ float fb_value(float rate) {
    float q = rate / 100.0f;
    q = q < 0.0f ? 0.0f : q;
    return q > 3.0f ? 3.0f : q;
}

// The output drive clamp(gain_out * y, -clip, clip).
__device__ __forceinline__ float drive_value(float y, float gain_out, float clip) {
    float d = gain_out * y;
    d = d < -clip ? -clip : d;
    return d > clip ? clip : d;
}

// dr = clip01(v + b) - b.
__device__ __forceinline__ float dr_value(float v, float base) {
    float s = v + base;
    return fh_clip01(s) - base;
}

__global__ __launch_bounds__(256) void k_optic_dr(const float* __restrict__ v, const float* __restrict__ b,
                                                  float* __restrict__ dr, int n, int64_t total) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    dr[t] = dr_value(v[t], b[t % n]);
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

// One substep fused with its recurrent product (flyverse/optic.py:474-483 then 525-551): y = W_rr dr_cur for the
// rows of group g (grouped order of launch_spmv_grouped), then each row's update as k_optic_update computes it, the
// next dr into dr_next (a different buffer: other rows' products still read dr_cur).  One wavefront per (group, fly).
template <class Matrix>
struct SubstepArgs {
    Matrix rr;
    const float* dr_cur;
    float *dr_next, *v, *adapt;
    const float *pr, *spk, *a, *b;
    int n_rate;
    float gr, ga, aad;
};

// The update of row i of the fly whose state starts at `base`, given its recurrent product total.
template <class Matrix>
__device__ __forceinline__ void substep_update(const SubstepArgs<Matrix>& s, int64_t base, int i, float total) {
    const int64_t t = base + i;
    const float old = s.dr_cur[t];
    float inp = s.gr * total;
    inp = inp + s.pr[t];
    float gad = s.ga * s.adapt[t];
    inp = inp - gad;
    inp = inp + s.spk[t];
    float d = s.v[t] - inp;
    d = d * s.a[i];
    float v = inp + d;
    s.v[t] = v;
    float da = s.adapt[t] - old;
    da = da * s.aad;
    s.adapt[t] = old + da;
    s.dr_next[t] = dr_value(v, s.b[i]);
}

__global__ __launch_bounds__(256) void k_optic_substep(SubstepArgs<fh_groups::Grouped> s) {
    __shared__ float sums[4][kWave];
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int g = blockIdx.x * 4 + wv;
    if (g >= s.rr.ngroups) return;
    const int64_t base = (int64_t)blockIdx.y * s.n_rate;
    const float* x = s.dr_cur + base;
    fh_groups::group_product(s.rr, g, lane, sums[wv], [&](int i) { return x[i]; },
                             [&](int i, float total) { substep_update(s, base, i, total); });
}

// The same substep with the W_rr product by persistent panels (panel.h): one 1024-thread workgroup per (row block,
// fly), x = dr_cur staged in LDS a panel at a time; each row's update runs when its total is folded.
__global__ __launch_bounds__(1024) void k_optic_substep_panel(SubstepArgs<fh_panel::Panel> s) {
    extern __shared__ float xs[];
    const int64_t base = (int64_t)blockIdx.y * s.n_rate;
    fh_panel::block_product(s.rr, blockIdx.x, s.dr_cur + base, xs,
                            [&](int i, float total) { substep_update(s, base, i, total); });
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

// Photoreceptor cell k of fly b; t = b * n_cells + k.
__device__ __forceinline__ void pr_cell(const PrCells& pc, float* __restrict__ I_lp, float* __restrict__ I_mean,
                                        const uint8_t* __restrict__ fresh, const float* __restrict__ cf,
                                        float* __restrict__ contrast, float* __restrict__ a_pr, int n_col, int n_pr,
                                        int b, int k) {
    const int64_t t = (int64_t)b * (n_col * 5) + k;
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

__global__ __launch_bounds__(256) void k_photoreceptor(PrCells pc, float* __restrict__ I_lp,
                                                       float* __restrict__ I_mean, const uint8_t* __restrict__ fresh,
                                                       const float* __restrict__ cf, float* __restrict__ contrast,
                                                       float* __restrict__ a_pr, int n_col, int n_pr,
                                                       int64_t total) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int n_cells = n_col * 5;
    pr_cell(pc, I_lp, I_mean, fresh, cf, contrast, a_pr, n_col, n_pr, (int)(t / n_cells), (int)(t % n_cells));
}

__global__ void k_clear_u8(uint8_t* f, int n) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) f[t] = 0;
}

// s[b, k] = feedback of rate[b, spk_idx[k]].
__global__ __launch_bounds__(256) void k_feedback(const float* __restrict__ rate, const int* __restrict__ spk_idx,
                                                  float* __restrict__ s, int n, int n_spk, int64_t total) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    int k = (int)(t % n_spk), b = (int)(t / n_spk);
    s[t] = fb_value(rate[(int64_t)b * n + spk_idx[k]]);
}

// drive[b, spk_idx[k]] = clamp(gain_out * y_sr[b, k], -clip, clip); the caller zeroes drive's other entries once.
__global__ __launch_bounds__(256) void k_drive_out(const float* __restrict__ y_sr, const int* __restrict__ spk_idx,
                                                   float* __restrict__ drive, int n, int n_spk, float gain_out,
                                                   float clip, int64_t total) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    int k = (int)(t % n_spk), b = (int)(t / n_spk);
    drive[(int64_t)b * n + spk_idx[k]] = drive_value(y_sr[t], gain_out, clip);
}

// ------------------------------------------------------------------ the frame's optic inputs and outputs, fused
// Frame start (flyverse/optic.py:496-517, 572, 579-580) in two launches.  k_frame_cells, per fly: the photoreceptor
// cells (state, contrast, activity a_pr), the feedback s of every spiking target, and the first substep's
// dr = clip01(v + b) - b.  k_frame_held, per fly, only in a frame with substeps: pr_in = gain_in * W_rp a_pr and
// spk_in = gain_fb * W_rs s, grouped products over a_pr and s as the first launch left them.  (Within one launch,
// ordering the products after every photoreceptor cell takes a device-wide count per workgroup; on this GPU those
// same-address atomics cost far more than the launch boundary.)  Values and order are those of the separate kernels.
struct FrameIn {
    PrCells pc;
    float *I_lp, *I_mean, *contrast, *a_pr;
    const uint8_t* fresh;
    const float* cf;
    const float* rate;      // (B, n) LIF rates
    const int* spk_idx;     // (n_spk)
    float* s;               // (B, n_spk)
    const float *v, *bvec;  // (B, n_rate), (n_rate)
    float* dr;              // (B, n_rate)
    fh_groups::Grouped rp, rs;
    float *pr_in, *spk_in;  // (B, n_rate)
    float gain_in, gain_fb;
    int n_col, n_pr, n, n_spk, n_rate;
    int j_cells, j_spk, j_rate, j_rp;   // workgroups per fly of each job kind
};

__global__ __launch_bounds__(256) void k_frame_cells(FrameIn a) {
    const int b = blockIdx.y;
    int j = blockIdx.x;
    if (j < a.j_cells) {
        const int k = j * 256 + threadIdx.x;
        if (k < a.n_col * 5) pr_cell(a.pc, a.I_lp, a.I_mean, a.fresh, a.cf, a.contrast, a.a_pr, a.n_col, a.n_pr, b, k);
    } else if ((j -= a.j_cells) < a.j_spk) {
        const int k = j * 256 + threadIdx.x;
        if (k < a.n_spk) a.s[(int64_t)b * a.n_spk + k] = fb_value(a.rate[(int64_t)b * a.n + a.spk_idx[k]]);
    } else {
        const int r = (j - a.j_spk) * 256 + threadIdx.x;
        if (r < a.n_rate) a.dr[(int64_t)b * a.n_rate + r] = dr_value(a.v[(int64_t)b * a.n_rate + r], a.bvec[r]);
    }
}

__global__ __launch_bounds__(256) void k_frame_held(FrameIn a) {
    __shared__ float sums[4][kWave];
    const int b = blockIdx.y, j = blockIdx.x;
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (j < a.j_rp) {
        const int g = j * 4 + wv;
        const float* x = a.a_pr + (int64_t)b * a.n_pr;
        float* y = a.pr_in + (int64_t)b * a.n_rate;
        if (g < a.rp.ngroups)
            fh_groups::group_product(a.rp, g, lane, sums[wv], [&](int i) { return x[i]; },
                                     [&](int row, float tot) { y[row] = tot * a.gain_in; });
    } else {
        const int g = (j - a.j_rp) * 4 + wv;
        const float* x = a.s + (int64_t)b * a.n_spk;
        float* y = a.spk_in + (int64_t)b * a.n_rate;
        if (g < a.rs.ngroups)
            fh_groups::group_product(a.rs, g, lane, sums[wv], [&](int i) { return x[i]; },
                                     [&](int row, float tot) { y[row] = tot * a.gain_fb; });
    }
}

// Frame end (flyverse/optic.py:599-604): drive[b, spk_idx[k]] = clamp(gain_out * (W_sr dr)[k]) with
// dr = clip01(v + b) - b, the rate change from r0 = b.  READY: dr_src is the last substep's output, read as is (the
// substep writes exactly this value), and copied into dr when it is another buffer; otherwise (a frame without
// substeps) the product forms dr from v as it gathers, and the launch's own rate-unit jobs store it.  No job reads
// what another writes.  Workgroup (0, 0) also clears the fresh flags, which the frame start read.
struct FrameOut {
    fh_groups::Grouped sr;
    const float *v, *bvec;  // (B, n_rate), (n_rate)
    const float* dr_src;    // (B, n_rate): the last substep's output (READY)
    float* dr;              // (B, n_rate)
    const int* spk_idx;     // (n_spk)
    float* drive;           // (B, n)
    uint8_t* fresh;         // (B)
    float gain_out, clip;
    int n, n_spk, n_rate, j_sr, batch;
};

template <bool READY>
__global__ __launch_bounds__(256) void k_frame_out(FrameOut a) {
    __shared__ float sums[4][kWave];
    const int b = blockIdx.y, j = blockIdx.x;
    if (b == 0 && j == 0)
        for (int i = threadIdx.x; i < a.batch; i += blockDim.x) a.fresh[i] = 0;
    const float* vb = a.v + (int64_t)b * a.n_rate;
    const float* srcb = a.dr_src + (int64_t)b * a.n_rate;
    float* drb = a.dr + (int64_t)b * a.n_rate;
    if (j < a.j_sr) {
        const int lane = threadIdx.x & 63;
        const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        const int g = j * 4 + wv;
        if (g >= a.sr.ngroups) return;
        float* drive = a.drive + (int64_t)b * a.n;
        auto finish = [&](int row, float tot) { drive[a.spk_idx[row]] = drive_value(tot, a.gain_out, a.clip); };
        if (READY)
            fh_groups::group_product(a.sr, g, lane, sums[wv], [&](int i) { return srcb[i]; }, finish);
        else
            fh_groups::group_product(a.sr, g, lane, sums[wv], [&](int i) { return dr_value(vb[i], a.bvec[i]); },
                                     finish);
    } else {
        const int r = (j - a.j_sr) * 256 + threadIdx.x;
        if (r < a.n_rate) drb[r] = READY ? srcb[r] : dr_value(vb[r], a.bvec[r]);
    }
}

__global__ __launch_bounds__(128) void k_means(const int* __restrict__ ptr, const int* __restrict__ idx,
                                               const float* __restrict__ rates, double* __restrict__ out, int n,
                                               int groups, int batch, int stride) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= groups * batch) return;
    int b = t / groups, k = t % groups;
    const float* rb = rates + (int64_t)b * n;
    double sum = 0.0;
    int j = ptr[k];
    const int e = ptr[k + 1];
    for (; j + 8 <= e; j += 8) {                       // eight gathers in flight, then their adds in order
        int c[8];
        float v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) c[q] = idx[j + q];
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = rb[c[q]];
#pragma unroll
        for (int q = 0; q < 8; ++q) sum += (double)v[q];
    }
    for (; j < e; ++j) sum += (double)rb[idx[j]];
    int cnt = e - ptr[k];
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

// One substep fused with its recurrent product: p = [W_rr group_row, chunk_ptr, pidx, pw (grouped storage of
// launch_spmv_grouped), dr_cur, dr_next, v, adapt, pr, spk (B, n_rate each), a, b (n_rate)]; ngroups = W_rr groups.
// The product of launch_spmv_grouped (scale 1) followed by launch_optic's update, writing the next dr into dr_next.
FH_EXPORT int launch_optic_substep(void** p, int n_rate, int ngroups, int batch, float gr, float ga, float aad,
                                   void* stream) {
    if (n_rate < 0 || ngroups < 0 || batch < 0 || batch > 65535) return FH_ERR_ARGS;
    if (!n_rate || !ngroups || !batch) return 0;
    SubstepArgs<fh_groups::Grouped> s{fh_groups::Grouped{(const int*)p[0], (const int*)p[1], (const int4*)p[2],
                                                         (const float4*)p[3], ngroups},
                  (const float*)p[4], (float*)p[5], (float*)p[6], (float*)p[7], (const float*)p[8],
                  (const float*)p[9], (const float*)p[10], (const float*)p[11], n_rate, gr, ga, aad};
    hipLaunchKernelGGL(k_optic_substep, dim3((ngroups + 3) / 4, batch), dim3(256), 0, (hipStream_t)stream, s);
    FH_LAUNCH_CHECK();
}

// The substep with its W_rr product by panels: p = [W_rr blk_row (nblk+1; blocks of launch_spmv_panel), chunk_ptr,
// pidx, pw (slot order), dr_cur, dr_next, v, adapt, pr, spk (B, n_rate each), a, b (n_rate), nat (slot -> chunk, or
// null)].  Same values as launch_optic_substep.
FH_EXPORT int launch_optic_substep_panel(void** p, int n_rate, int nblk, int batch, float gr, float ga, float aad,
                                         void* stream) {
    if (n_rate < 0 || nblk < 0 || batch < 0 || batch > 65535) return FH_ERR_ARGS;
    if (!n_rate || !nblk || !batch) return 0;
    const int pc = fh_panel::panel_columns(n_rate, 0);
    const int lds_floats = fh_panel::lds_floats(pc);
    hipError_t e = hipFuncSetAttribute((const void*)k_optic_substep_panel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, lds_floats * 4);
    if (e != hipSuccess) return (int)e;
    SubstepArgs<fh_panel::Panel> s{fh_panel::Panel{(const int*)p[0], (const int*)p[1], (const int4*)p[2],
                                                   (const float4*)p[3], (const int*)p[12], n_rate, pc},
                                   (const float*)p[4], (float*)p[5], (float*)p[6], (float*)p[7], (const float*)p[8],
                                   (const float*)p[9], (const float*)p[10], (const float*)p[11], n_rate, gr, ga, aad};
    hipLaunchKernelGGL(k_optic_substep_panel, dim3(nblk, batch), dim3(fh_panel::kThreads), (size_t)lds_floats * 4,
                       (hipStream_t)stream, s);
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

// The frame's optic inputs (flyverse/optic.py:496-517, 572, 579-580), two launches: the photoreceptor cells, feedback
// and first dr, then (products != 0) the held inputs.  p = [rad (B, n_col, 4), sens, pr_col, cell_ptr, cell_pr,
// cell_w, I_lp, I_mean, fresh uint8 (B), coefs (6), contrast, a_pr (B, n_pr), rate (B, n), spk_idx int32 (n_spk),
// s (B, n_spk), v (B, n_rate), b (n_rate), dr (B, n_rate), W_rp group_row, chunk_ptr, pidx, pw, W_rs group_row,
// chunk_ptr, pidx, pw, pr_in (B, n_rate), spk_in (B, n_rate)]; dims = [n_col, n_pr, n, n_spk, n_rate, batch,
// W_rp groups, W_rs groups, products].  Chunk-padded grouped storage as launch_spmv_grouped.  fresh is read, not
// cleared: launch_frame_out clears it.
FH_EXPORT int launch_frame_in(void** p, const int* dims, float gain_in, float gain_fb, void* stream) {
    for (int i = 0; i < 9; ++i)
        if (dims[i] < 0) return FH_ERR_ARGS;
    if (dims[5] > 65535) return FH_ERR_ARGS;
    FrameIn a;
    a.pc = PrCells{(const float*)p[0], (const float*)p[1], (const int*)p[2], (const int*)p[3], (const int*)p[4],
                   (const float*)p[5]};
    a.I_lp = (float*)p[6]; a.I_mean = (float*)p[7]; a.fresh = (const uint8_t*)p[8]; a.cf = (const float*)p[9];
    a.contrast = (float*)p[10]; a.a_pr = (float*)p[11]; a.rate = (const float*)p[12]; a.spk_idx = (const int*)p[13];
    a.s = (float*)p[14]; a.v = (const float*)p[15]; a.bvec = (const float*)p[16]; a.dr = (float*)p[17];
    a.rp = fh_groups::Grouped{(const int*)p[18], (const int*)p[19], (const int4*)p[20], (const float4*)p[21], dims[6]};
    a.rs = fh_groups::Grouped{(const int*)p[22], (const int*)p[23], (const int4*)p[24], (const float4*)p[25], dims[7]};
    a.pr_in = (float*)p[26]; a.spk_in = (float*)p[27];
    a.gain_in = gain_in; a.gain_fb = gain_fb;
    a.n_col = dims[0]; a.n_pr = dims[1]; a.n = dims[2]; a.n_spk = dims[3]; a.n_rate = dims[4];
    const int batch = dims[5];
    a.j_cells = (int)nblk((int64_t)a.n_col * 5);
    a.j_spk = (int)nblk(a.n_spk);
    a.j_rate = (int)nblk(a.n_rate);
    a.j_rp = (a.rp.ngroups + 3) / 4;
    const int g1 = a.j_cells + a.j_spk + a.j_rate, g2 = a.j_rp + (a.rs.ngroups + 3) / 4;
    if (!batch) return 0;
    if (g1) {
        hipLaunchKernelGGL(k_frame_cells, dim3(g1, batch), dim3(256), 0, (hipStream_t)stream, a);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return (int)e;
    }
    if (dims[8] && g2) hipLaunchKernelGGL(k_frame_held, dim3(g2, batch), dim3(256), 0, (hipStream_t)stream, a);
    FH_LAUNCH_CHECK();
}

// The frame's optic output into the LIF drive, fused (flyverse/optic.py:599-604):
// p = [W_sr group_row, chunk_ptr, pidx, pw, v (B, n_rate), b (n_rate), dr (B, n_rate), spk_idx int32 (n_spk),
//      drive (B, n), fresh uint8 (B), dr_src (B, n_rate)]; dims = [n, n_spk, n_rate, batch, W_sr groups, ready].
// ready: dr_src holds clip01(v + b) - b already (the last substep's output) and is copied into dr if it is another
// buffer; otherwise the launch forms it from v and stores it in dr (dr_src unused).  Only the spiking targets of
// drive are written.  Clears fresh.
FH_EXPORT int launch_frame_out(void** p, const int* dims, float gain_out, float clip, void* stream) {
    for (int i = 0; i < 6; ++i)
        if (dims[i] < 0) return FH_ERR_ARGS;
    if (dims[3] > 65535) return FH_ERR_ARGS;
    FrameOut a;
    a.sr = fh_groups::Grouped{(const int*)p[0], (const int*)p[1], (const int4*)p[2], (const float4*)p[3], dims[4]};
    a.v = (const float*)p[4]; a.bvec = (const float*)p[5]; a.dr = (float*)p[6]; a.spk_idx = (const int*)p[7];
    a.drive = (float*)p[8]; a.fresh = (uint8_t*)p[9]; a.dr_src = (const float*)p[10];
    a.gain_out = gain_out; a.clip = clip;
    a.n = dims[0]; a.n_spk = dims[1]; a.n_rate = dims[2]; a.batch = dims[3];
    a.j_sr = (a.sr.ngroups + 3) / 4;
    const bool ready = dims[5] != 0;
    const bool store = !ready || a.dr_src != a.dr;
    const int gx = max(1, a.j_sr + (store ? (int)nblk(a.n_rate) : 0));
    if (!dims[3]) return 0;
    if (ready)
        hipLaunchKernelGGL(k_frame_out<true>, dim3(gx, dims[3]), dim3(256), 0, (hipStream_t)stream, a);
    else
        hipLaunchKernelGGL(k_frame_out<false>, dim3(gx, dims[3]), dim3(256), 0, (hipStream_t)stream, a);
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
