/* CPU oracle for the brain step: a plain-C restatement of the reference's arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product path (the HIP library or the flyhip package) links,
 * loads or calls this file; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg do, as
 * the checker.  Built with -ffp-contract=off so every product and sum rounds separately, as the
 * reference's CPU torch ops and its CUDA kernels (nvcc --fmad=false, flyverse/cuda.py:51) do.
 *
 * Each function cites the reference lines it restates (paths relative to the reference repository).
 * Pinning (what checks each function against something other than itself) is listed in oracle/__init__.py.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Threads of the parallel loops below (every one writes disjoint elements: per element, per row or per fly, so the
 * results do not depend on the thread count). */
void orc_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------ Philox4x32-10
 * Counter-based RNG of Salmon et al., "Parallel random numbers: as easy as 1, 2, 3" (SC'11), the
 * published 10-round Philox-4x32 with its published multipliers and Weyl key increments.  The reference
 * draws torch.rand from a stateful generator (flyverse/brain.py:848, 877); a stateful stream cannot be
 * reproduced per fly inside one launch, so this build keys every uniform by (seed, step, neuron, fly).
 * Pinned by the Random123 known-answer vectors in tests/test_oracle.py. */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += PHILOX_W0; k1 += PHILOX_W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Uniform float in [0, 1) from the first Philox word: the top 24 bits times 2^-24 (exact in fp32). */
static inline float u01(uint32_t x) { return (float)(x >> 8) * 5.9604644775390625e-08f; }

/* The draw that replaces torch.rand(B, N) at LIF step `step` (flyverse/brain.py:848/877):
 * counter (neuron, step, fly, stream), key (seed low, seed high).  stream 0 = Poisson forcing. */
float orc_uniform1(uint64_t seed, uint32_t step, uint32_t neuron, uint32_t fly, uint32_t stream) {
    uint32_t ctr[4] = {neuron, step, fly, stream}, key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)}, o[4];
    orc_philox4x32_10(ctr, key, o);
    return u01(o[0]);
}

/* out[j] = the uniform of cell idx[j] (counter (idx[j], step, fly, stream)): the draws a Poisson step can use, since
 * a uniform in [0, 1) is never below a zero probability. */
void orc_uniform_at(float* out, const int32_t* idx, int m, uint64_t seed, uint32_t step, uint32_t fly, uint32_t stream) {
    for (int j = 0; j < m; ++j) out[j] = orc_uniform1(seed, step, (uint32_t)idx[j], fly, stream);
}

void orc_uniform(float* out, int batch, int n, uint64_t seed, uint32_t step, uint32_t stream) {
#pragma omp parallel for schedule(static)
    for (int b = 0; b < batch; ++b)
        for (int i = 0; i < n; ++i) out[(int64_t)b * n + i] = orc_uniform1(seed, step, (uint32_t)i, (uint32_t)b, stream);
}

/* ------------------------------------------------------------------ LIF update (A1-A3)
 * One fused step per (b, i) in the operation order of Brain._step_inference (flyverse/brain.py:838-868),
 * with the flags and parameter packing of the native kernel (flyverse/kernels/neural.cu:11-52):
 *   flags 1 Poisson, 2 STD, 4 record counts, 8 adaptation, 16 per-neuron phase coefficients (8, N),
 *   32 frozen-at-rest skip.  P = [v_rest, v_reset, v_th, a_m, dt, t_ref, adapt_jump, a_ad, a_std, a_r, rate_gain]
 *   (flyverse/brain.py:517-520).  phase rows: a_m, dt, a_ad, a_std, a_r, rate_gain, update mask, Poisson multiplier
 *   (flyverse/brain.py:636-645).
 * The caller has already decayed g and added the synaptic input (flyverse/brain.py:832-833). */
void orc_lif(int n, int batch, int flags, float* v, float* g, const float* drive, float* adapt, float* refrac,
             const float* active, const float* prob, const float* rnd, float* res, const float* std_u,
             float* spikes, float* out, float* rate, float* counts, const float* P, const float* phase) {
    int64_t total = (int64_t)n * batch;
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < total; ++t) {
        int i = (int)(t % n);
        if ((flags & 32) && active[i] == 0.0f && (!(flags & 1) || prob[t] == 0.0f) && v[t] == P[0] && g[t] == 0.0f &&
            drive[t] == 0.0f && adapt[t] == 0.0f && refrac[t] == 0.0f && rate[t] == 0.0f && spikes[t] == 0.0f &&
            out[t] == 0.0f && res[t] == 1.0f)
            continue;                                     /* neural.cu:24-26: frozen cell at rest keeps its state */
        float am = P[3], dt = P[4], aad = P[7], ast = P[8], ar = P[9], rg = P[10], u = 1.0f, pois = 1.0f;
        if (flags & 16) {
            am = phase[i]; dt = phase[n + i]; aad = phase[2 * n + i]; ast = phase[3 * n + i];
            ar = phase[4 * n + i]; rg = phase[5 * n + i]; u = phase[6 * n + i]; pois = phase[7 * n + i];
        }
        float target = P[0] + g[t];                       /* brain.py:572: v_rest + g + drive - adapt */
        target = target + drive[t];
        target = target - adapt[t];
        float dv = v[t] - target;
        dv = dv * am;
        float vv = target + dv;                           /* brain.py:840 */
        float ref = refrac[t];
        if (ref > 0.0f) vv = P[1];                        /* brain.py:841-842 */
        ref = ref - dt;                                   /* brain.py:843 */
        if (ref < 0.0f) ref = 0.0f;
        float spike = (vv >= P[2]) ? 1.0f : 0.0f;         /* brain.py:845 */
        spike = spike * u;
        spike = spike * active[i];
        if (flags & 1) {                                  /* brain.py:847-849 */
            float pp = prob[t] * pois;
            if (rnd[t] < pp) spike = fmaxf(spike, 1.0f);
        }
        int fired = spike > 0.0f;
        if (fired) { vv = P[1]; ref = P[5]; }             /* brain.py:850-852 */
        v[t] = vv; refrac[t] = ref; spikes[t] = spike;
        if (flags & 8) {                                  /* brain.py:855-856: spike is 0/1 so jump*spike is exact */
            float a = adapt[t] * aad;
            float j = spike * P[6];
            adapt[t] = a + j;
        }
        if (flags & 2) {                                  /* brain.py:858-862 */
            float r = res[t];
            out[t] = spike * r;
            if (fired) { float keep = 1.0f - std_u[i]; r = r * keep; }
            float d = 1.0f - r;
            d = d * ast;
            res[t] = 1.0f - d;
        } else {
            out[t] = spike;                               /* brain.py:864 */
        }
        if (flags & 4) counts[t] = counts[t] + spike;     /* brain.py:866-867 */
        float rr = rate[t] * ar;                          /* brain.py:868 */
        float rs = spike * rg;
        rate[t] = rr + rs;
    }
}

/* g *= a_s (flyverse/brain.py:832) */
void orc_scale(float* g, int64_t count, float a) {
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < count; ++t) g[t] = g[t] * a;
}

/* ------------------------------------------------------------------ event-driven synaptic input (A4)
 * The reference CPU path, flyverse/brain.py:679-692: nz = nonzero(x) in (brain, pre) row-major order; for each
 * (b, pre) its out-edges in CSC order (sum_duplicates + sort_indices at brain.py:530); one serial index_add_, so
 * every target accumulates w*x in ascending presynaptic order, silent presynaptic cells skipped, each product
 * rounded before the add.  out_ptr/out_post/out_w are the pre-major CSC arrays of set_weights (brain.py:535-538). */
void orc_events(int n, int batch, const int64_t* out_ptr, const int32_t* out_post, const float* out_w,
                const float* x, float* g) {
#pragma omp parallel for schedule(static)
    for (int b = 0; b < batch; ++b) {
        const float* xb = x + (int64_t)b * n;
        float* gb = g + (int64_t)b * n;
        for (int pre = 0; pre < n; ++pre) {
            float xv = xb[pre];
            if (xv == 0.0f) continue;
            for (int64_t e = out_ptr[pre]; e < out_ptr[pre + 1]; ++e) {
                float c = out_w[e] * xv;
                gb[out_post[e]] = gb[out_post[e]] + c;
            }
        }
    }
}

/* ------------------------------------------------------------------ CSR product (A8)
 * y[b, r] = sum_j W[r, j] x[b, j] for the optic products (flyverse/optic.py:161-165, 474-483).  The reference's CPU
 * order (torch sparse mm) is not pinned by any reference vector, so this build DEFINES the order: the row's stored
 * entries (ascending column) are cut into chunks of 16 consecutive entries; each chunk is a sequential fp32 fold of
 * its products from +0 (each product rounded first); the row total is a sequential fold of the chunk sums from +0.
 * The GPU kernels reproduce this order bit for bit; parity with the reference itself is to 1e-5 relative
 * (checked against float64 scipy in tests).  half != 0: weights are float16 values (given here as float) and x is
 * rounded to half before the product (neural.cu:99). */
#define ORC_CHUNK 16
static float orc_half_round(float x);
void orc_spmv(int rows, int cols, int batch, const int64_t* ptr, const int32_t* idx, const float* w,
              const float* x, float* y, int half) {
#pragma omp parallel for collapse(2) schedule(dynamic, 256)
    for (int b = 0; b < batch; ++b)
        for (int r = 0; r < rows; ++r) {
            float total = 0.0f;
            for (int64_t cs = ptr[r]; cs < ptr[r + 1]; cs += ORC_CHUNK) {
                int64_t ce = cs + ORC_CHUNK < ptr[r + 1] ? cs + ORC_CHUNK : ptr[r + 1];
                float s = 0.0f;
                for (int64_t e = cs; e < ce; ++e) {
                    float xv = x[(int64_t)b * cols + idx[e]];
                    if (half) xv = orc_half_round(xv);
                    float c = w[e] * xv;
                    s = s + c;
                }
                total = total + s;
            }
            y[(int64_t)b * rows + r] = total;
        }
}

/* Round a float to the nearest IEEE binary16 value (ties to even), returned as float; used for half-weight products.
 * Written from the binary16 format definition (IEEE 754-2008 s3.6): 5 exponent bits, 10 fraction bits. */
static float orc_half_round(float x) {
    if (isnan(x) || isinf(x)) return x;
    float ax = fabsf(x);
    if (ax >= 65520.0f) return copysignf(INFINITY, x);          /* rounds up past the largest half, 65504 */
    int e;
    frexpf(ax, &e);                                               /* ax = m * 2^e, m in [0.5, 1) */
    int exp_unb = e - 1;                                          /* ax = 1.f * 2^exp_unb */
    if (exp_unb < -14) exp_unb = -14;                             /* subnormal halves share the 2^-24 quantum */
    float q = ldexpf(1.0f, exp_unb - 10);                         /* spacing of halves near ax */
    float r = ax / q;                                             /* exact: power-of-two scaling */
    float fl = floorf(r), frac = r - fl;
    if (frac > 0.5f || (frac == 0.5f && fmodf(fl, 2.0f) != 0.0f)) fl += 1.0f;
    return copysignf(fl * q, x);
}

/* ------------------------------------------------------------------ optic substep (A9)
 * dr = clip(v + b, 0, 1) - b  (flyverse/optic.py:520-523, 528; neural.cu:62-66) */
static inline float clip01(float x) { return x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x); }

void orc_optic_dr(int n, int batch, const float* v, const float* bvec, float* dr) {
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < (int64_t)n * batch; ++t) {
        float b = bvec[t % n];
        float s = v[t] + b;
        dr[t] = clip01(s) - b;
    }
}

/* One substep given y = W_rr @ dr (flyverse/optic.py:525-551; neural.cu:68-79):
 *   inp = gain_rr*y + pr - adapt_gain*adapt; inp = inp + spk   (the addition order is fixed, optic.py:530-532)
 *   v = inp + (v - inp)*a; adapt = dr + (adapt - dr)*a_ad; then the next dr (fused, neural.cu:78). */
void orc_optic_update(int n, int batch, float* v, float* adapt, float* dr, const float* y, const float* pr,
                      const float* spk, const float* a, const float* bvec, float gain_rr, float adapt_gain, float a_ad) {
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < (int64_t)n * batch; ++t) {
        int i = (int)(t % n);
        float old = dr[t];
        float inp = gain_rr * y[t];
        inp = inp + pr[t];
        float ga = adapt_gain * adapt[t];
        inp = inp - ga;
        inp = inp + spk[t];
        float d = v[t] - inp;
        d = d * a[i];
        float vv = inp + d;
        v[t] = vv;
        float da = adapt[t] - old;
        da = da * a_ad;
        adapt[t] = old + da;
        float s = vv + bvec[i];
        dr[t] = clip01(s) - bvec[i];
    }
}

/* ------------------------------------------------------------------ photoreceptor stage (A10)
 * flyverse/optic.py:496-517 over (column, family) cells, cell = column * 5 + family (optic.py:315-325).  For batch
 * row b and cell k: each photoreceptor p of the cell has intensity I_p = sum over the 4 spectral channels, in
 * channel order, of rad[b, column(p), c] * sens[p, c] (sequential fold from +0); the cell's input is the averaging
 * matrix row applied to I (weights cell_w = 1/count as float32, photoreceptors ascending) in the CSR product's
 * chunked order (chunks of 16 folded from +0, chunk sums folded from +0); the first frame of a row initialises the
 * low-pass and the mean to that input (optic.py:505-508); then I_lp = a_lp*I_lp + (1-a_lp)*I;
 * I_mean = a_ad*I_mean + (1-a_ad)*I_lp; contrast = clamp((I_lp - I_mean)/(I_mean + eps), -1, clip), replaced by 0
 * for a cell without photoreceptors (optic.py:513-514); each photoreceptor's activity is its cell's contrast
 * (optic.py:517).  coefs = [a_lp, 1-a_lp, a_ad, 1-a_ad, eps, clip] as float32 (the host computes them in double,
 * optic.py:509).  rad (B, n_col, 4); I_lp, I_mean, contrast (B, n_col*5); a_pr (B, n_pr). */
void orc_photoreceptor(int n_col, int n_pr, int batch, const float* rad, const float* sens, const int32_t* pr_col,
                       const int32_t* cell_ptr, const int32_t* cell_pr, const float* cell_w, float* I_lp,
                       float* I_mean, uint8_t* fresh, const float* coefs, float* contrast, float* a_pr) {
    const int n_cells = n_col * 5;
#pragma omp parallel for collapse(2) schedule(static)
    for (int b = 0; b < batch; ++b) {
        for (int k = 0; k < n_cells; ++k) {
            int64_t t = (int64_t)b * n_cells + k;
            float I = 0.0f;
            for (int cs = cell_ptr[k]; cs < cell_ptr[k + 1]; cs += ORC_CHUNK) {
                int ce = cs + ORC_CHUNK < cell_ptr[k + 1] ? cs + ORC_CHUNK : cell_ptr[k + 1];
                float s = 0.0f;
                for (int e = cs; e < ce; ++e) {
                    int p = cell_pr[e];
                    const float* r = rad + ((int64_t)b * n_col + pr_col[p]) * 4;
                    float Ip = 0.0f;
                    for (int c = 0; c < 4; ++c) { float q = r[c] * sens[p * 4 + c]; Ip = Ip + q; }
                    float q = cell_w[e] * Ip;
                    s = s + q;
                }
                I = I + s;
            }
            float lp = fresh[b] ? I : I_lp[t];
            float mean = fresh[b] ? I : I_mean[t];
            float l1 = coefs[0] * lp, l2 = coefs[1] * I;
            lp = l1 + l2;
            float m1 = coefs[2] * mean, m2 = coefs[3] * lp;
            mean = m1 + m2;
            I_lp[t] = lp; I_mean[t] = mean;
            float num = lp - mean, den = mean + coefs[4];
            float c = num / den;
            if (c < -1.0f) c = -1.0f;
            if (c > coefs[5]) c = coefs[5];
            if (cell_ptr[k + 1] == cell_ptr[k]) c = 0.0f;
            contrast[t] = c;
            for (int e = cell_ptr[k]; e < cell_ptr[k + 1]; ++e) a_pr[(int64_t)b * n_pr + cell_pr[e]] = c;
        }
    }
    for (int b = 0; b < batch; ++b) fresh[b] = 0;
}

/* ------------------------------------------------------------------ motor group means (next row)
 * Fixed-order double accumulation of rates[b, idx[ptr[k]..ptr[k+1])] per group (neural.cu:134-140). */
void orc_group_means(int n, int groups, int batch, int stride, const int32_t* ptr, const int32_t* idx,
                     const float* rates, double* out) {
    for (int b = 0; b < batch; ++b)
        for (int k = 0; k < groups; ++k) {
            double s = 0.0;
            for (int j = ptr[k]; j < ptr[k + 1]; ++j) s += (double)rates[(int64_t)b * n + idx[j]];
            int cnt = ptr[k + 1] - ptr[k];
            out[(int64_t)b * stride + k] = cnt ? s / cnt : 0.0;
        }
}

/* ------------------------------------------------------------------ optic output drive (A10)
 * drive[b, spk_idx[k]] = clamp(gain_out * (W_sr @ dr)[b, k], -clip, clip); every other entry 0
 * (flyverse/optic.py:599-602 and fly.py:492, where drive is overwritten by the frame's output). */
void orc_drive_out(int n, int n_spk, int batch, const float* y_sr, const int32_t* spk_idx, float gain_out,
                   float clip, float* drive) {
    memset(drive, 0, sizeof(float) * (size_t)n * batch);
    for (int b = 0; b < batch; ++b)
        for (int k = 0; k < n_spk; ++k) {
            float d = gain_out * y_sr[(int64_t)b * n_spk + k];
            if (d < -clip) d = -clip;
            if (d > clip) d = clip;
            drive[(int64_t)b * n + spk_idx[k]] = d;
        }
}

/* feedback s = clamp(rate[spk_idx] / 100, 0, 3)  (flyverse/optic.py:572) */
void orc_feedback(int n, int n_spk, int batch, const float* rate, const int32_t* spk_idx, float* s) {
    for (int b = 0; b < batch; ++b)
        for (int k = 0; k < n_spk; ++k) {
            float q = rate[(int64_t)b * n + spk_idx[k]] / 100.0f;
            if (q < 0.0f) q = 0.0f;
            if (q > 3.0f) q = 3.0f;
            s[(int64_t)b * n_spk + k] = q;
        }
}
