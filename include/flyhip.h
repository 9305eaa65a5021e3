/* flyhip: the C ABI of the MI355X (gfx950) brain-step kernels, libflyhip.so.
 *
 * Conventions (those of the reference's flyverse/cuda.py:82-99 and flyverse/kernels/neural.cu):
 *   - every entry point returns int: 0 = ok, > 0 = the hipError_t of the failed launch, < 0 = an argument error
 *     found on the host before any launch (-1 bad count, -2 size out of range);
 *   - buffers are device pointers owned by the caller; no entry point allocates, frees or synchronises, and none
 *     reads index values on the host, so every call can be captured into a HIP graph;
 *   - `stream` is a hipStream_t passed as void*; work is ordered on it only;
 *   - float state is (batch, n) row-major ("fly-major") unless stated otherwise; indices are int32.
 * Empty shapes are valid no-ops.
 */
#ifndef FLYHIP_H
#define FLYHIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ reference entry points (same signatures) */

/* Fused LIF update.  Replaces launch_lif, flyverse/kernels/neural.cu:54 (bound at flyverse/cuda.py:73, called
 * from cuda.lif_update, cuda.py:102-113).  p = 16 pointers [v, g, drive, adapt, refrac, active(n), prob, rnd, res,
 * std_u(n), spikes, out, rate, counts, P(11), phase(8,n) or P].  flags: 1 Poisson, 2 STD, 4 counts, 8 adaptation,
 * 16 phase coefficients, 32 frozen-at-rest skip. */
int launch_lif(void** p, int n, int batch, int flags, void* stream);

/* dr = clip(v + b, 0, 1) - b.  Replaces launch_optic_dr, neural.cu:80 (cuda.py:116-120).  p = [v, b(n), dr]. */
int launch_optic_dr(void** p, int n, int batch, void* stream);

/* One optic substep given y = W_rr dr.  Replaces launch_optic, neural.cu:84 (cuda.py:123-128).
 * p = [v, adapt, dr, y, pr, spk, a(n), b(n)]. */
int launch_optic(void** p, int n, int batch, float gain_rr, float adapt_gain, float a_ad, void* stream);

/* y = W x for a row-major CSR, deterministic chunked summation order.  Replaces launch_csr, neural.cu:105
 * (cuda.py:131-153).  p = [ptr(rows+1), idx, values (float, or half if half), x (batch, cols), y (batch, rows)]. */
int launch_csr(void** p, int rows, int cols, int batch, int half, void* stream);

/* g += W x for the pre-major CSC of Brain.set_weights, deterministic: each target accumulates in ascending
 * presynaptic order (the reference CPU order).  Replaces launch_events, neural.cu:124 (cuda.py:156-168), whose float
 * atomicAdd does not repeat.  p = [ptr(n+1), post, w, x, g, pre_ids]; count < 0 walks every presynaptic cell. */
int launch_events(void** p, int n, int batch, int half, int count, void* stream);

/* Fixed-order double means of rates over index groups.  Replaces launch_means, neural.cu:141 (cuda.py:171-184).
 * p = [ptr(groups+1), idx, rates (batch, n), out double (batch, stride)]. */
int launch_means(void** p, int n, int groups, int batch, int stride, void* stream);

/* ------------------------------------------------------------------ additions of this backend */

/* launch_lif drawing the Poisson uniform in-kernel: Philox4x32-10 keyed (seed), counter (neuron, step, fly, 0) with
 * step = step + *p[16] (p[16] a device uint32, may be null).  Replaces the torch.rand draw that precedes
 * cuda.lif_update at flyverse/brain.py:877. */
int launch_lif_rng(void** p, int n, int batch, int flags, uint64_t seed, uint32_t step, void* stream);

/* out[b, i] = Philox uniform (seed, step + *step_base, i, b, stream_tag) in [0, 1).  Replaces torch.rand(B, N)
 * (flyverse/brain.py:848, 877) where an injected uniform is wanted. */
int launch_uniform(float* out, int n, int batch, uint64_t seed, const uint32_t* step_base, uint32_t step,
                   uint32_t stream_tag, void* stream);

/* x *= a over count floats: g.mul_(a_s), flyverse/brain.py:832. */
int launch_scale(float* x, int64_t count, float a, void* stream);

/* Fused decay + deterministic pull: g = g*a_s (unless flags & 2) then, per target, the products w*x of its
 * transmitting presynaptic cells (x != 0) folded in ascending presynaptic order.  Replaces the pair
 * g.mul_(a_s); cuda.event_scatter(...) of Brain._step_cuda, flyverse/brain.py:875-876.
 * p = [ptr(n+1), idx, w (float, or half if flags & 1), x (batch, n), g (batch, n)], post-major CSR. */
int launch_pull(void** p, int n, int batch, float a_s, int flags, void* stream);

/* The launch_csr product through a precomputed chunk table (one lane per 16-entry chunk, then one lane per row),
 * y = scale * W x (+ y if accumulate).  Replaces the _mv products of flyverse/optic.py:161-165, 579-580, 602.
 * p = [ptr, chunk_row(nchunks), chunk_ptr(rows+1), idx, w, x (batch, cols), y (batch, rows), partial (batch,
 * nchunks) scratch]. */
int launch_spmv_chunked(void** p, int rows, int cols, int nchunks, int batch, float scale, int accumulate,
                        void* stream);

/* Photoreceptor low-pass, adaptation and contrast, one cell per photoreceptor (flyverse/optic.py:496-517).
 * p = [rad (batch, n_pr, 4), sens (n_pr, 4), I_lp, I_mean, fresh uint8 (batch), coefs(6), a_pr]; clears fresh. */
int launch_photoreceptor(void** p, int n_pr, int batch, void* stream);

/* s = clamp(rate[:, spk_idx] / 100, 0, 3), flyverse/optic.py:572.  p = [rate (batch, n), spk_idx, s]. */
int launch_feedback(void** p, int n, int n_spk, int batch, void* stream);

/* drive[:, spk_idx] = clamp(gain_out * y_sr, -clip, clip), flyverse/optic.py:602.  p = [y_sr, spk_idx, drive]. */
int launch_drive_out(void** p, int n, int n_spk, int batch, float gain_out, float clip, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* FLYHIP_H */
