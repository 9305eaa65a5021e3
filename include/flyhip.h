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

/* The synaptic input of the clocked step (per-module clocks): g = g * a_s[row] (this phase's decay, 1 for rows whose
 * clock does not tick), then, when tick_mask is nonzero, g = g + y with y the row's product over its clock class's
 * input (w * x for x != 0 folded from +0 in ascending presynaptic order) if that class ticks, +0 otherwise.
 * Replaces g.mul_(c["a_s"]) and the _matmul_add calls of Brain._step_clocked, flyverse/brain.py:905-921.
 * p = [ptr (n+1), idx, w (float), g (batch, n), row_class uint8 (n), a_s (n), x_0 .. x_{ncls-1} (batch, n)];
 * ncls <= 8; bit j of tick_mask: class j forms its product this step. */
int launch_pull_clocked(void** p, int n, int batch, int ncls, int tick_mask, void* stream);

/* The clock accumulators of the clocked step (flyverse/brain.py:914-918): acc_j += x for j in add_mask, then
 * acc_j = 0 for j in clear_mask.  p = [x (count), acc_0 .. acc_{nacc-1} (count each)]; nacc <= 8. */
int launch_clock_acc(void** p, int64_t count, int nacc, int add_mask, int clear_mask, void* stream);

/* The launch_csr product through a precomputed chunk table (one lane per 16-entry chunk, then one lane per row),
 * y = scale * W x (+ y if accumulate).  Replaces the _mv products of flyverse/optic.py:161-165, 579-580, 602.
 * p = [ptr, chunk_row(nchunks), chunk_ptr(rows+1), idx, w, x (batch, cols), y (batch, rows), partial (batch,
 * nchunks) scratch]. */
int launch_spmv_chunked(void** p, int rows, int cols, int nchunks, int batch, float scale, int accumulate,
                        void* stream);

/* launch_spmv_chunked without the workspace argument, for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc); it has no counterpart in the reference.  The chunk partial
 * sums go to a device array of the library itself (4,194,304 floats per GPU, held until the process ends), so
 * batch * nchunks must not exceed 4,194,304 (FH_ERR_SIZE otherwise).  Calls that overlap on one GPU share that array
 * and race.  p = [ptr, chunk_row(nchunks), chunk_ptr(rows+1), idx, w, x (batch, cols), y (batch, rows)]; the stream
 * belongs to the current device.  Allocates nothing. */
int launch_spmv_chunked_noalloc(void** p, int rows, int cols, int nchunks, int batch, float scale, int accumulate,
                                void* stream);

/* The launch_csr product in one pass over chunk-padded storage (row r's entries start at 16 * chunk_ptr[r], pads
 * have idx -1 and add nothing): one wavefront per row group (consecutive rows of at most 64 chunks and 64 rows
 * together, or one longer row) and fly; same additions in the same order as launch_csr.
 * p = [group_row (ngroups+1), chunk_ptr (rows+1), pidx (16*nchunks), pw (16*nchunks), x (batch, cols),
 * y (batch, rows)].  y = scale * W x (+ y if accumulate).  Replaces the _mv products of flyverse/optic.py:161-165. */
int launch_spmv_grouped(void** p, int rows, int cols, int ngroups, int batch, float scale, int accumulate,
                        void* stream);

/* The launch_csr product with x staged in LDS, range of columns by range (a 4-byte gather through the vector cache
 * moves a whole line from L2; from LDS it moves 4 bytes).  One 1024-thread workgroup per tile of at most 1024
 * chunks and 1024 rows; same additions in the same order as launch_csr.  p = [tile_row (ntiles+1), chunk_ptr
 * (rows+1), pidx (16*nchunks), pw (16*nchunks), x (batch, cols), y (batch, rows)]; range_cols = columns staged at
 * a time (0: as many as the LDS holds).  Replaces the _mv products of flyverse/optic.py:161-165. */
int launch_spmv_lds(void** p, int rows, int cols, int ntiles, int batch, int range_cols, float scale, int accumulate,
                    void* stream);

/* The launch_csr product by persistent panels: one 1024-thread workgroup per (row block, fly) holds the block's
 * chunks (at most 3 per thread) in registers for the whole launch, and x is staged in LDS one panel of columns at a
 * time, each workgroup reading x once; same additions in the same order as launch_csr.  p = [blk_row (nblk+1): first
 * row of each block of at most 3 * 1024 chunks, chunk_ptr (rows+1), pidx (16*nchunks), pw (16*nchunks), x (batch,
 * cols), y (batch, rows), nat (nchunks: the chunk each slot of pidx / pw holds, a permutation within each block, or
 * null)]; panel = columns staged at a time (0: as few equal panels as the LDS allows).
 * y = scale * W x (+ y if accumulate).  Replaces the _mv products of flyverse/optic.py:161-165. */
int launch_spmv_panel(void** p, int rows, int cols, int nblk, int batch, int panel, float scale, int accumulate,
                      void* stream);

/* Photoreceptor stage per (column, family) cell (flyverse/optic.py:496-517): the averaging matrix of
 * optic.py:320-325 (cell_w = 1/count) over the photoreceptor intensities, low-pass, adaptation and contrast (0 for an
 * empty cell), each photoreceptor's activity = its cell's contrast.  p = [rad (batch, n_col, 4), sens (n_pr, 4),
 * pr_col int32 (n_pr), cell_ptr int32 (5 n_col + 1), cell_pr int32 (n_pr), cell_w (n_pr), I_lp, I_mean
 * (batch, 5 n_col), fresh uint8 (batch), coefs (6), contrast (batch, 5 n_col), a_pr (batch, n_pr)]; clears fresh. */
int launch_photoreceptor(void** p, int n_col, int n_pr, int batch, void* stream);

/* s = clamp(rate[:, spk_idx] / 100, 0, 3), flyverse/optic.py:572.  p = [rate (batch, n), spk_idx, s]. */
int launch_feedback(void** p, int n, int n_spk, int batch, void* stream);

/* drive[:, spk_idx] = clamp(gain_out * y_sr, -clip, clip), flyverse/optic.py:602.  p = [y_sr, spk_idx, drive]. */
int launch_drive_out(void** p, int n, int n_spk, int batch, float gain_out, float clip, void* stream);

/* One optic substep with its recurrent product fused (flyverse/optic.py:474-483 then 525-551; neural.cu:68-79 for
 * the update): y = W_rr dr_cur in the grouped order of launch_spmv_grouped, then launch_optic's update, the next dr
 * into dr_next (another buffer).  p = [W_rr group_row, chunk_ptr, pidx, pw, dr_cur, dr_next, v, adapt, pr, spk
 * (batch, n_rate each), a, b (n_rate)]; ngroups = W_rr groups.  Replaces a launch_csr + launch_optic pair. */
int launch_optic_substep(void** p, int n_rate, int ngroups, int batch, float gr, float ga, float aad, void* stream);

/* launch_optic_substep with the W_rr product by persistent panels (launch_spmv_panel's blocks and order, which are
 * launch_csr's additions): p = [W_rr blk_row (nblk+1), chunk_ptr, pidx, pw, dr_cur, dr_next, v, adapt, pr, spk
 * (batch, n_rate each), a, b (n_rate), nat (as launch_spmv_panel's)].  Same values as launch_optic_substep.
 * Replaces a launch_csr + launch_optic pair (flyverse/optic.py:474-483 then 525-551). */
int launch_optic_substep_panel(void** p, int n_rate, int nblk, int batch, float gr, float ga, float aad,
                               void* stream);

/* The frame's optic inputs (flyverse/optic.py:496-517 photoreceptor stage, :572 feedback, :579-580 the held
 * inputs) in two launches: the photoreceptor cells (as launch_photoreceptor), s = clamp(rate[:, spk_idx] / 100, 0, 3)
 * and the first substep's dr = clip(v + b, 0, 1) - b in one; then (products != 0) pr_in = gain_in * W_rp a_pr and
 * spk_in = gain_fb * W_rs s (grouped storage and order of launch_spmv_grouped).
 * p = [rad (batch, n_col, 4), sens, pr_col, cell_ptr, cell_pr, cell_w, I_lp, I_mean, fresh uint8 (batch), coefs (6),
 * contrast, a_pr (batch, n_pr), rate (batch, n), spk_idx int32 (n_spk), s (batch, n_spk), v (batch, n_rate),
 * b (n_rate), dr (batch, n_rate), W_rp group_row, chunk_ptr, pidx, pw, W_rs group_row, chunk_ptr, pidx, pw, pr_in,
 * spk_in (batch, n_rate)]; dims = [n_col, n_pr, n, n_spk, n_rate, batch, W_rp groups, W_rs groups, products].
 * fresh is read here and cleared by launch_frame_out. */
int launch_frame_in(void** p, const int* dims, float gain_in, float gain_fb, void* stream);

/* The frame's optic output in one launch (flyverse/optic.py:599-604): drive[:, spk_idx] = clamp(gain_out * W_sr dr,
 * -clip, clip) with dr = clip(v + b, 0, 1) - b = rates() - r0.  p = [W_sr group_row, chunk_ptr, pidx, pw, v, b, dr
 * (batch, n_rate), spk_idx, drive (batch, n), fresh uint8 (batch), dr_src (batch, n_rate)]; dims = [n, n_spk, n_rate,
 * batch, W_sr groups, ready]: ready when dr_src already holds that value (the last substep's output; copied into dr
 * when it is another buffer), else the launch forms it from v and stores it in dr.  Clears fresh. */
int launch_frame_out(void** p, const int* dims, float gain_out, float clip, void* stream);

/* launch_lif_rng that also writes the transmitted 0/1 spikes as packed bits: p[17] = uint64 words (batch, ceil(n/64)),
 * bit i % 64 of word i / 64 set when cell i fired (no short-term depression).  With launch_pull_bits it replaces
 * the delay-ring write spike_buf[pos] = spike of flyverse/brain.py:864-865 for the plain single-fly path. */
int launch_lif_bits(void** p, int n, int batch, int flags, uint64_t seed, uint32_t step, void* stream);

/* launch_pull over packed spike bits: p = [ptr(n+1), idx, w, bits uint64 (batch, ceil(n/64)), g (batch, n),
 * sched (n: rows with in-edges first, longest first, then the rest)]; n_full = rows with in-edges; grid =
 * workgroups of the persistent launch.  Replaces g.mul_(a_s) + event_scatter, flyverse/brain.py:875-876. */
int launch_pull_bits(void** p, int n, int n_full, int batch, float a_s, int flags, int grid, void* stream);

/* K <= n_delay LIF steps of the plain (batch, n) brain in one window, from global step *step_base + step (step_base a
 * device uint32 or null).  Replaces K iterations of Brain._step_cuda (flyverse/brain.py:873-881: g.mul_(a_s),
 * cuda.event_scatter, torch.rand, cuda.lif_update), with the same additions and updates as K steps of launch_pull +
 * launch_lif_rng.  Spikes are read from 64-step history words (bit t % 64 of hist[b, c]: cell c of fly b fired at
 * step t), so each in-edge's weight and presynaptic word is read once per window; the spikes of every step are also
 * written into the delay ring spike_buf from slot slot0 on.  p = [ptr (n+1), idx, w (float), sched (ceil(n/8)
 * blocks of 8 cells), v, g, refrac, rate, counts, adapt, spikes, drive, pois, active (n), spike_buf (delay, batch,
 * n), hist uint64 (batch, n), P (11), long_rows (n_long: every cell with more than 1024 in-edges, most first)];
 * dims = [n, batch, K, delay, slot0, n_long, n_vlong (the first n_vlong long cells have more than 2048 in-edges),
 * flags (1 Poisson, 4 counts, 8 adaptation)].  Requires 1 <= K <= delay, 2 * delay + K <= 64 and 0 <= slot0 < delay. */
int launch_lif_window(void** p, const int* dims, uint64_t seed, float a_s, const uint32_t* step_base, uint32_t step,
                      void* stream);

/* ------------------------------------------------------------------ batched engine: 64 flies, fly-minor (rows, 64)
 * State is (rows, 64): lane b of row r is fly b's copy of cell r.  Spikes are 64-bit masks (one bit per fly) in a
 * ring of ring_slots >= 2 * n_delay entries.  One launch runs K <= n_delay LIF steps of every row.  Together these
 * replace BatchSim.step / FlyBrain._graph_frame over B flies (flyverse/batch_sim.py:203-254, flyverse/fly.py:509-610)
 * with per-fly step counters, so rollouts of uneven length share one captured graph. */

/* K LIF steps (global steps *engine_step + step_offset ...).  p = [ptr, pre_off, w, sched, neuron_id, row_flags, v,
 * g, refrac, rate, counts, drive, pois, ring, seed, fly_step, budget, engine_step, P(12), act];
 * dims = [n_rows, lanes, ring_slots, delay, record, grid, fold mode 0-6, row_major ring].  Every fold mode performs
 * the same additions in the same order.  Mode 6 (slot-major ring, K <= 8) first writes one activity byte per row into
 * act (bit k: the mask step k reads is nonzero) and then skips the mask gathers of the presynaptic cells that fired
 * in no fly; act holds (n_rows rounded up to a multiple of 8) + 8 bytes, the last 8 zero.  Other modes ignore p[19]. */
int launch_batch_lif(void** p, const int* dims, int step_offset, int K, void* stream);

/* y (rows, 64) = scale * W x (cols, 64), chunked order of launch_csr.  p = [ptr, idx, w, sched, x, y].
 * Replaces the per-frame optic products W_rp@a_pr, W_rs@s, W_sr@dr (flyverse/optic.py:579-580, 602). */
int launch_spmm_fm(void** p, int rows, float scale, int grid, void* stream);

/* One optic substep for 64 flies: y = W_rr dr_cur fused with the update of launch_optic; writes dr_next.
 * p = [ptr, idx, w, sched, dr_cur, dr_next, v, adapt, pr, spk, a, b, live_mask].  Replaces OpticLobe._substep
 * (flyverse/optic.py:525-551). */
int launch_optic_substep_fm(void** p, int rows, float gr, float ga, float aad, int grid, void* stream);

/* dr = clip(v + b, 0, 1) - b on (rows, 64).  p = [v, b, dr].  launch_optic_dr in the fly-minor layout. */
int launch_optic_dr_fm(void** p, int rows, void* stream);

/* Photoreceptor stage with the movie drawn in-kernel (Philox, stream 1, counter column * 4 + channel).  p = [sens,
 * pr_col, cell_ptr, cell_pr, cell_w, I_lp (5 n_col, 64), I_mean, a_pr (n_pr, 64), movie_seed, frame, fresh_mask,
 * live_mask, coefs]; n_cells = 5 n_col.  launch_photoreceptor in the fly-minor layout. */
int launch_photoreceptor_fm(void** p, int n_cells, void* stream);

/* s = clamp(rate / 100, 0, 3) for the spiking rows.  p = [rate (rows, 64), spk_row (n_spk), s (n_spk, 64)]. */
int launch_feedback_fm(void** p, int n_spk, void* stream);

/* drive[spk_row] = clamp(gain_out * y_sr, -clip, clip).  p = [y_sr (n_spk, 64), spk_row, drive (rows, 64)]. */
int launch_drive_out_fm(void** p, int n_spk, float gain_out, float clip, void* stream);

/* launch_frame_in in the fly-minor layout: launch_photoreceptor_fm, launch_feedback_fm and the first dr in one launch;
 * then (grid > 0) pr_in = gain_in * W_rp a_pr and spk_in = gain_fb * W_rs s, each streamed by grid workgroups
 * (flyverse/optic.py:496-517, 572, 579-580).  p = [the 13 pointers of launch_photoreceptor_fm, rate (rows, 64),
 * spk_row, s (n_spk, 64), v, b, dr (n_rate, 64), W_rp ptr, idx, w, sched, W_rs ptr, idx, w, sched, pr_in, spk_in
 * (n_rate, 64)]; dims = [n_cells, n_spk, n_rate, grid]. */
int launch_frame_in_fm(void** p, const int* dims, float gain_in, float gain_fb, void* stream);

/* drive[spk_row] = clamp(gain_out * W_sr dr, -clip, clip) in one launch, dr (n_rate, 64) the last substep's output
 * (flyverse/optic.py:599-604).  p = [W_sr ptr, idx, w, sched, dr, spk_row, drive (rows, 64)]. */
int launch_frame_out_fm(void** p, int n_spk, float gain_out, float clip, int grid, void* stream);

/* End of a frame: engine step += steps, live flies' step / budget / frame counters advance.
 * p = [engine_step, fly_step, budget, frame, fresh_mask, live_mask]. */
int launch_frame_end(void** p, int lanes, int steps, void* stream);

/* Reset the lanes in `mask` to rest (a fly finished and a new rollout takes its lane; flyverse/brain.py:719-733,
 * flyverse/fly.py:682-707) and set the live mask to live_set.  p = [v, g, refrac, rate, counts, drive, pois,
 * optic v, optic adapt, ring, fresh_mask, live_mask]. */
int launch_reset_lanes(void** p, uint64_t mask, uint64_t live_set, float v_rest, int n_rows, int n_rate,
                       int ring_slots, void* stream);

/* ------------------------------------------------------------------ packed uneven subgraphs
 * Many LIF graphs packed block-diagonally into one CSR with offset tables (instance / local index per cell, one
 * Philox key per instance).  The reference would run one Brain per Connectome.subset (flyverse/connectome.py:469-481).
 * p = [ptr, idx, w, sched (ceil(n/64) blocks of 64 cells, visiting order), v, g, refrac, rate, counts, drive, pois,
 * inst, local, seeds uint64, ring uint64 (n spike history words: bit t % 64 of word c = cell c fired at step t),
 * bar (2 uint32), error (4 int32), P(12)]. */

/* `steps` steps from step t0 in one cooperative launch; dims = [n, steps, window, delay, grid, t0, barrier bound ms];
 * window <= delay, 2*delay + window <= 64, grid <= packed_max_grid. */
int launch_packed_run(void** p, const int* dims, void* stream);

/* One window of K <= window steps from step *t_base + off (t_base: device int32, so a captured graph of windows
 * replays at advancing steps); p continues with [long_rows (n_long: cells with more than 512 in-edges, most first),
 * long_id (n: index in long_rows or -1), long_g (n_long * 21 floats, scratch)]; dims = [n, K, window, delay, grid,
 * off, n_long, n_vlong] (the first n_vlong long cells have more than 2048 in-edges; n_long = 0: no long-cell kernels,
 * those cells are streamed in the window kernel). */
int launch_packed_window(void** p, const int* dims, const int* t_base, void* stream);

/* launch_packed_window in parts, same arguments: parts bit 1 = the very long cells' chains, bit 2 = the long cells'
 * chains, bit 4 = the window kernel (7 = launch_packed_window).  The two chain kernels are independent, so a caller
 * can run them on two streams; the window kernel reads what both wrote and must follow them. */
int launch_packed_window_parts(void** p, const int* dims, const int* t_base, int parts, void* stream);

/* *t_base += steps on the stream. */
int launch_packed_advance(int* t_base, int steps, void* stream);

/* The largest grid launch_packed_run accepts. */
int packed_max_grid(int* out);

/* ------------------------------------------------------------------ optic-lobe adjoint (32 movies, fly-minor)
 * Forward and hand-written backward of the optic rate network for fitting per-type tau, per-type bias and type-pair
 * gains.  The reference has no adjoint: torch autograd through OpticLobe._step_frame_grad (flyverse/optic.py:606-676)
 * with tau_by_type / baseline_by_type / pair_gain as the parameters (flyverse/optic.py:82-87, 251-267, 343-347). */

/* a = exp(-dt / tau[type]), b = bias[type].  p = [tau, bias, unit_type (n), a (n), b (n)]. */
int launch_adj_params(void** p, int n, float dt, void* stream);

/* w = gain[pair_bin] * w0.  p = [gain (types*types), pair_bin (nnz), w0 (nnz), w (nnz)]. */
int launch_adj_weights(void** p, int64_t nnz, void* stream);

/* dr = clip(v + b, 0, 1) - b on (n, 32).  p = [v, b, dr]. */
int launch_adj_dr(void** p, int n, void* stream);

/* One forward step.  p = [ptr, idx, sched, w, a, b, v, adapt, dr, P, v1, adapt1, dr1, inp|null, readout,
 * target|null, loss|null]; f = [g_rr, g_ad, a_ad]. */
int launch_adj_forward(void** p, int rows, const float* f, int grid, void* stream);

/* Adjoint through v_t+1 = inp + (v - inp)*a: ly = g_rr * lv * (1 - a) (the adjoint of y = W dr), and
 * ga[r] += sum over movies of lv * (v - inp) (the partial for a_r, in a fixed butterfly order).
 * p = [lv, v, inp, a, ly, ga]. */
int launch_adj_back_ly(void** p, int rows, float g_rr, void* stream);

/* Adjoint of one step through the transpose CSR (W^T ly) and the clip gates.  p = [ptrT, idxT, schedT, wT, a, b, ly,
 * v, dr, lv, la, readout, target|null, gb, eg|null, ly_next|null, inp|null, ga|null]; f = [g_ad, a_ad, loss_scale,
 * g_rr].  With eg (16 floats per chunk of the chunk-padded W^T) each real entry also gets eg += sum over movies of
 * ly[row] * dr[col], from the same gathers.  With inp (step t's inputs) the launch also does launch_adj_back_ly's ga
 * term for step t, and with ly_next (a buffer other than ly) launch_adj_back_ly's ly for step t - 1 from the new lv,
 * so a backward step needs this launch alone. */
int launch_adj_back_rows(void** p, int rows, const float* f, int grid, void* stream);

/* out[s] = fixed-order sum over segment s of vals[perm[j]] (* mult[perm[j]]).  p = [vals, mult|null, perm,
 * seg (nseg+1), out]. */
int launch_segment_sum(void** p, int nseg, void* stream);

/* out[0] = scale * fixed-order sum of n partial losses.  p = [parts, out]. */
int launch_loss_total(void** p, int64_t n, float scale, void* stream);

/* A bandwidth probe, not part of the brain step and with no counterpart in the reference: y = x for 4 * n4 floats
 * (16-byte aligned, not overlapping) by a streaming copy with non-temporal loads and stores, `grid` workgroups of
 * 256 threads.  bench.py times it to state the copy bandwidth this GPU attains beside the vendor peak. */
int launch_copy_stream(const void* x, void* y, int64_t n4, int grid, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* FLYHIP_H */
