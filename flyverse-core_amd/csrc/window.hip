// The plain (B, N) brain stepped a window of K <= n_delay steps per launch: the single-fly path (and any plain batch
// of flies sharing one graph).
//
// The reference advances one step at a time: g *= a_s, the synaptic input of the spikes emitted n_delay steps ago in
// ascending presynaptic order (flyverse/brain.py:832-833, 679-692), then the LIF update (brain.py:838-868).  The
// input of a step only reads spikes at least n_delay steps old, so within a window of K <= n_delay steps every
// synaptic input is known before the window starts.  Each cell keeps a 64-bit spike history word (bit t mod 64: the
// cell fired at step t), and a window computes, per cell, first the synaptic chain g_1..g_K of the window (fh_chain:
// lanes over the cell's in-edges, each in-edge's weight and presynaptic history word read once per window instead of
// once per step), then K LIF updates with those g_k in the operation order of brain.py:838-868, the Poisson uniform
// drawn in-kernel (Philox keyed by the brain's seed, counter (cell, step, fly, 0), the draws of launch_lif_rng); the
// spike of every step goes into the delay ring spike_buf, the window's bits into the cell's own history word (one
// plain store), and the state is stored once.  Every step's additions and updates are those of launch_pull +
// launch_lif_rng, so the results are the same bits.
#include "chain.h"

namespace {

using fh_chain::Edges;
using fh_chain::kMaxWindow;
using fh_chain::kShortEdges;
using fh_chain::kSlots;
using fh_chain::Window;
constexpr int kVlongChunks = 36;        // chunks per wavefront for very long cells (up to 9216 in-edges in registers)

enum { kPoisson = 1, kCounts = 4, kAdapt = 8 };

struct Plain {
    const int* ptr;          // (n + 1) post-major CSR shared by the flies, columns ascending
    const int* idx;          // (nnz) presynaptic cell
    const float* w;          // (nnz)
    const int* sched;        // (ceil(n / 16)) blocks of 16 cells, most short-cell in-edges first
    float *v, *g, *refrac, *rate, *counts, *adapt, *spikes;   // (B, n)
    const float* drive;      // (B, n)
    const float* pois;       // (B, n) per-step Poisson probability
    const float* active;     // (n)
    float* ring;             // (n_delay, B, n) spike_buf: the slot of step t holds the spikes emitted at step t
    uint64_t* hist;          // (B, n) spike history words
    const float* P;          // (11) v_rest, v_reset, v_th, a_m, dt, t_ref, adapt_jump, a_ad, -, a_r, rate_gain
    const int* long_rows;    // (n_long) cells with more than 512 in-edges, most first
    uint64_t seed;
    int n, batch, delay, slot0, n_long, n_vlong, flags;
    float a_s;
};

__device__ __forceinline__ Edges edges_of(const Plain& a, int fly) {
    return Edges{a.idx, a.w, a.hist + (int64_t)fly * a.n};
}

// The LIF update of one cell over the window, brain.py:838-868, with g_k = gk(k) (the cell's synaptic chain), in the
// operation order of launch_lif_rng; spikes go into the delay ring, the window's bits into the cell's history word.
// `store` selects the lane that writes (every lane computes).
template <class G>
__device__ __forceinline__ void lif_window_cell(const Plain& a, uint32_t T0, int K, int fly, int c, float g, G gk,
                                                bool store) {
    const float v_rest = a.P[0], v_reset = a.P[1], v_th = a.P[2], am = a.P[3], dt = a.P[4], t_ref = a.P[5];
    const float jump = a.P[6], aad = a.P[7], ar = a.P[9], rg = a.P[10];
    const int64_t q = (int64_t)fly * a.n + c;
    float v = a.v[q], ref = a.refrac[q], rate = a.rate[q], cnt = a.counts[q], ad = a.adapt[q];
    const float drive = a.drive[q], p = a.pois[q], act = a.active[c];
    const uint64_t old_hist = a.hist[q];
    uint64_t mine = 0, fired_bits = 0;
    float spike = 0.0f;
    int slot = a.slot0;
    for (int k = 0; k < K; ++k) {
        const uint32_t t = T0 + (uint32_t)k;
        const float acc = gk(k, g);
        float target = v_rest + acc;                       // brain.py:572: v_rest + g + drive - adapt
        target = target + drive;
        target = target - ad;
        float dv = v - target;
        dv = dv * am;
        float vn = target + dv;                            // brain.py:840
        float rn = ref;
        if (rn > 0.0f) vn = v_reset;                       // brain.py:841-842
        rn = rn - dt;                                      // brain.py:843
        if (rn < 0.0f) rn = 0.0f;
        spike = (vn >= v_th) ? 1.0f : 0.0f;                // brain.py:845
        spike = spike * act;
        if ((a.flags & kPoisson) && p > 0.0f && fh_uniform(a.seed, t, (uint32_t)c, (uint32_t)fly, 0u) < p)
            spike = fmaxf(spike, 1.0f);                    // brain.py:847-849
        const bool fired = spike > 0.0f;
        if (fired) { vn = v_reset; rn = t_ref; }           // brain.py:850-852
        if (a.flags & kAdapt) {                            // brain.py:853-856
            const float d = ad * aad;
            const float j = spike * jump;
            ad = d + j;
        }
        if (a.flags & kCounts) cnt = cnt + spike;          // brain.py:866-867
        const float rr = rate * ar;                        // brain.py:868
        const float rs = spike * rg;
        g = acc; v = vn; ref = rn; rate = rr + rs;
        if (store) a.ring[((int64_t)slot * a.batch + fly) * a.n + c] = spike;     // brain.py:864
        slot = slot + 1 == a.delay ? 0 : slot + 1;
        const uint64_t bit = 1ull << (t & (kSlots - 1));
        mine |= bit;
        if (fired) fired_bits |= bit;
    }
    if (store) {
        a.v[q] = v; a.g[q] = g; a.refrac[q] = ref; a.rate[q] = rate; a.counts[q] = cnt; a.adapt[q] = ad;
        a.spikes[q] = spike;
        a.hist[q] = (old_hist & ~mine) | fired_bits;
    }
}

// One launch per window, three kinds of independent workgroups (nothing waits on anything):
//   [0, n_vlong * batch)            one very long cell (more than 2048 in-edges) of one fly per workgroup: its chain
//                                   over the four wavefronts (fh_chain::g_chain_compact: each step's fired weights
//                                   compacted in parallel, folded in edge order by wavefront 0), then its LIF update;
//   next ceil(long_items / 4)        one long cell (513-2048 in-edges) of one fly per wavefront: chain with the edges
//                                   in registers, then its LIF update;
//   the rest                         one block of kBlock consecutive cells of one fly per wavefront (blocks in `sched`
//                                   order, most in-edges first).  Phase A: the block's short cells' in-edges form one
//                                   contiguous range of the CSR; it is staged through LDS kSeg edges at a time (all
//                                   index, weight and history loads of a segment in flight together), and the chain of
//                                   each cell in the segment runs from LDS, g_k into LDS.  Phase B: lanes over the
//                                   block's cells, the LIF updates (a cell without in-edges only decays: g_k =
//                                   g_{k-1} * a_s).  Long cells inside a block are left to their own workgroups.
// A cell's history word and state are written by the one wavefront that owns it; the bits other wavefronts read
// from it (steps T0 - delay .. T0 - delay + K - 1) are not among those it rewrites (T0 .. T0 + K - 1).
constexpr int kBlock = 16;
constexpr int kSegChunks = 16;                 // a staged segment: 16 chunks of 64 in-edges
constexpr int kSeg = kSegChunks * kWave;
constexpr int kVlongCompact = 24;              // chunks per wavefront of a very long cell (up to 6144 in-edges)

// Stage in-edges [s0, s1) (at most kSeg) of one fly into LDS: window bits and weights, one entry per edge.
__device__ __forceinline__ void stage_segment(const Edges& E, const Window& W, int s0, int s1, int lane,
                                              uint32_t* sbits, float* sw) {
    int pre[kSegChunks];
    float wv[kSegChunks];
#pragma unroll
    for (int u = 0; u < kSegChunks; ++u) {
        const int e = s0 + u * kWave + lane;
        const bool val = s0 + u * kWave < s1 && e < s1;
        pre[u] = val ? E.idx[e] : -1;
        wv[u] = val ? E.w[e] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < kSegChunks; ++u) {
        const uint32_t b = pre[u] >= 0 ? W.bits(E.hist[pre[u]]) : 0u;
        if (s0 + u * kWave < s1) {
            sbits[u * kWave + lane] = b;
            sw[u * kWave + lane] = wv[u];
        }
    }
}

// The chain of one staged cell: its edges are entries [o0, o1) of the segment (at most 8 chunks).
__device__ __forceinline__ float chain_from_lds(const Window& W, int o0, int o1, float g0, int lane,
                                                const uint32_t* sbits, const float* sw) {
    constexpr int NC = kShortEdges / kWave;
    float wc[NC];
    uint32_t hist[NC];
    const int nch = (o1 - o0 + kWave - 1) / kWave;
#pragma unroll
    for (int u = 0; u < NC; ++u) {
        const int o = o0 + u * kWave + lane;
        const bool val = u < nch && o < o1;
        hist[u] = val ? sbits[o] : 0u;
        wc[u] = val ? sw[o] : 0.0f;
    }
    return fh_chain::chain_steps_bits<NC>(W, o0, o1, g0, lane, wc, hist);
}

__global__ __launch_bounds__(256) void k_win(Plain a, const uint32_t* step_base, uint32_t step, int K, int vlong_wgs,
                                             int long_wgs) {
    __shared__ float gbuf[4][kBlock * kMaxWindow];
    __shared__ uint32_t stage[4 * 2 * kSeg];       // short blocks: per wavefront kSeg bits + kSeg weights
    __shared__ int counts[4];
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t T0 = step + (step_base ? *step_base : 0u);
    const Window W(T0, K, a.delay, a.a_s);
    const int wg = blockIdx.x;
    if (wg < vlong_wgs) {
        const int fly = wg % a.batch, i = wg / a.batch;
        const int r = a.long_rows[i];
        const int e0 = a.ptr[r], e1 = a.ptr[r + 1];
        const float g0 = a.g[(int64_t)fly * a.n + r];
        const Edges E = edges_of(a, fly);
        float gk;
        if (e1 - e0 <= 4 * kVlongCompact * kWave) {
            gk = fh_chain::g_chain_compact<kVlongCompact>(E, W, e0, e1, g0, lane, wv, (float*)stage, counts);
        } else {
            gk = wv == 0 ? fh_chain::g_chain_stream(E, W, e0, e1, g0, lane) : 0.0f;
        }
        if (wv == 0)
            lif_window_cell(a, T0, K, fly, r, g0, [&](int k, float) { return fh_readlane(gk, k); }, lane == 0);
        return;
    }
    if (wg < vlong_wgs + long_wgs) {
        const int64_t it = (int64_t)(wg - vlong_wgs) * 4 + wv;
        if (it >= (int64_t)(a.n_long - a.n_vlong) * a.batch) return;
        const int fly = (int)(it % a.batch), i = a.n_vlong + (int)(it / a.batch);
        const int r = a.long_rows[i];
        const int e0 = a.ptr[r], e1 = a.ptr[r + 1];
        const float g0 = a.g[(int64_t)fly * a.n + r];
        const Edges E = edges_of(a, fly);
        const float gk = e1 - e0 <= 16 * kWave ? fh_chain::g_chain_cached<16>(E, W, e0, e1, g0, lane)
                                               : fh_chain::g_chain_cached<32>(E, W, e0, e1, g0, lane);
        lif_window_cell(a, T0, K, fly, r, g0, [&](int k, float) { return fh_readlane(gk, k); }, lane == 0);
        return;
    }
    const int64_t it = (int64_t)(wg - vlong_wgs - long_wgs) * 4 + wv;
    const int blocks = (a.n + kBlock - 1) / kBlock;
    if (it >= (int64_t)blocks * a.batch) return;
    const int fly = (int)(it % a.batch);
    const int c0 = a.sched[it / a.batch] * kBlock;
    const int nc = min(kBlock, a.n - c0);
    const Edges E = edges_of(a, fly);
    float* const gb = gbuf[wv];
    uint32_t* const sbits = stage + wv * 2 * kSeg;
    float* const sw = (float*)(sbits + kSeg);
    const int ll = min(lane, nc);
    const int pl0 = a.ptr[c0 + ll], pl1 = a.ptr[c0 + min(lane + 1, nc)];
    const bool is_long = lane < nc && pl1 - pl0 > kShortEdges;
    const float gl = lane < nc ? a.g[(int64_t)fly * a.n + c0 + lane] : 0.0f;
    // ---------------------------------------------------------------- phase A: chains of the short cells
    // cells with in-edges that are not long, in order; segments of consecutive such cells of at most kSeg edges
    uint64_t todo = __builtin_amdgcn_ballot_w64(lane < nc && pl1 > pl0 && !is_long);
    while (todo) {
        const int i0 = __builtin_ctzll(todo);
        const int s0 = __builtin_amdgcn_readlane(pl0, i0);
        // the segment: cells i0, i0 + 1, ... up to the first long cell or the first that ends beyond kSeg edges
        // (cells without in-edges inside it add nothing to the range)
        const uint64_t fits = __builtin_amdgcn_ballot_w64(lane >= i0 && lane < nc && pl1 - s0 <= kSeg && !is_long);
        const uint64_t brk = ~fits & ~((1ull << i0) - 1);
        const int i1 = brk ? __builtin_ctzll(brk) : 64;
        const int last = min(i1, nc) - 1;
        const int s1 = __builtin_amdgcn_readlane(pl1, last);
        stage_segment(E, W, s0, s1, lane, sbits, sw);
        __builtin_amdgcn_wave_barrier();
        for (int i = i0; i <= last; ++i) {
            const int e0 = __builtin_amdgcn_readlane(pl0, i), e1 = __builtin_amdgcn_readlane(pl1, i);
            if (e1 == e0) continue;
            const float gk = chain_from_lds(W, e0 - s0, e1 - s0, fh_readlane(gl, i), lane, sbits, sw);
            if (lane < K) gb[i * kMaxWindow + lane] = gk;
        }
        __builtin_amdgcn_wave_barrier();
        todo &= ~((last + 1 < 64 ? (1ull << (last + 1)) : 0ull) - 1);
    }
    __builtin_amdgcn_wave_barrier();
    // ---------------------------------------------------------------- phase B: LIF, lane = cell c0 + lane
    if (lane < nc && !is_long) {
        const bool empty = pl1 == pl0;
        const float a_s = a.a_s;
        lif_window_cell(a, T0, K, fly, c0 + lane, gl,
                        [&](int k, float g) { return empty ? g * a_s : gb[lane * kMaxWindow + k]; }, true);
    }
}

}  // namespace

// K LIF steps of the plain (batch, n) brain from global step *step_base + step (step_base: device uint32 or null).
// p = [ptr (n+1), idx, w (float), sched (ceil(n/16) blocks of 16 cells), v, g, refrac, rate, counts, adapt, spikes,
//      drive, pois, active (n), spike_buf (delay, batch, n), hist uint64 (batch, n), P (11), long_rows (n_long: every
//      cell with more than 512 in-edges, most first, the n_vlong with more than 2048 first)];
// dims = [n, batch, K, delay, slot0 (spike_buf slot of the first step), n_long, n_vlong, flags (1 Poisson, 4 counts,
//         8 adaptation)].  Requires 1 <= K <= delay, 2 * delay + K <= 64 (so K <= 21) and 0 <= slot0 < delay.
FH_EXPORT int launch_lif_window(void** p, const int* dims, uint64_t seed, float a_s, const uint32_t* step_base,
                                uint32_t step, void* stream) {
    if (!p || !dims) return FH_ERR_ARGS;
    Plain a;
    a.ptr = (const int*)p[0]; a.idx = (const int*)p[1]; a.w = (const float*)p[2]; a.sched = (const int*)p[3];
    a.v = (float*)p[4]; a.g = (float*)p[5]; a.refrac = (float*)p[6]; a.rate = (float*)p[7]; a.counts = (float*)p[8];
    a.adapt = (float*)p[9]; a.spikes = (float*)p[10]; a.drive = (const float*)p[11]; a.pois = (const float*)p[12];
    a.active = (const float*)p[13]; a.ring = (float*)p[14]; a.hist = (uint64_t*)p[15]; a.P = (const float*)p[16];
    a.long_rows = (const int*)p[17];
    a.seed = seed; a.a_s = a_s;
    a.n = dims[0]; a.batch = dims[1];
    const int K = dims[2];
    a.delay = dims[3]; a.slot0 = dims[4];
    a.n_long = dims[5]; a.n_vlong = dims[6]; a.flags = dims[7];
    if (a.n < 0 || a.batch < 0 || K < 1 || K > a.delay || K > kMaxWindow || 2 * a.delay + K > kSlots ||
        a.slot0 < 0 || a.slot0 >= a.delay || a.n_long < 0 || a.n_long > a.n || a.n_vlong < 0 ||
        a.n_vlong > a.n_long || (a.flags & ~(kPoisson | kCounts | kAdapt)))
        return FH_ERR_ARGS;
    if (a.n_long > 0 && !a.long_rows) return FH_ERR_ARGS;
    if ((int64_t)a.n * a.batch == 0) return 0;
    if ((int64_t)a.n * a.batch >= (1ll << 31)) return FH_ERR_SIZE;
    const int64_t vlong_wgs = (int64_t)a.n_vlong * a.batch;
    const int64_t long_wgs = ((int64_t)(a.n_long - a.n_vlong) * a.batch + 3) / 4;
    const int64_t block_wgs = (((int64_t)(a.n + kBlock - 1) / kBlock) * a.batch + 3) / 4;
    const int64_t wgs = vlong_wgs + long_wgs + block_wgs;
    if (wgs >= (1ll << 31)) return FH_ERR_SIZE;
    hipLaunchKernelGGL(k_win, dim3((unsigned)wgs), dim3(256), 0, (hipStream_t)stream, a, step_base, step, K,
                       (int)vlong_wgs, (int)long_wgs);
    FH_LAUNCH_CHECK();
}
