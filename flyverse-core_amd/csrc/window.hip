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
using fh_chain::kSlots;
using fh_chain::Window;

enum { kPoisson = 1, kCounts = 4, kAdapt = 8 };

struct Plain {
    const int* ptr;          // (n + 1) post-major CSR shared by the flies, columns ascending
    const int* idx;          // (nnz) presynaptic cell
    const float* w;          // (nnz)
    const int* sched;        // (ceil(n / 8)) blocks of 8 cells, most short-cell in-edges first
    float *v, *g, *refrac, *rate, *counts, *adapt, *spikes;   // (B, n)
    const float* drive;      // (B, n)
    const float* pois;       // (B, n) per-step Poisson probability
    const float* active;     // (n)
    float* ring;             // (n_delay, B, n) spike_buf: the slot of step t holds the spikes emitted at step t
    uint64_t* hist;          // (B, n) spike history words
    const float* P;          // (11) v_rest, v_reset, v_th, a_m, dt, t_ref, adapt_jump, a_ad, -, a_r, rate_gain
    const int* long_rows;    // (n_long) cells with more than 1024 in-edges, most first
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

// A window is one launch of independent work items (no item waits on another):
//  long cells  one workgroup per (fly, cell with more than 1024 in-edges): the cell's chain over the four wavefronts
//              (fh_chain::g_chain_compact: each step's fired weights compacted in edge order by the wavefronts in
//              parallel, folded by wavefront 0 with the lane fold), then its LIF update on wavefront 0;
//  blocks      one wavefront per (fly, block of kBlock consecutive cells), blocks in `sched` order (most in-edges
//              first).  Phase A: for each short cell of the block with in-edges, lanes over its in-edges (edges and
//              window bits in registers, the register count fitted to the cell's chunks), the chain with each step's
//              fired weights added in edge order from the step's ballot; g_k into LDS.  Phase B: lanes over the block's
//              cells, the LIF updates (a cell without in-edges only decays: g_k = g_{k-1} * a_s).  Long cells are left
//              to their own workgroups, which come first in the grid.
// A cell's history word and state are written by the one wavefront that owns it; the bits other wavefronts read
// from it (steps T0 - delay .. T0 - delay + K - 1) are not among those it rewrites (T0 .. T0 + K - 1).
constexpr int kBlock = 8;
// Cells with more in-edges than this take a workgroup of their own (the host's long_rows).  One wavefront walking up
// to 16 chunks in registers beats the four-wavefront compaction up to here: B=1 frame 1.44 ms at 512, 1.17 at 1024,
// 1.27 at 1536 (24 chunks' registers lower the whole kernel's occupancy).
constexpr int kWinShort = 16 * kWave;
constexpr int kVlongCompact = 24;              // chunks per wavefront of a very long cell (up to 6144 in-edges)
constexpr int kLongCompact = 8;               // chunks per wavefront of a long cell (up to 2048 in-edges)

// The chain of one short cell with its edges in registers, NC fitted to its chunk count.  At a few fired in-edges
// per step, walking each step's ballot (chain_steps_bits) is cheaper than compacting it through LDS for the lane
// fold (chain_steps_fold): B=1 frame 1.67 -> 1.47 ms.
template <int NC>
__device__ __forceinline__ float short_chain(const Edges& E, const Window& W, int e0, int e1, float g0, int lane) {
    float wc[NC];
    uint32_t hist[NC];
    fh_chain::load_window_edges<NC>(E, W, e0, e1, lane, wc, hist);
    return fh_chain::chain_steps_bits<NC>(W, e0, e1, g0, lane, wc, hist);
}

// The LDS of a workgroup: a long cell's four compaction lists, or a block wavefront's g_k table.
constexpr int kLongLds = 4 * (kVlongCompact * kWave + kWave);
constexpr int kBlockLds = kBlock * kMaxWindow;
static_assert(4 * kBlockLds <= kLongLds, "the block wavefronts' LDS fits in the long cells' buffer");

// One launch per window: workgroups [0, long_wgs) take one (fly, long cell) each, the rest four blocks each.
__global__ __launch_bounds__(256) void k_win(Plain a, const uint32_t* step_base, uint32_t step, int K, int long_wgs) {
    __shared__ float lds[kLongLds];
    __shared__ int counts[4];
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t T0 = step + (step_base ? *step_base : 0u);
    const Window W(T0, K, a.delay, a.a_s);
    if ((int)blockIdx.x < long_wgs) {
        const int fly = blockIdx.x % a.batch, i = blockIdx.x / a.batch;
        const int r = a.long_rows[i];
        const int e0 = a.ptr[r], e1 = a.ptr[r + 1];
        const float g0 = a.g[(int64_t)fly * a.n + r];
        const Edges E = edges_of(a, fly);
        float gk;
        if (e1 - e0 <= 4 * kLongCompact * kWave)
            gk = fh_chain::g_chain_compact<kLongCompact>(E, W, e0, e1, g0, lane, wv, lds, counts);
        else if (e1 - e0 <= 4 * kVlongCompact * kWave)
            gk = fh_chain::g_chain_compact<kVlongCompact>(E, W, e0, e1, g0, lane, wv, lds, counts);
        else
            gk = wv == 0 ? fh_chain::g_chain_stream(E, W, e0, e1, g0, lane) : 0.0f;
        if (wv == 0)
            lif_window_cell(a, T0, K, fly, r, g0, [&](int k, float) { return fh_readlane(gk, k); }, lane == 0);
        return;
    }
    const int64_t it = (int64_t)(blockIdx.x - long_wgs) * 4 + wv;
    const int blocks = (a.n + kBlock - 1) / kBlock;
    if (it >= (int64_t)blocks * a.batch) return;
    const int fly = (int)(it % a.batch);
    const int c0 = a.sched[it / a.batch] * kBlock;
    const int nc = min(kBlock, a.n - c0);
    const Edges E = edges_of(a, fly);
    float* const gb = lds + wv * kBlockLds;
    const int pl0 = a.ptr[c0 + min(lane, nc)], pl1 = a.ptr[c0 + min(lane + 1, nc)];
    const bool is_long = lane < nc && pl1 - pl0 > kWinShort;
    const float gl = lane < nc ? a.g[(int64_t)fly * a.n + c0 + lane] : 0.0f;
    // ---------------------------------------------------------------- phase A: chains of the short cells
    uint64_t todo = __builtin_amdgcn_ballot_w64(lane < nc && pl1 > pl0 && !is_long);
    while (todo) {
        const int i = __builtin_ctzll(todo);
        todo &= todo - 1;
        const int e0 = __builtin_amdgcn_readlane(pl0, i), e1 = __builtin_amdgcn_readlane(pl1, i);
        const float g0 = fh_readlane(gl, i);
        const int nch = (e1 - e0 + kWave - 1) / kWave;
        float gk;
        if (nch <= 1)
            gk = short_chain<1>(E, W, e0, e1, g0, lane);
        else if (nch <= 2)
            gk = short_chain<2>(E, W, e0, e1, g0, lane);
        else if (nch <= 4)
            gk = short_chain<4>(E, W, e0, e1, g0, lane);
        else if (nch <= 8)
            gk = short_chain<8>(E, W, e0, e1, g0, lane);
        else
            gk = short_chain<16>(E, W, e0, e1, g0, lane);
        if (lane < K) gb[i * kMaxWindow + lane] = gk;
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
// p = [ptr (n+1), idx, w (float), sched (ceil(n/8) blocks of 8 cells), v, g, refrac, rate, counts, adapt, spikes,
//      drive, pois, active (n), spike_buf (delay, batch, n), hist uint64 (batch, n), P (11), long_rows (n_long: every
//      cell with more than 1024 in-edges, most first, the n_vlong with more than 2048 first)];
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
    const int64_t long_wgs = (int64_t)a.n_long * a.batch;
    const int64_t block_wgs = (((int64_t)(a.n + kBlock - 1) / kBlock) * a.batch + 3) / 4;
    if (long_wgs >= (1ll << 31) || block_wgs >= (1ll << 31)) return FH_ERR_SIZE;
    if (long_wgs + block_wgs >= (1ll << 31)) return FH_ERR_SIZE;
    hipLaunchKernelGGL(k_win, dim3((unsigned)(long_wgs + block_wgs)), dim3(256), 0, (hipStream_t)stream, a, step_base,
                       step, K, (int)long_wgs);
    FH_LAUNCH_CHECK();
}
