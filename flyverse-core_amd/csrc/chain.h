// Synaptic chains over a window of steps from spike-history words (shared by the packed stepper and the plain
// window stepper).
//
// A history word per cell holds its last 64 steps: bit (t mod 64) says whether the cell fired (transmitted a 0/1
// spike) at step t.  The input of step t is the spikes of step t - delay (flyverse/brain.py:832-833, the delay ring),
// so for a window of K <= delay steps starting at T0, every bit a step reads was written before the window started:
// the synaptic chain of a cell over the whole window,
//     g_k = g_{k-1} * a_s + (weights of the in-edges whose presynaptic cell fired at step T0 + k - delay),
// added in edge order (ascending presynaptic index, the reference CPU order of flyverse/brain.py:679-692), needs
// nothing of the membrane and reads each in-edge's weight and history word once for the whole window.
#pragma once
#include "common.h"

namespace fh_chain {

constexpr int kSlots = 64;              // steps per history word
constexpr int kMaxWindow = 21;          // 2 * window + delay <= 64 with window <= delay
constexpr int kShortEdges = 8 * kWave;  // cells with more in-edges are "long" (their chains run in kernels of their own)

// Where a cell's in-edges and its presynaptic cells' history words are: post-major CSR columns (presynaptic cells,
// ascending within a row) and weights, and the history words of the fly the row belongs to.
struct Edges {
    const int* idx;
    const float* w;
    const uint64_t* hist;
};

// The window's view of the spike history: bit k of bits() = the cell fired at step T0 - delay + k (k < 32).
struct Window {
    uint32_t T0;
    int K, s0;
    float a_s;
    __device__ Window(uint32_t t0, int k, int delay, float as)
        : T0(t0), K(k), s0((int)((t0 - (uint32_t)delay) & (kSlots - 1))), a_s(as) {}
    __device__ __forceinline__ uint32_t bits(uint64_t w) const {
        return (uint32_t)(s0 ? (w >> s0) | (w << (kSlots - s0)) : w);
    }
};

// (((acc + x_0) + x_1) + ... + x_{n-1}) for the n <= 63 values in lanes 0..n-1 of x, every addition rounded as in
// a sequential loop, without moving values to scalar registers: a wavefront shift by one lane (DPP wave_shr:1) per
// round hands lane i's running value to lane i + 1, which adds x_i; lane 0 keeps acc (it adds -0.0, which leaves
// every value unchanged).  After round r lanes 0..r hold their final values, so after n rounds lane n holds the fold.
__device__ __forceinline__ float fold_lanes(float acc, float x, int n) {
    constexpr int kWaveShr1 = 0x138;
    const float xs = __int_as_float(
        __builtin_amdgcn_update_dpp(__float_as_int(-0.0f), __float_as_int(x), kWaveShr1, 0xf, 0xf, false));
    float y = acc;
    for (int i = 0; i < n; ++i) {
        const float t = __int_as_float(
            __builtin_amdgcn_update_dpp(__float_as_int(y), __float_as_int(y), kWaveShr1, 0xf, 0xf, false));
        y = t + xs;
    }
    return fh_readlane(y, n);
}

// The fired weights of one step, compacted in edge order into `list` (a wavefront's own LDS, at least nch * 64 + 64
// floats) and folded onto acc: per chunk one ballot, the lanes' prefix counts and one masked store, then fold_lanes
// over the list 63 values at a time.
template <int NC>
__device__ __forceinline__ float step_fold(float acc, int k, int nch, const float* wc, const uint32_t* hist, float* list,
                                           int lane) {
    int cnt = 0;
#pragma unroll
    for (int u = 0; u < NC; ++u) {
        if (u < nch) {
            const bool fired = (hist[u] >> k) & 1u;
            const uint64_t m = __builtin_amdgcn_ballot_w64(fired);
            if (fired)
                list[cnt + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u))] =
                    wc[u];
            cnt += __builtin_popcountll(m);
        }
    }
    for (int b0 = 0; b0 < cnt; b0 += 63) acc = fold_lanes(acc, list[b0 + lane], min(63, cnt - b0));
    return acc;
}

// The in-edges [e0, e1) (at most NC * 64) of one cell, lanes over edges: presynaptic cell and weight per chunk;
// lanes past the row end repeat its last entry (their bits are masked off by the caller).
template <int NC>
__device__ __forceinline__ void load_edges(const Edges& a, int e0, int e1, int lane, int* pre, float* wc) {
    const int last = max(e1 - 1, e0);
#pragma unroll
    for (int u = 0; u < NC; ++u) {
        const int e = e0 + u * kWave + lane;
        pre[u] = e1 > e0 ? a.idx[min(e, last)] : 0;
        wc[u] = e1 > e0 ? a.w[min(e, last)] : 0.0f;
    }
}

// The chain's steps from the window's bits of each edge (bit k: the presynaptic cell fired at step T0 - delay + k);
// returns g_k in lane k.
template <int NC>
__device__ __forceinline__ float chain_steps_bits(const Window& W, int e0, int e1, float acc, int lane,
                                                  const float* wc, const uint32_t* hist) {
    const int nch = (e1 - e0 + kWave - 1) / kWave;
    float gk = 0.0f;
    for (int k = 0; k < W.K; ++k) {
        acc = acc * W.a_s;
#pragma unroll
        for (int u = 0; u < NC; ++u) {
            if (u < nch) {
                uint64_t m = __builtin_amdgcn_ballot_w64((hist[u] >> k) & 1u);
                while (m) {
                    const int j = __builtin_ctzll(m);
                    acc = acc + fh_readlane(wc[u], j);
                    m &= m - 1;
                }
            }
        }
        gk = lane == k ? acc : gk;
    }
    return gk;
}

// chain_steps_bits with each step's fired weights folded through `list` (step_fold): the same additions in the same
// order.
template <int NC>
__device__ __forceinline__ float chain_steps_fold(const Window& W, int e0, int e1, float acc, int lane, const float* wc,
                                                  const uint32_t* hist, float* list) {
    const int nch = (e1 - e0 + kWave - 1) / kWave;
    float gk = 0.0f;
    for (int k = 0; k < W.K; ++k) {
        acc = step_fold<NC>(acc * W.a_s, k, nch, wc, hist, list, lane);
        gk = lane == k ? acc : gk;
    }
    return gk;
}

// The edges of a cell of at most NC * 64 in-edges in registers: weights and the window's bits of each presynaptic
// history word (0 past the row end).
template <int NC>
__device__ __forceinline__ void load_window_edges(const Edges& a, const Window& W, int e0, int e1, int lane, float* wc,
                                                  uint32_t* hist) {
    int pre[NC];
    load_edges<NC>(a, e0, e1, lane, pre, wc);
#pragma unroll
    for (int u = 0; u < NC; ++u) {
        const uint64_t raw = a.hist[pre[u]];
        hist[u] = e0 + u * kWave + lane < e1 ? W.bits(raw) : 0u;
    }
}

// The chain of a cell of at most NC * 64 in-edges with its edges and history words held in registers.
template <int NC>
__device__ __forceinline__ float g_chain_cached(const Edges& a, const Window& W, int e0, int e1, float acc, int lane) {
    int pre[NC];
    float wc[NC];
    load_edges<NC>(a, e0, e1, lane, pre, wc);
    // history words 8 chunks at a time, folded to the window's bits at once (few registers for long cells)
    uint32_t hist[NC];
#pragma unroll
    for (int u0 = 0; u0 < NC; u0 += 8) {
        uint64_t raw[8];
#pragma unroll
        for (int v = 0; v < 8 && u0 + v < NC; ++v) raw[v] = a.hist[pre[u0 + v]];
#pragma unroll
        for (int v = 0; v < 8 && u0 + v < NC; ++v)
            hist[u0 + v] = e0 + (u0 + v) * kWave + lane < e1 ? W.bits(raw[v]) : 0u;
    }
    return chain_steps_bits<NC>(W, e0, e1, acc, lane, wc, hist);
}

// The same chain for a cell of any length, streaming its in-edges (256 at a time) once per step.
__device__ __forceinline__ float g_chain_stream(const Edges& a, const Window& W, int e0, int e1, float acc, int lane) {
    float gk = 0.0f;
    for (int k = 0; k < W.K; ++k) {
        acc = acc * W.a_s;
        for (int base = e0; base < e1; base += 4 * kWave) {
            int pre[4];
            bool act[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int e = base + u * kWave + lane;
                pre[u] = e < e1 ? a.idx[e] : -1;
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) act[u] = pre[u] >= 0 && ((W.bits(a.hist[max(pre[u], 0)]) >> k) & 1u);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (base + u * kWave >= e1) break;
                const float c = act[u] ? a.w[base + u * kWave + lane] : 0.0f;
                uint64_t m = __builtin_amdgcn_ballot_w64(act[u]);
                while (m) {
                    const int j = __builtin_ctzll(m);
                    acc = acc + fh_readlane(c, j);
                    m &= m - 1;
                }
            }
        }
        gk = lane == k ? acc : gk;
    }
    return gk;
}

// The chains of a very long cell (NCW * 64 < in-edges <= 4 * NCW * 64) by a workgroup of four wavefronts:
// wavefront wv holds chunks [wv * NCW, (wv + 1) * NCW) of the in-edges in registers, and for each step the running
// sum passes from wavefront to wavefront in edge order through LDS (`carry`, one barrier per segment).  Longer cells
// are streamed by wavefront 0.  Every wavefront of the workgroup calls this together; g_k ends in lane k of
// wavefront 0 (returned there; other wavefronts return 0).
template <int NCW>
__device__ __forceinline__ float g_chain_workgroup(const Edges& a, const Window& W, int e0, int e1, float g0, int lane,
                                                   int wv, float* carry) {
    const int nch = (e1 - e0 + kWave - 1) / kWave;
    if (nch > 4 * NCW) {
        float gk = 0.0f;
        if (wv == 0) gk = g_chain_stream(a, W, e0, e1, g0, lane);
        __syncthreads();
        return gk;
    }
    const int mych = min(max(nch - wv * NCW, 0), NCW);
    int pre[NCW];
    float wc[NCW];
    uint32_t hist[NCW];
#pragma unroll
    for (int u = 0; u < NCW; ++u) {
        const int e = e0 + (wv * NCW + u) * kWave + lane;
        const bool val = u < mych && e < e1;
        pre[u] = val ? a.idx[e] : 0;
        wc[u] = val ? a.w[e] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < NCW; ++u) {
        const bool val = u < mych && e0 + (wv * NCW + u) * kWave + lane < e1;
        hist[u] = val ? W.bits(a.hist[pre[u]]) : 0u;
    }
    float acc = g0, gk = 0.0f;
    for (int k = 0; k < W.K; ++k) {
        for (int seg = 0; seg < 4; ++seg) {
            if (wv == seg) {
                float x = seg == 0 ? acc * W.a_s : *carry;
#pragma unroll
                for (int u = 0; u < NCW; ++u) {
                    if (u < mych) {
                        uint64_t m = __builtin_amdgcn_ballot_w64((hist[u] >> k) & 1u);
                        while (m) {
                            const int j = __builtin_ctzll(m);
                            x = x + fh_readlane(wc[u], j);
                            m &= m - 1;
                        }
                    }
                }
                if (lane == 0) *carry = x;
            }
            __syncthreads();
        }
        acc = *carry;
        gk = lane == k ? acc : gk;
        __syncthreads();
    }
    return wv == 0 ? gk : 0.0f;
}

// The chains of a very long cell by a workgroup of four wavefronts, the per-step fold kept short: wavefront wv holds
// chunks [wv * NCW, (wv + 1) * NCW) of the in-edges (at most 4 * NCW * 64; the caller streams longer cells) with
// their window bits in registers.  For each step every wavefront compacts the weights of its edges that fired into
// its own LDS list, in edge order (ballot + lane prefix count, one masked store per chunk); after a barrier
// wavefront 0 folds the four lists in wavefront order onto g_{k-1} * a_s (fold_lanes), which is edge order, the same
// additions as the one-wavefront chain.  counts holds 4 ints; g_k ends in lane k of wavefront 0
// (returned there; 0 elsewhere).  lists: 4 * (NCW * 64 + 64) floats.
template <int NCW>
__device__ __forceinline__ float g_chain_compact(const Edges& a, const Window& W, int e0, int e1, float g0, int lane,
                                                 int wv, float* lists, int* counts) {
    const int nch = (e1 - e0 + kWave - 1) / kWave;
    const int mych = min(max(nch - wv * NCW, 0), NCW);
    float wc[NCW];
    uint32_t hist[NCW];
    {
        int pre[NCW];
#pragma unroll
        for (int u = 0; u < NCW; ++u) {
            const int e = e0 + (wv * NCW + u) * kWave + lane;
            const bool val = u < mych && e < e1;
            pre[u] = val ? a.idx[e] : -1;
            wc[u] = val ? a.w[e] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < NCW; ++u) hist[u] = pre[u] >= 0 ? W.bits(a.hist[pre[u]]) : 0u;
    }
    constexpr int kList = NCW * kWave + kWave;      // a wavefront's list, with room for the fold's last read
    float* const mine = lists + wv * kList;
    float acc = g0, gk = 0.0f;
    for (int k = 0; k < W.K; ++k) {
        int cnt = 0;
#pragma unroll
        for (int u = 0; u < NCW; ++u) {
            if (u < mych) {
                const bool fired = (hist[u] >> k) & 1u;
                const uint64_t m = __builtin_amdgcn_ballot_w64(fired);
                if (fired) {
                    const int below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                 __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                    mine[cnt + below] = wc[u];
                }
                cnt += __builtin_popcountll(m);
            }
        }
        if (lane == 0) counts[wv] = cnt;
        __syncthreads();
        if (wv == 0) {
            acc = acc * W.a_s;
            for (int w = 0; w < 4; ++w) {
                const int c = counts[w];
                const float* list = lists + w * kList;
                for (int b0 = 0; b0 < c; b0 += 63) acc = fold_lanes(acc, list[b0 + lane], min(63, c - b0));
            }
            gk = lane == k ? acc : gk;
        }
        __syncthreads();
    }
    return wv == 0 ? gk : 0.0f;
}

}  // namespace fh_chain
