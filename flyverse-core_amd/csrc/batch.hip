// Batched multi-fly brain step: 64 flies per wavefront, one lane per fly ("fly-minor" layout).
//
// State is (rows, 64): row r, lane b holds fly b's value of cell r, so a wavefront reads and writes one row of all
// flies with one coalesced 256-byte access.  A spike is one bit: after a LIF step the wavefront's ballot of "fired"
// is the 64-bit mask of which flies' copy of this cell spiked, stored in a ring of masks.  Two layouts: slot-major
// (ring_slots, rows), where all wavefronts reading the same step share one small slot, and row-major
// (rows, ring_slots), where a cache line holds one cell's masks for consecutive steps; the autotuner picks.
//
// The synaptic delay (n_delay steps, 18 at dt 0.1 ms) makes every row independent for n_delay consecutive steps:
// the input of step t comes from spikes of step t - n_delay, which a previous launch wrote.  So one launch runs up
// to n_delay steps of every row with no synchronisation between wavefronts, and each row's membrane state stays in
// registers across those steps.  The ring has ring_slots >= 2 * n_delay entries so the slots read during a launch
// (steps T0 - D .. T0 + K - 1 - D) are never the slots written (T0 .. T0 + K - 1).
//
// Synaptic input keeps the reference CPU order (flyverse/brain.py:832-833, 679-692): g = g*a_s, then the in-edges
// of the row in ascending presynaptic order, each adding w to the flies whose presynaptic bit is set.  The bit mask
// is the select condition of the add (inverse ballot: an SGPR pair read as a per-lane predicate), so each edge costs
// one add and one select for all 64 flies, and lanes whose bit is clear are untouched, exactly like the reference
// skipping silent presynaptic cells.
#include "common.h"

namespace {

// Constant-address-space views: a launch never writes the CSR or the ring slots it reads, so these reads may go
// through the scalar cache (s_load with an SGPR offset: no per-edge address arithmetic).
typedef const __attribute__((address_space(4))) uint32_t* cu32p;
typedef const __attribute__((address_space(4))) float* cf32p;
typedef const __attribute__((address_space(4))) uint64_t* cu64p;

// acc + w[e] for every edge e in [e0, e1) whose presynaptic mask has this lane's bit, in edge order.  The mask is
// the select condition (an SGPR pair), so a lane whose bit is clear keeps acc unchanged, as the reference skips a
// silent presynaptic cell; 8 edges per round batch the scalar loads.
__device__ __forceinline__ float fold_edges(float acc, cu32p off, cf32p w, const char* slot, int e0, int e1) {
    int e = e0;
    for (; e + 8 <= e1; e += 8) {
        uint32_t o[8];
        float wv[8];
        uint64_t m[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) { o[j] = off[e + j]; wv[j] = w[e + j]; }
#pragma unroll
        for (int j = 0; j < 8; ++j) m[j] = *(cu64p)(slot + o[j]);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            float t = acc + wv[j];
            acc = __builtin_amdgcn_inverse_ballot_w64(m[j]) ? t : acc;
        }
    }
    for (; e < e1; ++e) {
        uint64_t m = *(cu64p)(slot + off[e]);
        float t = acc + w[e];
        acc = __builtin_amdgcn_inverse_ballot_w64(m) ? t : acc;
    }
    return acc;
}

// The same fold with the masks gathered by vector loads: lane j loads the mask of edge j of a 64-edge chunk (one
// instruction for 64 scattered 8-byte reads through the vector cache), then each edge's mask is read back into an
// SGPR pair with v_readlane and used as the select condition, in edge order.
__device__ __forceinline__ float fold_edges_gather(float acc, const uint32_t* __restrict__ off, cf32p w,
                                                   const char* slot, int e0, int e1, int lane) {
    for (int cb = e0; cb < e1; cb += 64) {
        const int e = cb + lane;
        uint64_t mj = 0;
        if (e < e1) mj = *(const uint64_t*)(slot + off[e]);
        const uint32_t lo = (uint32_t)mj, hi = (uint32_t)(mj >> 32);
        const int n = min(64, e1 - cb);
        int j = 0;
        for (; j + 8 <= n; j += 8) {
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const uint64_t m = (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)lo, j + u) |
                                   ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)hi, j + u) << 32);
                float t = acc + w[cb + j + u];
                acc = __builtin_amdgcn_inverse_ballot_w64(m) ? t : acc;
            }
        }
        for (; j < n; ++j) {
            const uint64_t m = (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)lo, j) |
                               ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)hi, j) << 32);
            float t = acc + w[cb + j];
            acc = __builtin_amdgcn_inverse_ballot_w64(m) ? t : acc;
        }
    }
    return acc;
}

// Transpose of the 32x32 bit block held by each half-wave (lane = row, bit = column), in place: five exchange stages
// with the lane at distance s (ds_swizzle, xor mode), each moving the off-diagonal s x s bit blocks with one rotate
// (v_alignbit) and one bit-field insert (v_bfi).
template <int S, uint32_t M>
__device__ __forceinline__ uint32_t transpose_stage(uint32_t x, int lane) {
    const uint32_t t = (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, (S << 10) | 0x1F);
    const bool lower = (lane & S) != 0;
    const uint32_t ts = __builtin_amdgcn_alignbit(t, t, lower ? S : 32 - S);     // rotr(t, S) or rotl(t, S)
    const uint32_t keep = lower ? ~M : M;
    return (x & keep) | (ts & ~keep);
}

__device__ __forceinline__ uint32_t transpose32(uint32_t x, int lane) {
    x = transpose_stage<16, 0x0000FFFFu>(x, lane);
    x = transpose_stage<8, 0x00FF00FFu>(x, lane);
    x = transpose_stage<4, 0x0F0F0F0Fu>(x, lane);
    x = transpose_stage<2, 0x33333333u>(x, lane);
    x = transpose_stage<1, 0x55555555u>(x, lane);
    return x;
}

// The fold with the 64 x 64 bit matrix of a 64-edge chunk transposed: lane j gathers the mask of edge j (bit b =
// fly b's copy of the presynaptic cell fired); after the transpose lane b holds the bits of its own fly over the 64
// edges, and walks its set bits in ascending edge order adding the staged weights.  The work per chunk is the
// transpose plus as many rounds as the busiest lane has set bits, instead of two lane reads per edge.
__device__ __forceinline__ float fold_chunk_transposed(float acc, uint64_t mj, float wj, int lane, float* wlds) {
    wlds[lane] = wj;
    uint32_t lo = (uint32_t)mj, hi = (uint32_t)(mj >> 32);
    // 64 x 64 = four 32 x 32 blocks: swap the off-diagonal halves between lane j and j ^ 32, then transpose each
    const uint32_t send = lane < 32 ? hi : lo;
    const uint32_t got = (uint32_t)__shfl_xor((int)send, 32);
    if (lane < 32) hi = got; else lo = got;
    lo = transpose32(lo, lane);
    hi = transpose32(hi, lane);
    __builtin_amdgcn_wave_barrier();
    while (__builtin_amdgcn_ballot_w64((lo | hi) != 0u)) {
        if ((lo | hi) != 0u) {
            const int j = lo ? __builtin_ctz(lo) : 32 + __builtin_ctz(hi);
            acc = acc + wlds[j];
            if (lo) lo &= lo - 1u; else hi &= hi - 1u;
        }
    }
    __builtin_amdgcn_wave_barrier();
    return acc;
}

// The fold with the 64 x 64 bit matrix of a 64-edge chunk transposed: lane j gathers the mask of edge j (bit b =
// fly b's copy of the presynaptic cell fired); after the transpose lane b holds the bits of its own fly over the 64
// edges, and walks its set bits in ascending edge order adding the staged weights.  The work per chunk is the
// transpose plus as many rounds as the busiest lane has set bits, instead of two lane reads per edge.  The gathers
// of four chunks are issued together: they do not depend on the running sum, so only the adds form the chain.
__device__ __forceinline__ float fold_edges_transposed(float acc, const uint32_t* __restrict__ off,
                                                       const float* __restrict__ w, const char* slot, int e0, int e1,
                                                       int lane, float* wlds) {
    constexpr int U = 4;
    for (int cb = e0; cb < e1; cb += U * 64) {
        uint64_t mj[U];
        float wj[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = cb + u * 64 + lane;
            mj[u] = 0;
            wj[u] = 0.0f;
            if (e < e1) {
                mj[u] = *(const uint64_t*)(slot + off[e]);
                wj[u] = w[e];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (cb + u * 64 >= e1) break;
            acc = fold_chunk_transposed(acc, mj[u], wj[u], lane, wlds);
        }
    }
    return acc;
}

// fold_chunk_transposed with the weights read by lane permute instead of staged in LDS: lane j holds the weight of
// edge j, and the lane of fly b fetches the weight of its next set bit with ds_bpermute.  The half swap before the
// two 32 x 32 transposes is one v_permlane32_swap (gfx950): lanes 0-31 get (edges 0-31, edges 32-63) of flies 0-31,
// lanes 32-63 those of flies 32-63.  A lane's adds go over edges 0-31, then edges 32-63, each ascending.
__device__ __forceinline__ float fold_chunk_permute(float acc, uint64_t mj, float wj, int lane) {
    const auto sw = __builtin_amdgcn_permlane32_swap((uint32_t)mj, (uint32_t)(mj >> 32), false, false);
    uint32_t lo = transpose32(sw[0], lane);
    uint32_t hi = transpose32(sw[1], lane);
    const int wi = __float_as_int(wj);
    while (__builtin_amdgcn_ballot_w64(lo != 0u)) {
        const int j = __builtin_ctz(lo | 0x80000000u);          // 31 (unused) once a lane is done
        const float wv = __int_as_float(__builtin_amdgcn_ds_bpermute(j << 2, wi));
        if (lo != 0u) {
            acc = acc + wv;
            lo &= lo - 1u;
        }
    }
    while (__builtin_amdgcn_ballot_w64(hi != 0u)) {
        const int j = 32 + __builtin_ctz(hi | 0x80000000u);
        const float wv = __int_as_float(__builtin_amdgcn_ds_bpermute(j << 2, wi));
        if (hi != 0u) {
            acc = acc + wv;
            hi &= hi - 1u;
        }
    }
    return acc;
}

// The two halves of fold_chunk_permute, so the transposes of several chunks can be issued before their adds: the
// transpose of one 64-edge chunk (lane b gets lo = edges 0-31, hi = edges 32-63 of fly b), then the adds of one
// chunk over a lane's set bits, ascending, two bits per round (both weights fetched before either is added).
__device__ __forceinline__ void transpose_chunk(uint64_t mj, int lane, uint32_t& lo, uint32_t& hi) {
    const auto sw = __builtin_amdgcn_permlane32_swap((uint32_t)mj, (uint32_t)(mj >> 32), false, false);
    lo = transpose32(sw[0], lane);
    hi = transpose32(sw[1], lane);
}

__device__ __forceinline__ float add_bits(float acc, uint32_t bits, int wi, int lane_base) {
    while (__builtin_amdgcn_ballot_w64(bits != 0u)) {
        const uint32_t b1 = bits & (bits - 1u);
        const int j0 = lane_base + __builtin_ctz(bits | 0x80000000u);
        const int j1 = lane_base + __builtin_ctz(b1 | 0x80000000u);
        const float w0 = __int_as_float(__builtin_amdgcn_ds_bpermute(j0 << 2, wi));
        const float w1 = __int_as_float(__builtin_amdgcn_ds_bpermute(j1 << 2, wi));
        if (bits != 0u) acc = acc + w0;
        if (b1 != 0u) acc = acc + w1;
        bits = b1 & (b1 - 1u);
    }
    return acc;
}

// Mode 5's form of the transpose: the same five exchange stages, each merging the kept and received bits with one
// bit-field select (v_bitop3 / v_bfi from the xor form) instead of an and plus an and-or.
template <int S, uint32_t M>
__device__ __forceinline__ uint32_t transpose_stage_x(uint32_t x, int lane) {
    const uint32_t t = (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, (S << 10) | 0x1F);
    const bool lower = (lane & S) != 0;
    const uint32_t ts = __builtin_amdgcn_alignbit(t, t, lower ? S : 32 - S);
    const uint32_t keep = lower ? ~M : M;
    return ts ^ ((x ^ ts) & keep);
}

__device__ __forceinline__ void transpose_chunk_x(uint64_t mj, int lane, uint32_t& lo, uint32_t& hi) {
    const auto sw = __builtin_amdgcn_permlane32_swap((uint32_t)mj, (uint32_t)(mj >> 32), false, false);
    uint32_t x = sw[0], y = sw[1];
    x = transpose_stage_x<16, 0x0000FFFFu>(x, lane); y = transpose_stage_x<16, 0x0000FFFFu>(y, lane);
    x = transpose_stage_x<8, 0x00FF00FFu>(x, lane);  y = transpose_stage_x<8, 0x00FF00FFu>(y, lane);
    x = transpose_stage_x<4, 0x0F0F0F0Fu>(x, lane);  y = transpose_stage_x<4, 0x0F0F0F0Fu>(y, lane);
    x = transpose_stage_x<2, 0x33333333u>(x, lane);  y = transpose_stage_x<2, 0x33333333u>(y, lane);
    x = transpose_stage_x<1, 0x55555555u>(x, lane);  y = transpose_stage_x<1, 0x55555555u>(y, lane);
    lo = x;
    hi = y;
}

// Index of the lowest set bit, -1 for a zero word (v_ffbl_b32 as the hardware defines it).
__device__ __forceinline__ int lowest_bit(uint32_t x) {
    int r;
    __asm__("v_ffbl_b32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// Mode 5's adds of one 32-edge half of a chunk: seg points at the half's 32 weights in LDS and seg[-1] holds -0.0,
// so a lane with no bit left reads -0.0 and its add leaves acc unchanged bit for bit (x + -0.0 = x for every x,
// signed zeros included).  The adds therefore need no select; a lane's adds still go over its set bits in
// ascending edge order, two per round.
__device__ __forceinline__ float add_bits_lds(float acc, uint32_t bits, const float* seg) {
    while (__builtin_amdgcn_ballot_w64(bits != 0u)) {
        const uint32_t b1 = bits & (bits - 1u);
        const float w0 = seg[lowest_bit(bits)];
        const float w1 = seg[lowest_bit(b1)];
        acc = acc + w0;
        acc = acc + w1;
        bits = b1 & (b1 - 1u);
    }
    return acc;
}

struct BatchLif {
    const int* ptr;          // (n_rows + 1) post-major CSR over LIF rows
    const uint32_t* pre_off; // (nnz) byte offset of the presynaptic row's masks in the ring (row * ring_slots * 8)
    const float* w;          // (nnz) weights (mV)
    const int* sched;        // (n_rows) visiting order
    const int* neuron_id;    // (n_rows) global neuron index (Philox counter)
    const uint8_t* row_flags;// (n_rows) bit 0: some fly has Poisson forcing on this row
    float *v, *g, *refrac, *rate, *counts;
    const float *drive, *pois;
    uint64_t* ring;
    const uint64_t* seed;    // (64) per-lane Philox key
    const uint32_t* fly_step;// (64) steps this lane's fly has taken at the start of the frame
    const uint32_t* budget;  // (64) steps this lane's fly may take in total
    const uint32_t* engine_step; // (1) engine step at the start of the frame (ring slot clock)
    const float* P;          // (12) v_rest, v_reset, v_th, a_m, dt, t_ref, -, -, -, a_r, rate_gain, a_s
    const uint8_t* act;      // mode 6: (n_rows) bit k = the mask step k of the launch reads is nonzero
    const uint64_t* zero;    // mode 6: one zero word, the address of every gather whose mask is known to be zero
    int n_rows, lanes, ring_slots, delay, record, mode, row_major;
};

__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

template <int MODE>
__global__ __launch_bounds__(256) void k_batch_lif(BatchLif a, int step_offset, int K) {
    constexpr bool L5 = MODE >= 5;                  // LDS-staged weights
    constexpr bool ACT = MODE == 6;                 // gathers skipped where the activity bits say the mask is zero
    __shared__ float wstage[MODE == 2 ? 4 : 1][64];
    __shared__ uint64_t cmask[MODE == 4 ? 4 : 1][256];
    __shared__ float cweight[MODE == 4 ? 4 : 1][256];
    // mode 5: per wavefront 16 segments of 33 floats, [-0.0, the weights of 32 consecutive edges]; segments 2c and
    // 2c + 1 hold the two halves of chunk c of the row (rows of up to 512 edges) or of the current work item
    __shared__ float wseg[L5 ? 4 : 1][L5 ? 16 * 33 : 1];
    const int lane = threadIdx.x & 63;
    float* const ws = wseg[L5 ? threadIdx.x >> 6 : 0];
    const int ws_lane = (lane >> 5) * 33 + 1 + (lane & 31);      // this lane's weight slot within a chunk's segments
    if (L5 && lane < 16) ws[lane * 33] = -0.0f;
    const int gw = blockIdx.x * 4 + wave_id();
    const int nw = gridDim.x * 4;
    const uint32_t T0 = *a.engine_step + (uint32_t)step_offset;
    const float v_rest = a.P[0], v_reset = a.P[1], v_th = a.P[2], am = a.P[3], dt = a.P[4], t_ref = a.P[5];
    const float ar = a.P[9], rg = a.P[10], a_s = a.P[11];
    const uint32_t R = (uint32_t)a.ring_slots, D = (uint32_t)a.delay;
    uint64_t seed = 0;
    uint32_t fstep = 0;
    int live_steps = 0;
    if (lane < a.lanes) {
        seed = a.seed[lane];
        fstep = a.fly_step[lane] + (uint32_t)step_offset;
        uint32_t bud = a.budget[lane];
        live_steps = bud > fstep ? (int)min(bud - fstep, (uint32_t)K) : 0;
    }
    for (int si = gw; si < a.n_rows; si += nw) {
        const int r = a.sched[si];
        const int e0 = a.ptr[r], e1 = a.ptr[r + 1];
        const uint32_t nid = (uint32_t)a.neuron_id[r];
        const bool has_pois = a.row_flags[r] & 1;
        const int64_t q = (int64_t)r * 64 + lane;
        float v = a.v[q], g = a.g[q], ref = a.refrac[q], rate = a.rate[q];
        float cnt = a.record ? a.counts[q] : 0.0f;
        const float drive = a.drive[q];
        const float p = has_pois ? a.pois[q] : 0.0f;
        // LIF update of flyverse/brain.py:838-868 with active = 1, no adaptation or depression, for step k
        auto lif_step = [&](int k, float acc, uint32_t wslot) {
            float target = v_rest + acc;
            target = target + drive;
            float dv = v - target;
            dv = dv * am;
            float vn = target + dv;
            float rn = ref;
            if (rn > 0.0f) vn = v_reset;
            rn = rn - dt;
            if (rn < 0.0f) rn = 0.0f;
            float spike = (vn >= v_th) ? 1.0f : 0.0f;
            if (p > 0.0f && fh_uniform(seed, fstep + (uint32_t)k, nid, 0u, 0u) < p) spike = 1.0f;
            const bool fired = spike > 0.0f;
            if (fired) { vn = v_reset; rn = t_ref; }
            float rr = rate * ar;
            float rs = spike * rg;
            const bool live = k < live_steps;
            if (live) {
                g = acc; v = vn; ref = rn; rate = rr + rs;
                cnt = cnt + spike;
            }
            const uint64_t bits = __builtin_amdgcn_ballot_w64(live && fired);
            if (lane == 0) a.ring[a.row_major ? (size_t)r * R + wslot : (size_t)wslot * a.n_rows + r] = bits;
        };
        auto slot_of = [&](int k) {
            const uint32_t rslot = (T0 + (uint32_t)k + R - D) % R;
            return (const char*)(a.row_major ? a.ring + rslot : a.ring + (size_t)rslot * a.n_rows);
        };
        if (MODE >= 3) {
            // Work items (step k, group of 4 chunks of 64 edges), groups inner.  The mask gathers of item i + 1 are
            // issued before item i is folded (two mask buffers that swap roles, so no load is waited for early).
            // Rows of at most 8 chunks keep their offsets and weights in registers for the whole launch; longer
            // rows load those of item i + 1 just before its gathers.
            const int ne = e1 - e0;
            if (ne == 0) {
                for (int k = 0; k < K; ++k) lif_step(k, g * a_s, (T0 + (uint32_t)k) % R);
            } else {
                const int G = (ne + 255) >> 8;
                // a cursor over the work items (step k, group grp), groups inner, with the ring slots of step k
                // (read: step t - D, write: step t) advanced by increments instead of divisions
                struct Cursor { int k, grp; uint32_t rslot, wslot; };
                const uint32_t r0 = (T0 + R - D) % R, w0 = T0 % R;
                const size_t slot_bytes = a.row_major ? 8 : (size_t)a.n_rows * 8;
                auto advance = [&](Cursor& c) {
                    if (c.k == K - 1 && c.grp == G - 1) return;             // stays on the last item
                    if (++c.grp == G) {
                        c.grp = 0;
                        ++c.k;
                        c.rslot = c.rslot + 1 == R ? 0 : c.rslot + 1;
                        c.wslot = c.wslot + 1 == R ? 0 : c.wslot + 1;
                    }
                };
                // raw mask words; lanes past the row's end are cleared when folded, so nothing waits on the loads here.
                // Mode 6: a lane whose activity bit for step k is clear (its presynaptic mask is zero) reads the one
                // shared zero word instead, so such lanes cost the gather one cache line together; the loads stay
                // unconditional, so the in-order wait counts are unchanged.
                uint32_t act_lo = ~0u, act_hi = ~0u;       // activity bytes of chunks 0-3 and 4-7 (rows of G <= 2)
                auto gather = [&](const Cursor& c, const uint32_t* o, uint64_t* m) {
                    const char* rb = (const char*)a.ring + (size_t)c.rslot * slot_bytes;
                    const uint32_t aw = c.grp ? act_hi : act_lo;
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const uint64_t* src = (const uint64_t*)(rb + o[u]);
                        if (ACT && !((aw >> (8 * u + c.k)) & 1u)) src = a.zero;
                        m[u] = *src;
                    }
                };
                float acc = 0.0f;
                auto fold = [&](const Cursor& c, const uint64_t* m, const float* wv) {
                    if (c.grp == 0) acc = g * a_s;
                    const int nch = min(4, (ne - c.grp * 256 + 63) >> 6);
                    const int base = c.grp * 256 + lane;
                    if (MODE == 4) {
                        // the item's edges with a nonzero mask word (some fly's presynaptic cell fired), compacted
                        // in edge order through LDS; an all-zero word adds nothing in any fly
                        uint64_t* cm = cmask[threadIdx.x >> 6];
                        float* cw = cweight[threadIdx.x >> 6];
                        int n = 0;
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const uint64_t mv = u < nch && base + u * 64 < ne ? m[u] : 0ull;
                            const uint64_t bal = __builtin_amdgcn_ballot_w64(mv != 0ull);
                            const int pos = n + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                    __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
                            if (mv != 0ull) { cm[pos] = mv; cw[pos] = wv[u]; }
                            n += __builtin_popcountll(bal);
                        }
                        __builtin_amdgcn_wave_barrier();
                        for (int cb = 0; cb < n; cb += 64) {
                            const uint64_t mj = cb + lane < n ? cm[cb + lane] : 0ull;
                            acc = fold_chunk_permute(acc, mj, cw[cb + lane], lane);
                        }
                        __builtin_amdgcn_wave_barrier();
                    } else if (L5) {
                        uint32_t lo[4], hi[4];
#pragma unroll
                        for (int u = 0; u < 4; ++u)
                            if (u < nch) transpose_chunk_x(base + u * 64 < ne ? m[u] : 0ull, lane, lo[u], hi[u]);
                        const float* sg = ws + 1;
                        if (G > 2) {
                            // weights of this item's chunks, in the item's segments (the previous item's reads of
                            // them were issued earlier by this wavefront, and LDS serves a wavefront in order)
#pragma unroll
                            for (int u = 0; u < 4; ++u) ws[u * 66 + ws_lane] = wv[u];
                            __builtin_amdgcn_wave_barrier();
                        } else {
                            sg += c.grp * 8 * 33;
                        }
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            if (u < nch) {
                                acc = add_bits_lds(acc, lo[u], sg + (2 * u) * 33);
                                acc = add_bits_lds(acc, hi[u], sg + (2 * u + 1) * 33);
                            }
                        }
                    } else {
                        uint32_t lo[4], hi[4];
#pragma unroll
                        for (int u = 0; u < 4; ++u)
                            if (u < nch) transpose_chunk(base + u * 64 < ne ? m[u] : 0ull, lane, lo[u], hi[u]);
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            if (u < nch) {
                                acc = add_bits(acc, lo[u], __float_as_int(wv[u]), 0);
                                acc = add_bits(acc, hi[u], __float_as_int(wv[u]), 32);
                            }
                        }
                    }
                    if (c.grp == G - 1) lif_step(c.k, acc, c.wslot);
                };
                // the prefetches are unconditional (past the last item they repeat it) so the wait before a fold
                // counts only the gathers issued after its own
                auto run = [&](auto item_ow) {
                    uint32_t oA[4], oB[4];
                    float wA[4], wB[4];
                    uint64_t mA[4], mB[4];
                    const int items = K * G;
                    Cursor f{0, 0, r0, w0}, p = f;                        // fold and prefetch cursors
                    item_ow(p, oA, wA);
                    gather(p, oA, mA);
                    for (int it = 0; it < items; it += 2) {
                        advance(p);
                        item_ow(p, oB, wB);
                        gather(p, oB, mB);
                        fold(f, mA, wA);
                        advance(f);
                        advance(p);
                        item_ow(p, oA, wA);
                        gather(p, oA, mA);
                        if (it + 1 < items) {
                            fold(f, mB, wB);
                            advance(f);
                        }
                    }
                };
                if (G <= 2) {
                    // offsets and weights of the row's (at most 8) chunks stay in registers for the whole launch
                    uint32_t oc[8];
                    float wc[8];
#pragma unroll
                    for (int c = 0; c < 8; ++c) {
                        const int e = min(e0 + c * 64 + lane, e1 - 1);
                        oc[c] = a.pre_off[e];
                        wc[c] = a.w[e];
                    }
                    if (ACT) {
                        // slot-major ring: the byte offset of a mask within a slot is 8 x its row
                        act_lo = 0u;
                        act_hi = 0u;
#pragma unroll
                        for (int c = 0; c < 4; ++c) {
                            act_lo |= (uint32_t)a.act[oc[c] >> 3] << (8 * c);
                            act_hi |= (uint32_t)a.act[oc[4 + c] >> 3] << (8 * c);
                        }
                    }
                    if (L5) {
                        // the row's weights go to LDS once per launch (after this wavefront's reads of the previous
                        // row's, in program order)
                        __builtin_amdgcn_wave_barrier();
#pragma unroll
                        for (int c = 0; c < 8; ++c) ws[c * 66 + ws_lane] = wc[c];
                        __builtin_amdgcn_wave_barrier();
                    }
                    run([&](const Cursor& c, uint32_t* o, float* wv) {
                        const bool hi = c.grp != 0;
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            o[u] = hi ? oc[4 + u] : oc[u];
                            wv[u] = hi ? wc[4 + u] : wc[u];
                        }
                    });
                } else {
                    run([&](const Cursor& c, uint32_t* o, float* wv) {
                        const int base = e0 + c.grp * 256 + lane;
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const int e = min(base + u * 64, e1 - 1);
                            o[u] = a.pre_off[e];
                            wv[u] = a.w[e];
                        }
                    });
                }
            }
        } else {
            for (int k = 0; k < K; ++k) {
                const char* rb = slot_of(k);
                float acc = g * a_s;
                if (MODE == 2)
                    acc = fold_edges_transposed(acc, a.pre_off, a.w, rb, e0, e1, lane, wstage[threadIdx.x >> 6]);
                else if (MODE == 1)
                    acc = fold_edges_gather(acc, a.pre_off, (cf32p)a.w, rb, e0, e1, lane);
                else
                    acc = fold_edges(acc, (cu32p)a.pre_off, (cf32p)a.w, rb, e0, e1);
                lif_step(k, acc, (T0 + (uint32_t)k) % R);
            }
        }
        a.v[q] = v; a.g[q] = g; a.refrac[q] = ref; a.rate[q] = rate;
        if (a.record) a.counts[q] = cnt;
    }
}

// ------------------------------------------------------------------ optic lobe, fly-minor
// Chunked products over a (cols, 64) input with chunk-padded storage: row r owns chunks [chunk_ptr[r],
// chunk_ptr[r+1]) (at least one: an empty row has one all-pad chunk), chunk c is entries [16 c, 16 c + 16) of
// pidx / pw, pads have idx -1 and are skipped.  Wavefront `wave` of `nw` walks the chunks of its rows (sched order
// from `wave`, stride nw) as one stream, three chunks deep: the indices and weights of chunk c + 2 arrive by one vector
// load (lanes 0-15 the indices, lanes 16-31 the weights), the 16 row gathers of chunk c + 1 (256 bytes each) are in
// flight, and chunk c is folded.  Vector loads complete in issue order, so each wait leaves the younger stages' loads
// outstanding.  The row after the leading cursor's is fetched (scalar loads) when the cursor enters a row, so moving
// to the next row waits on nothing.  Each chunk is folded from +0 in entry order and the chunk sums of a row in chunk
// order (csrc/spmv.hip order); finish(row, total) consumes the row total.
struct FmCursor { int si, r, c, ce; };

template <class Finish>
__device__ __forceinline__ void fm_chunk_stream(const int* __restrict__ chunk_ptr, const int* __restrict__ pidx,
                                                const float* __restrict__ pw, const int* __restrict__ sched,
                                                const float* __restrict__ x, int rows, int lane, int wave, int nw,
                                                Finish finish) {
    auto row_at = [&](int si) {
        FmCursor q;
        q.si = si;
        q.r = sched[min(si, rows - 1)];
        q.c = chunk_ptr[q.r];
        q.ce = chunk_ptr[q.r + 1];
        return q;
    };
    FmCursor q0 = row_at(wave);
    if (q0.si >= rows) return;
    FmCursor nr = row_at(q0.si + nw);                  // the row after the leading cursor's
    // the leading cursor's next chunk (past the last row it stays past the end: si >= rows)
    auto next = [&](FmCursor q) {
        if (q.c + 1 < q.ce) { ++q.c; return q; }
        if (q.si >= rows) return q;
        q = nr;
        nr = row_at(min(q.si + nw, rows));
        return q;
    };
    const int vl = lane & 31;
    // one chunk's indices (lanes 0-15) and weights (lanes 16-31; lanes 32-63 repeat lanes 0-31)
    auto load_v = [&](const FmCursor& q) -> uint32_t {
        const uint32_t* src = vl < 16 ? (const uint32_t*)pidx + q.c * 16 + vl
                                      : (const uint32_t*)pw + q.c * 16 + (vl - 16);
        return *src;
    };
    // the chunk's 16 row gathers (pads read row 0), its weights read out into scalars and the mask of its real
    // entries, so the index/weight register is free once the gathers are issued
    struct Stage {
        float x[16];
        float w[16];
        uint32_t valid;
    };
    auto load_x = [&](uint32_t V, Stage& st) {
        st.valid = (uint32_t)__builtin_amdgcn_ballot_w64((int)V >= 0) & 0xFFFFu;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int ij = __builtin_amdgcn_readlane((int)V, j);
            st.w[j] = __int_as_float(__builtin_amdgcn_readlane((int)V, 16 + j));
            st.x[j] = x[(size_t)max(ij, 0) * 64 + lane];
        }
    };
    // the barriers keep the issue order load_v, load_x, fold (the waits depend on it): pin() stops the compiler
    // from sinking a load below it, the scheduling barrier stops the instruction scheduler from moving code across
    auto pin = [] { __asm__ volatile("" ::: "memory"); __builtin_amdgcn_sched_barrier(0); };
    float total = 0.0f;
    auto fold = [&](const FmCursor& q, const Stage& st) {
        float s = 0.0f;
#pragma unroll
        for (int j = 0; j < 16; ++j)
            if ((st.valid >> j) & 1u) { const float p = st.w[j] * st.x[j]; s = s + p; }
        total = total + s;
        if (q.c + 1 == q.ce) { finish(q.r, total); total = 0.0f; }
    };
    // stages past the end repeat chunk q0's loads and are never folded; two of each buffer alternate, so no register
    // that a load is still writing is ever copied
    // (the prologue issues the loads in the order the loop does, V of q1 before the gathers of q0, so the loop is
    // entered with the same loads outstanding as around its back edge)
    FmCursor q1 = next(q0);
    Stage SA, SB;
    const uint32_t V0 = load_v(q0);
    uint32_t VA, VB = load_v(q1.si < rows ? q1 : q0);
    pin();
    load_x(V0, SA);
    pin();
    FmCursor q2 = next(q1);
    for (;;) {
        VA = load_v(q2.si < rows ? q2 : q0);
        pin();
        load_x(VB, SB);
        pin();
        fold(q0, SA);
        if (q1.si >= rows) break;
        q0 = q1; q1 = q2; q2 = next(q2);
        VB = load_v(q2.si < rows ? q2 : q0);
        pin();
        load_x(VA, SA);
        pin();
        fold(q0, SB);
        if (q1.si >= rows) break;
        q0 = q1; q1 = q2; q2 = next(q2);
    }
}

// y[r, b] = scale * row total (chunk-padded CSR, csrc/spmv.hip order) over the (cols, 64) input.
__global__ __launch_bounds__(256) void k_spmm_fm(const int* __restrict__ chunk_ptr, const int* __restrict__ pidx,
                                                 const float* __restrict__ pw, const int* __restrict__ sched,
                                                 const float* __restrict__ x, float* __restrict__ y, int rows,
                                                 float scale) {
    const int lane = threadIdx.x & 63;
    fm_chunk_stream(chunk_ptr, pidx, pw, sched, x, rows, lane, blockIdx.x * 4 + wave_id(), gridDim.x * 4,
                    [&](int r, float total) {
        if (scale != 1.0f) total = total * scale;
        y[(int64_t)r * 64 + lane] = total;
    });
}

// The optic update of flyverse/optic.py:525-551 (neural.cu:68-79) for row r, lane q = r * 64 + lane, given the
// recurrent product total = (W_rr dr_cur)[r]: writes v, adapt and the next dr; lanes of finished flies only copy dr.
__device__ __forceinline__ void optic_update_one(int r, int64_t q, float total, bool live, const float* dr_cur,
                                                 float* dr_next, float* v, float* adapt, const float* pr,
                                                 const float* spk, const float* a, const float* b, float gr, float ga,
                                                 float aad) {
    const float old = dr_cur[q];
    if (!live) { dr_next[q] = old; return; }
    float inp = gr * total;
    inp = inp + pr[q];
    float gad = ga * adapt[q];
    inp = inp - gad;
    inp = inp + spk[q];
    float d = v[q] - inp;
    d = d * a[r];
    float vn = inp + d;
    v[q] = vn;
    float da = adapt[q] - old;
    da = da * aad;
    adapt[q] = old + da;
    float s2 = vn + b[r];
    dr_next[q] = fh_clip01(s2) - b[r];
}

// One optic substep fused with its recurrent product: y = W_rr dr_cur (chunked order), then the update, writing the
// next dr into dr_next.
__global__ __launch_bounds__(256) void k_optic_substep_fm(const int* __restrict__ chunk_ptr, const int* __restrict__ pidx,
                                                          const float* __restrict__ pw, const int* __restrict__ sched,
                                                          const float* __restrict__ dr_cur, float* __restrict__ dr_next,
                                                          float* __restrict__ v, float* __restrict__ adapt,
                                                          const float* __restrict__ pr, const float* __restrict__ spk,
                                                          const float* __restrict__ a, const float* __restrict__ b,
                                                          const uint64_t* __restrict__ live_mask, int rows, float gr,
                                                          float ga, float aad) {
    const int lane = threadIdx.x & 63;
    const bool live = (*live_mask >> lane) & 1ull;
    fm_chunk_stream(chunk_ptr, pidx, pw, sched, dr_cur, rows, lane, blockIdx.x * 4 + wave_id(), gridDim.x * 4,
                    [&](int r, float total) {
        optic_update_one(r, (int64_t)r * 64 + lane, total, live, dr_cur, dr_next, v, adapt, pr, spk, a, b, gr, ga,
                         aad);
    });
}

__global__ __launch_bounds__(256) void k_optic_dr_fm(const float* __restrict__ v, const float* __restrict__ b,
                                                     float* __restrict__ dr, int64_t total) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    float base = b[t >> 6];
    float s = v[t] + base;
    dr[t] = fh_clip01(s) - base;
}

// Photoreceptor stage over (column, family) cells with the radiance drawn in-kernel, one thread per (cell k, lane b):
// rad[col, c] of lane b = Philox(movie_seed[b], frame[b], col*4 + c, 0, stream 1), the uniform the plain path's
// launch_uniform gives fly 0 of a one-fly batch; the averaging, state update and contrast are those of
// csrc/optic.hip's k_photoreceptor.  I_lp, I_mean (5 n_col, 64); a_pr (n_pr, 64).
struct PrFm {
    const float* sens;
    const int *pr_col, *cell_ptr, *cell_pr;
    const float* cell_w;
    float *I_lp, *I_mean, *a_pr;
    const uint64_t* movie_seed;
    const uint32_t* frame;
    const uint64_t *fresh_mask, *live_mask;
    const float* cf;
};

__device__ __forceinline__ void pr_cell_fm(const PrFm& a, int k, int lane) {
    const int64_t t = (int64_t)k * 64 + lane;
    const int e0 = a.cell_ptr[k], e1 = a.cell_ptr[k + 1];
    // a cell without photoreceptors feeds nothing: its state stays at the zeros it starts from (its input is 0)
    if (e0 == e1 || !((*a.live_mask >> lane) & 1ull)) return;
    const uint64_t ms = a.movie_seed[lane];
    const uint32_t fr = a.frame[lane];
    float I = 0.0f;
    for (int cs = e0; cs < e1; cs += 16) {
        const int ce = min(cs + 16, e1);
        float s = 0.0f;
        for (int e = cs; e < ce; ++e) {
            const int p = a.cell_pr[e];
            const uint32_t base = (uint32_t)a.pr_col[p] * 4u;
            float Ip = 0.0f;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                float rad = fh_uniform(ms, fr, base + (uint32_t)c, 0u, 1u);
                float q = rad * a.sens[p * 4 + c];
                Ip = Ip + q;
            }
            float q = a.cell_w[e] * Ip;
            s = s + q;
        }
        I = I + s;
    }
    const bool fr0 = (*a.fresh_mask >> lane) & 1ull;
    float lp = fr0 ? I : a.I_lp[t];
    float mean = fr0 ? I : a.I_mean[t];
    float l1 = a.cf[0] * lp, l2 = a.cf[1] * I;
    lp = l1 + l2;
    float m1 = a.cf[2] * mean, m2 = a.cf[3] * lp;
    mean = m1 + m2;
    a.I_lp[t] = lp; a.I_mean[t] = mean;
    float num = lp - mean, den = mean + a.cf[4];
    float c = num / den;
    c = c < -1.0f ? -1.0f : c;
    c = c > a.cf[5] ? a.cf[5] : c;
    for (int e = e0; e < e1; ++e) a.a_pr[(int64_t)a.cell_pr[e] * 64 + lane] = c;
}

__global__ __launch_bounds__(256) void k_photoreceptor_fm(PrFm a, int n_cells) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n_cells * 64) return;
    pr_cell_fm(a, (int)(t >> 6), (int)(t & 63));
}

__device__ __forceinline__ float fb_value_fm(float rate) {
    float q = rate / 100.0f;
    q = q < 0.0f ? 0.0f : q;
    return q > 3.0f ? 3.0f : q;
}

__device__ __forceinline__ float drive_value_fm(float y, float gain_out, float clip) {
    float d = gain_out * y;
    d = d < -clip ? -clip : d;
    return d > clip ? clip : d;
}

__global__ __launch_bounds__(256) void k_feedback_fm(const float* __restrict__ rate, const int* __restrict__ spk_row,
                                                     float* __restrict__ s, int64_t total) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    s[t] = fb_value_fm(rate[(int64_t)spk_row[t >> 6] * 64 + (t & 63)]);
}

__global__ __launch_bounds__(256) void k_drive_out_fm(const float* __restrict__ y, const int* __restrict__ spk_row,
                                                      float* __restrict__ drive, float gain_out, float clip,
                                                      int64_t total) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    drive[(int64_t)spk_row[t >> 6] * 64 + (t & 63)] = drive_value_fm(y[t], gain_out, clip);
}

// The frame's optic inputs (flyverse/optic.py:496-517, 572, 579-580) in two launches.  k_frame_cells_fm, workgroups of
// 4 rows x 64 lanes: the photoreceptor cells, the feedback s of every spiking target and the first substep's
// dr = clip01(v + b) - b.  k_frame_held_fm, only in a frame with substeps: pr_in = gain_in * W_rp a_pr and
// spk_in = gain_fb * W_rs s, each streamed by `j_prod` workgroups.  Values and order are those of the separate
// kernels.
struct FrameInFm {
    PrFm pr;
    const float* rate;     // (rows, 64) LIF rates
    const int* spk_row;    // (n_spk)
    float* s;              // (n_spk, 64)
    const float *v, *b;    // (n_rate, 64), (n_rate)
    float* dr;             // (n_rate, 64)
    const int *rp_ptr, *rp_idx, *rp_sched;
    const float* rp_w;
    const int *rs_ptr, *rs_idx, *rs_sched;
    const float* rs_w;
    float *pr_in, *spk_in; // (n_rate, 64)
    float gain_in, gain_fb;
    int n_cells, n_spk, n_rate, j_cells, j_spk, j_prod;
};

__global__ __launch_bounds__(256) void k_frame_cells_fm(FrameInFm a) {
    const int lane = threadIdx.x & 63, row = (int)(threadIdx.x >> 6);
    int j = blockIdx.x;
    if (j < a.j_cells) {
        const int k = j * 4 + row;
        if (k < a.n_cells) pr_cell_fm(a.pr, k, lane);
    } else if ((j -= a.j_cells) < a.j_spk) {
        const int k = j * 4 + row;
        if (k < a.n_spk) a.s[(int64_t)k * 64 + lane] = fb_value_fm(a.rate[(int64_t)a.spk_row[k] * 64 + lane]);
    } else {
        const int r = (j - a.j_spk) * 4 + row;
        if (r < a.n_rate) {
            const int64_t q = (int64_t)r * 64 + lane;
            const float base = a.b[r];
            const float x = a.v[q] + base;
            a.dr[q] = fh_clip01(x) - base;
        }
    }
}

__global__ __launch_bounds__(256) void k_frame_held_fm(FrameInFm a) {
    const int lane = threadIdx.x & 63;
    const int u = blockIdx.x, nw = a.j_prod * 4;
    if (u < a.j_prod)
        fm_chunk_stream(a.rp_ptr, a.rp_idx, a.rp_w, a.rp_sched, a.pr.a_pr, a.n_rate, lane, u * 4 + wave_id(), nw,
                        [&](int r, float tot) { a.pr_in[(int64_t)r * 64 + lane] = tot * a.gain_in; });
    else
        fm_chunk_stream(a.rs_ptr, a.rs_idx, a.rs_w, a.rs_sched, a.s, a.n_rate, lane, (u - a.j_prod) * 4 + wave_id(),
                        nw, [&](int r, float tot) { a.spk_in[(int64_t)r * 64 + lane] = tot * a.gain_fb; });
}

// The frame's optic output into the LIF drive, fused (flyverse/optic.py:599-604): drive[spk_row[k], b] =
// clamp(gain_out * (W_sr dr)[k, b]) for dr (n_rate, 64) = clip01(v + b) - b (the last substep's output).
__global__ __launch_bounds__(256) void k_frame_out_fm(const int* __restrict__ chunk_ptr, const int* __restrict__ pidx,
                                                      const float* __restrict__ pw, const int* __restrict__ sched,
                                                      const float* __restrict__ dr, const int* __restrict__ spk_row,
                                                      float* __restrict__ drive, int rows, float gain_out, float clip) {
    const int lane = threadIdx.x & 63;
    fm_chunk_stream(chunk_ptr, pidx, pw, sched, dr, rows, lane, blockIdx.x * 4 + wave_id(), gridDim.x * 4,
                    [&](int r, float tot) {
                        drive[(int64_t)spk_row[r] * 64 + lane] = drive_value_fm(tot, gain_out, clip);
                    });
}

// End of a frame: the engine clock and every live fly advance by `steps`; live and fresh masks are refreshed.
__global__ void k_frame_end(uint32_t* engine_step, uint32_t* fly_step, const uint32_t* budget, uint32_t* frame,
                            uint64_t* fresh_mask, uint64_t* live_mask, int lanes, int steps) {
    const int lane = threadIdx.x;
    bool live_next = false;
    if (lane < lanes) {
        uint32_t fs = fly_step[lane], bud = budget[lane];
        if (fs < bud) {
            fs = fs + min((uint32_t)steps, bud - fs);
            fly_step[lane] = fs;
            frame[lane] = frame[lane] + 1;
        }
        live_next = fs < bud;
    }
    const uint64_t live = __builtin_amdgcn_ballot_w64(live_next);
    if (lane == 0) {
        *engine_step = *engine_step + (uint32_t)steps;
        *live_mask = live;
        *fresh_mask = 0;
    }
}

// Start new flies on the lanes of `mask`: LIF and optic state to rest, their spike bits cleared from every ring
// slot, photoreceptor stage marked fresh.  Per-lane seeds, budgets and step counters are written by the caller.
__global__ __launch_bounds__(256) void k_reset_lanes(float* v, float* g, float* refrac, float* rate, float* counts,
                                                     float* drive, float* pois, float* ov, float* oadapt, uint64_t* ring,
                                                     uint64_t* fresh_mask, uint64_t* live_mask, uint64_t mask,
                                                     uint64_t live_set, float v_rest, int n_rows, int n_rate,
                                                     int ring_slots) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = (int)(t & 63);
    const bool mine = (mask >> lane) & 1ull;
    const int64_t lif = (int64_t)n_rows * 64, opt = (int64_t)n_rate * 64;
    if (t < lif && mine) {
        v[t] = v_rest; g[t] = 0.0f; refrac[t] = 0.0f; rate[t] = 0.0f; counts[t] = 0.0f; drive[t] = 0.0f;
        pois[t] = 0.0f;
    }
    if (t < opt && mine) { ov[t] = 0.0f; oadapt[t] = 0.0f; }
    const int64_t words = (int64_t)n_rows * ring_slots;
    for (int64_t k = t; k < words; k += (int64_t)gridDim.x * blockDim.x) ring[k] &= ~mask;
    if (t == 0) {
        *fresh_mask |= mask;
        *live_mask = (*live_mask & ~mask) | (live_set & mask);
    }
}

inline unsigned nblk(int64_t total) { return (unsigned)((total + 255) / 256); }

// The photoreceptor pointers p[0..12] of launch_photoreceptor_fm.
inline PrFm pr_fm(void** p) {
    return PrFm{(const float*)p[0], (const int*)p[1], (const int*)p[2], (const int*)p[3], (const float*)p[4],
                (float*)p[5], (float*)p[6], (float*)p[7], (const uint64_t*)p[8], (const uint32_t*)p[9],
                (const uint64_t*)p[10], (const uint64_t*)p[11], (const float*)p[12]};
}

// The activity bits of one k_batch_lif<6> launch (slot-major ring): act[r] bit k is set when the mask of row r in
// the slot that step k of the launch reads, (T0 + k - delay) mod ring_slots, has any fly's bit set.  Those slots
// were written by earlier launches (K <= delay), so the bits describe exactly what the launch will read.
__global__ __launch_bounds__(256) void k_batch_act(const uint64_t* __restrict__ ring, const uint32_t* engine_step,
                                                   int step_offset, int K, int n_rows, int R, int D,
                                                   uint8_t* __restrict__ act) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= n_rows) return;
    const uint32_t T0 = *engine_step + (uint32_t)step_offset;
    uint32_t slot = (T0 + (uint32_t)R - (uint32_t)D) % (uint32_t)R;
    uint32_t bits = 0u;
    for (int k = 0; k < K; ++k) {
        if (ring[(size_t)slot * n_rows + r] != 0ull) bits |= 1u << k;
        slot = slot + 1 == (uint32_t)R ? 0u : slot + 1;
    }
    act[r] = (uint8_t)bits;
}

}  // namespace

// Up to `delay` LIF steps of every row for 64 flies.  p = [ptr, pre_off, w, sched, neuron_id, row_flags, v, g,
// refrac, rate, counts, drive, pois, ring, seed, fly_step, budget, engine_step, P(12)];
// dims = [n_rows, lanes, ring_slots, delay, record, grid, mode, row_major]; mode 0 reads the spike masks through the
// scalar cache, mode 1 gathers them with vector loads, mode 2 gathers and transposes them, mode 3 transposes with
// the gathers of the next work item in flight, mode 4 also drops the edges no fly transmits before transposing,
// mode 5 is mode 3 with the weights staged in LDS behind a -0.0 sentinel (adds without selects), mode 6 is mode 5
// with the mask gathers of presynaptic cells that fired in no fly skipped: it needs a slot-major ring, K <= 8 and
// p[19] = an activity buffer of (n_rows rounded up to a multiple of 8) + 8 bytes whose last 8 bytes are zero (the
// bits are written by a k_batch_act launch just before the LIF launch)
// (fold_edges, fold_edges_gather, fold_edges_transposed, fold_chunk_permute: the same additions in the same order,
// so the same result; a skipped gather is of a zero mask, which adds nothing).
FH_EXPORT int launch_batch_lif(void** p, const int* dims, int step_offset, int K, void* stream) {
    BatchLif a;
    a.ptr = (const int*)p[0]; a.pre_off = (const uint32_t*)p[1]; a.w = (const float*)p[2]; a.sched = (const int*)p[3];
    a.neuron_id = (const int*)p[4]; a.row_flags = (const uint8_t*)p[5];
    a.v = (float*)p[6]; a.g = (float*)p[7]; a.refrac = (float*)p[8]; a.rate = (float*)p[9]; a.counts = (float*)p[10];
    a.drive = (const float*)p[11]; a.pois = (const float*)p[12]; a.ring = (uint64_t*)p[13];
    a.seed = (const uint64_t*)p[14]; a.fly_step = (const uint32_t*)p[15]; a.budget = (const uint32_t*)p[16];
    a.engine_step = (const uint32_t*)p[17]; a.P = (const float*)p[18];
    a.n_rows = dims[0]; a.lanes = dims[1]; a.ring_slots = dims[2]; a.delay = dims[3]; a.record = dims[4];
    const int grid = dims[5];
    a.mode = dims[6];
    a.row_major = dims[7];
    if (a.n_rows < 0 || a.lanes < 0 || a.lanes > 64 || K < 0 || K > a.delay || a.ring_slots < 2 * a.delay ||
        grid <= 0 || a.mode < 0 || a.mode > 6)
        return FH_ERR_ARGS;
    a.act = nullptr;
    a.zero = nullptr;
    if (a.mode == 6) {
        if (a.row_major || K > 8 || p[19] == nullptr) return FH_ERR_ARGS;
        a.act = (const uint8_t*)p[19];
        a.zero = (const uint64_t*)(a.act + (((size_t)a.n_rows + 7) & ~(size_t)7));
    }
    if (a.n_rows == 0 || K == 0) return 0;
    if (a.mode == 6) {
        hipLaunchKernelGGL(k_batch_act, dim3(nblk(a.n_rows)), dim3(256), 0, (hipStream_t)stream, a.ring,
                           a.engine_step, step_offset, K, a.n_rows, a.ring_slots, a.delay, (uint8_t*)p[19]);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return (int)e;
    }
    switch (a.mode) {
        case 0: hipLaunchKernelGGL(k_batch_lif<0>, dim3(grid), dim3(256), 0, (hipStream_t)stream, a, step_offset, K); break;
        case 1: hipLaunchKernelGGL(k_batch_lif<1>, dim3(grid), dim3(256), 0, (hipStream_t)stream, a, step_offset, K); break;
        case 2: hipLaunchKernelGGL(k_batch_lif<2>, dim3(grid), dim3(256), 0, (hipStream_t)stream, a, step_offset, K); break;
        case 3: hipLaunchKernelGGL(k_batch_lif<3>, dim3(grid), dim3(256), 0, (hipStream_t)stream, a, step_offset, K); break;
        case 4: hipLaunchKernelGGL(k_batch_lif<4>, dim3(grid), dim3(256), 0, (hipStream_t)stream, a, step_offset, K); break;
        case 5: hipLaunchKernelGGL(k_batch_lif<5>, dim3(grid), dim3(256), 0, (hipStream_t)stream, a, step_offset, K); break;
        default: hipLaunchKernelGGL(k_batch_lif<6>, dim3(grid), dim3(256), 0, (hipStream_t)stream, a, step_offset, K);
    }
    FH_LAUNCH_CHECK();
}

// y (rows, 64) = scale * W x (cols, 64).  p = [chunk_ptr (rows+1), pidx (16 per chunk, pad -1), pw, sched, x, y].
FH_EXPORT int launch_spmm_fm(void** p, int rows, float scale, int grid, void* stream) {
    if (rows < 0 || grid <= 0) return FH_ERR_ARGS;
    if (!rows) return 0;
    hipLaunchKernelGGL(k_spmm_fm, dim3(grid), dim3(256), 0, (hipStream_t)stream, (const int*)p[0], (const int*)p[1],
                       (const float*)p[2], (const int*)p[3], (const float*)p[4], (float*)p[5], rows, scale);
    FH_LAUNCH_CHECK();
}

// p = [chunk_ptr, pidx, pw, sched, dr_cur, dr_next, v, adapt, pr, spk, a, b, live_mask] (chunk-padded CSR as
// launch_spmm_fm).
FH_EXPORT int launch_optic_substep_fm(void** p, int rows, float gr, float ga, float aad, int grid, void* stream) {
    if (rows < 0 || grid <= 0) return FH_ERR_ARGS;
    if (!rows) return 0;
    hipLaunchKernelGGL(k_optic_substep_fm, dim3(grid), dim3(256), 0, (hipStream_t)stream, (const int*)p[0],
                       (const int*)p[1], (const float*)p[2], (const int*)p[3], (const float*)p[4], (float*)p[5],
                       (float*)p[6], (float*)p[7], (const float*)p[8], (const float*)p[9], (const float*)p[10],
                       (const float*)p[11], (const uint64_t*)p[12], rows, gr, ga, aad);
    FH_LAUNCH_CHECK();
}

// p = [v, b, dr], (rows, 64).
FH_EXPORT int launch_optic_dr_fm(void** p, int rows, void* stream) {
    if (rows < 0) return FH_ERR_ARGS;
    if (!rows) return 0;
    int64_t total = (int64_t)rows * 64;
    hipLaunchKernelGGL(k_optic_dr_fm, dim3(nblk(total)), dim3(256), 0, (hipStream_t)stream, (const float*)p[0],
                       (const float*)p[1], (float*)p[2], total);
    FH_LAUNCH_CHECK();
}

// p = [sens (n_pr, 4), pr_col, cell_ptr (n_cells + 1), cell_pr, cell_w, I_lp, I_mean, a_pr, movie_seed, frame,
// fresh_mask, live_mask, coefs]; n_cells = 5 * columns.
FH_EXPORT int launch_photoreceptor_fm(void** p, int n_cells, void* stream) {
    if (n_cells < 0) return FH_ERR_ARGS;
    if (!n_cells) return 0;
    hipLaunchKernelGGL(k_photoreceptor_fm, dim3(nblk((int64_t)n_cells * 64)), dim3(256), 0, (hipStream_t)stream,
                       pr_fm(p), n_cells);
    FH_LAUNCH_CHECK();
}

// p = [rate (rows, 64), spk_row (n_spk), s (n_spk, 64)].
FH_EXPORT int launch_feedback_fm(void** p, int n_spk, void* stream) {
    if (n_spk < 0) return FH_ERR_ARGS;
    if (!n_spk) return 0;
    int64_t total = (int64_t)n_spk * 64;
    hipLaunchKernelGGL(k_feedback_fm, dim3(nblk(total)), dim3(256), 0, (hipStream_t)stream, (const float*)p[0],
                       (const int*)p[1], (float*)p[2], total);
    FH_LAUNCH_CHECK();
}

// p = [y_sr (n_spk, 64), spk_row, drive (rows, 64)].
FH_EXPORT int launch_drive_out_fm(void** p, int n_spk, float gain_out, float clip, void* stream) {
    if (n_spk < 0) return FH_ERR_ARGS;
    if (!n_spk) return 0;
    int64_t total = (int64_t)n_spk * 64;
    hipLaunchKernelGGL(k_drive_out_fm, dim3(nblk(total)), dim3(256), 0, (hipStream_t)stream, (const float*)p[0],
                       (const int*)p[1], (float*)p[2], gain_out, clip, total);
    FH_LAUNCH_CHECK();
}

// The frame's optic inputs, two launches: p = [the 13 pointers of launch_photoreceptor_fm, rate (rows, 64),
// spk_row (n_spk), s (n_spk, 64), v (n_rate, 64), b (n_rate), dr (n_rate, 64), W_rp chunk_ptr, pidx, pw, sched,
// W_rs chunk_ptr, pidx, pw, sched, pr_in (n_rate, 64), spk_in (n_rate, 64)]; dims = [n_cells, n_spk, n_rate, grid]:
// grid workgroups stream each product, 0 in a frame without substeps.
FH_EXPORT int launch_frame_in_fm(void** p, const int* dims, float gain_in, float gain_fb, void* stream) {
    for (int i = 0; i < 4; ++i)
        if (dims[i] < 0) return FH_ERR_ARGS;
    FrameInFm a;
    a.pr = pr_fm(p);
    a.rate = (const float*)p[13]; a.spk_row = (const int*)p[14]; a.s = (float*)p[15]; a.v = (const float*)p[16];
    a.b = (const float*)p[17]; a.dr = (float*)p[18];
    a.rp_ptr = (const int*)p[19]; a.rp_idx = (const int*)p[20]; a.rp_w = (const float*)p[21];
    a.rp_sched = (const int*)p[22];
    a.rs_ptr = (const int*)p[23]; a.rs_idx = (const int*)p[24]; a.rs_w = (const float*)p[25];
    a.rs_sched = (const int*)p[26];
    a.pr_in = (float*)p[27]; a.spk_in = (float*)p[28];
    a.gain_in = gain_in; a.gain_fb = gain_fb;
    a.n_cells = dims[0]; a.n_spk = dims[1]; a.n_rate = dims[2];
    a.j_cells = (a.n_cells + 3) / 4; a.j_spk = (a.n_spk + 3) / 4;
    a.j_prod = a.n_rate ? dims[3] : 0;
    const int64_t g1 = (int64_t)a.j_cells + a.j_spk + (a.n_rate + 3) / 4;
    if (g1 >= (1ll << 31) || dims[3] >= (1 << 30)) return FH_ERR_SIZE;
    if (g1) {
        hipLaunchKernelGGL(k_frame_cells_fm, dim3((unsigned)g1), dim3(256), 0, (hipStream_t)stream, a);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return (int)e;
    }
    if (a.j_prod) hipLaunchKernelGGL(k_frame_held_fm, dim3(2 * a.j_prod), dim3(256), 0, (hipStream_t)stream, a);
    FH_LAUNCH_CHECK();
}

// The frame's optic output into the LIF drive in one launch: p = [W_sr chunk_ptr, pidx, pw, sched, dr (n_rate, 64),
// spk_row (n_spk), drive (rows, 64)]; dr must hold clip01(v + b) - b (the last substep's output).
FH_EXPORT int launch_frame_out_fm(void** p, int n_spk, float gain_out, float clip, int grid, void* stream) {
    if (n_spk < 0 || grid <= 0) return FH_ERR_ARGS;
    if (!n_spk) return 0;
    hipLaunchKernelGGL(k_frame_out_fm, dim3(grid), dim3(256), 0, (hipStream_t)stream, (const int*)p[0],
                       (const int*)p[1], (const float*)p[2], (const int*)p[3], (const float*)p[4], (const int*)p[5],
                       (float*)p[6], n_spk, gain_out, clip);
    FH_LAUNCH_CHECK();
}

// p = [engine_step, fly_step, budget, frame, fresh_mask, live_mask].
FH_EXPORT int launch_frame_end(void** p, int lanes, int steps, void* stream) {
    if (lanes < 0 || lanes > 64 || steps < 0) return FH_ERR_ARGS;
    hipLaunchKernelGGL(k_frame_end, dim3(1), dim3(64), 0, (hipStream_t)stream, (uint32_t*)p[0], (uint32_t*)p[1],
                       (const uint32_t*)p[2], (uint32_t*)p[3], (uint64_t*)p[4], (uint64_t*)p[5], lanes, steps);
    FH_LAUNCH_CHECK();
}

// p = [v, g, refrac, rate, counts, drive, pois, optic v, optic adapt, ring, fresh_mask, live_mask].
FH_EXPORT int launch_reset_lanes(void** p, uint64_t mask, uint64_t live_set, float v_rest, int n_rows, int n_rate,
                                 int ring_slots, void* stream) {
    if (n_rows < 0 || n_rate < 0 || ring_slots < 0) return FH_ERR_ARGS;
    int64_t total = (int64_t)(n_rows > n_rate ? n_rows : n_rate) * 64;
    if (total < 64) total = 64;
    hipLaunchKernelGGL(k_reset_lanes, dim3(nblk(total)), dim3(256), 0, (hipStream_t)stream, (float*)p[0],
                       (float*)p[1], (float*)p[2], (float*)p[3], (float*)p[4], (float*)p[5], (float*)p[6],
                       (float*)p[7], (float*)p[8], (uint64_t*)p[9], (uint64_t*)p[10], (uint64_t*)p[11], mask, live_set,
                       v_rest, n_rows, n_rate, ring_slots);
    FH_LAUNCH_CHECK();
}
