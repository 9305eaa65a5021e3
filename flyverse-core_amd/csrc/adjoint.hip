// Optic-lobe rate network: forward substeps and the hand-written adjoint (an addition: the reference has no
// adjoint, only torch autograd through OpticLobe._step_frame_grad, flyverse/optic.py:606-676).
//
// Model per step t (flyverse/optic.py:525-551 with no spiking feedback), units i, movies b:
//   dr_t  = clip(v_t + b_i, 0, 1) - b_i
//   y_t   = W dr_t,            W_ij = G[type_i, type_j] * W0_ij
//   inp_t = g_rr*y_t + P_t - g_ad*adapt_t
//   v_t+1 = inp_t + (v_t - inp_t)*a_i,   a_i = exp(-dt/tau_type_i)
//   adapt_t+1 = dr_t + (adapt_t - dr_t)*a_ad
// Loss: mean over (t = 1..T, movie, readout unit) of (clip(v_t + b) - target_t)^2.
//
// Layout: (units, 32) "fly-minor" with 32 movies; a wavefront works on two rows, one per half-wave (lane >> 5), in a
// longest-row-first order so the two rows of a wavefront have similar lengths.  Every product uses the chunked order
// of csrc/spmv.hip (16-entry chunks folded in order, chunk sums folded in order), every reduction a fixed order, so
// forward and backward are deterministic and a recomputed segment reproduces the first forward bit for bit.
#include "common.h"

#include <type_traits>

namespace {

constexpr int NL = 32;   // movies per row: one half-wavefront

// sum over the 32 lanes of a half-wave in a fixed butterfly order (the same value on every lane of the half)
__device__ __forceinline__ float half_sum(float x) {
    x = x + __shfl_xor(x, 16, 32);
    x = x + __shfl_xor(x, 8, 32);
    x = x + __shfl_xor(x, 4, 32);
    x = x + __shfl_xor(x, 2, 32);
    x = x + __shfl_xor(x, 1, 32);
    return x;
}

// DPP lane moves (gfx9 encodings): lane i of a 16-lane row reads lane i + 4 (row_shl:4), i - 4 (row_shr:4) or
// (i + 8) mod 16 = i ^ 8 (row_ror:8); within a quad, lane i ^ 1 or i ^ 2 (quad_perm).  BANKS selects the groups of
// four lanes of each row that are written; the others keep `old`.
constexpr int kDppShl4 = 0x104, kDppShr4 = 0x114, kDppRor8 = 0x128, kDppXor1 = 0xB1, kDppXor2 = 0x4E;

template <int CTRL, int BANKS = 0xf>
__device__ __forceinline__ float dpp_move(float old, float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(x), CTRL, 0xf, BANKS,
                                                      false));
}

// The 16 sums over the 32 lanes of a half-wave of v[0..15] at once, by a transposing butterfly: at each of four
// exchange steps (partners l ^ 16, l ^ 8, l ^ 4, l ^ 2) a lane keeps one half of its values and adds its partner's
// copy of them, so the values per lane halve; a last exchange (l ^ 1) completes the sums.  Lanes 2m and 2m + 1 of
// the half end with the sum of v[m] (the same fixed tree of additions for every m).  Every exchange is a VALU lane
// move: one v_permlane16_swap per pair at the first step (lane l < 16 of the half gets lane l + 16's copy of the
// value it keeps and gives its own copy of the other, so the swap itself does the keep/give choice), DPP after it.
__device__ __forceinline__ float half_sums16(const float (&v)[16], int l) {
    float a8[8], a4[4], a2[2];
    const bool u8 = l & 8, u4 = l & 4, u2 = l & 2;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const auto sw = __builtin_amdgcn_permlane16_swap(__float_as_uint(v[k]), __float_as_uint(v[k + 8]), false,
                                                         false);
        a8[k] = __uint_as_float(sw[0]) + __uint_as_float(sw[1]);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float keep = u8 ? a8[k + 4] : a8[k], give = u8 ? a8[k] : a8[k + 4];
        a4[k] = keep + dpp_move<kDppRor8>(0.0f, give);
    }
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const float keep = u4 ? a4[k + 2] : a4[k], give = u4 ? a4[k] : a4[k + 2];
        a2[k] = keep + dpp_move<kDppShl4, 0x5>(dpp_move<kDppShr4, 0xa>(0.0f, give), give);
    }
    const float keep = u2 ? a2[1] : a2[0], give = u2 ? a2[0] : a2[1];
    float a1 = keep + dpp_move<kDppXor2>(0.0f, give);
    a1 = a1 + dpp_move<kDppXor1>(0.0f, a1);
    return a1;
}

typedef const __attribute__((address_space(4))) int* ci32p;
typedef const __attribute__((address_space(4))) float* cf32p;

// Chunked fold for the two rows of a wavefront: half 0 folds entries [a0, a1), half 1 entries [b0, b1), each
// chunk of 16 in order from +0, chunk sums in order from +0.  Both rows' indices and weights are read with scalar
// loads (the CSR is constant during a launch) and each half selects its own, so an entry costs one vector gather
// of 32 contiguous floats per half.
__device__ __forceinline__ float fold32_pair(ci32p idx, cf32p w, const float* __restrict__ x, int a0, int a1, int b0,
                                             int b1, int l, bool h) {
    const int n0 = a1 - a0, n1 = b1 - b0, nmax = max(n0, n1);
    const int nh = h ? n1 : n0;
    float total = 0.0f;
    for (int cs = 0; cs < nmax; cs += 16) {
        const int m0 = n0 - cs, m1 = n1 - cs;
        int c0[16], c1[16];
        float w0[16], w1[16];
        if (m0 >= 16) {
#pragma unroll
            for (int j = 0; j < 16; ++j) { c0[j] = idx[a0 + cs + j]; w0[j] = w[a0 + cs + j]; }
        } else {
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                c0[j] = j < m0 ? idx[a0 + cs + j] : 0;
                w0[j] = j < m0 ? w[a0 + cs + j] : 0.0f;
            }
        }
        if (m1 >= 16) {
#pragma unroll
            for (int j = 0; j < 16; ++j) { c1[j] = idx[b0 + cs + j]; w1[j] = w[b0 + cs + j]; }
        } else {
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                c1[j] = j < m1 ? idx[b0 + cs + j] : 0;
                w1[j] = j < m1 ? w[b0 + cs + j] : 0.0f;
            }
        }
        const int m = h ? m1 : m0;
        float xv[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int c = h ? c1[j] : c0[j];
            xv[j] = j < m ? x[(size_t)c * NL + l] : 0.0f;
        }
        float s = 0.0f;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            if (j < m) {
                const float wv = h ? w1[j] : w0[j];
                float p = wv * xv[j];
                s = s + p;
            }
        }
        if (cs < nh) total = total + s;
    }
    return total;
}

// Lane J of the caller's own half-wavefront: ds_swizzle in bit mode, and-mask 0, or-mask J.
template <int J>
__device__ __forceinline__ int half_lane(int v) {
    return __builtin_amdgcn_ds_swizzle(v, J << 5);
}

// f(std::integral_constant<int, J>) for J = J0 .. 15, so J can be a constant expression inside f.
template <int J, class F>
__device__ __forceinline__ void unroll16(F&& f) {
    if constexpr (J < 16) {
        f(std::integral_constant<int, J>{});
        unroll16<J + 1>(f);
    }
}

// fold32_pair over chunk-padded storage (row r's chunks [cptr[r], cptr[r+1]), chunk c = entries [16 c, 16 c + 16) of
// pidx / pw, pads idx -1 and skipped), with the same additions in the same order.  Per chunk one vector load brings
// both halves' indices and weights (in each half, lanes 0-15 the indices, lanes 16-31 the weights); each half reads
// its own through lane swizzles, the indices when it gathers its 16 rows of x, the weights when it folds them.  The
// stream is three chunks deep: the index/weight load of chunk k + 2 and the gathers of chunk k + 1 are in flight
// while chunk k is folded (vector loads complete in issue order, so each wait leaves the younger loads
// outstanding).  Past a row's last chunk a half repeats that chunk's loads and folds nothing.
//
// With EG, each chunk also adds to eg[16 c + m] (the chunk-padded order of this CSR) the sum over the half's 32 lanes
// of x[i_m, l] * own[l], own being the half's own row of another (rows, 32) array: the per-edge products of the
// transposed product's gathers with the row's own values (half_sums16), written once per real entry.  The old
// eg[16 c + m] is loaded with the chunk's gathers, so the fold's add waits for that stage only: a load issued in
// the fold would wait for every younger stage's loads too (vector loads complete in issue order).
struct HalfStage {
    float x[16];
    float eg;                            // EG: eg[16 c + m] of this chunk, m = (l >> 1) & 15
    uint32_t real;                       // bit j: entry j of this half's chunk is real (not a pad)
};

template <bool EG>
__device__ __forceinline__ float fold32_pair_padded(const int* __restrict__ pidx, const float* __restrict__ pw,
                                                    const float* __restrict__ x, int ca0, int ca1, int cb0, int cb1,
                                                    int l, bool h, float own, float* __restrict__ eg) {
    const int n0 = ca1 - ca0, n1 = cb1 - cb0, nmax = max(n0, n1);
    if (nmax == 0) return 0.0f;
    const int c0 = h ? cb0 : ca0, nh = h ? n1 : n0;
    auto load_v = [&](int k) -> uint32_t {
        const int c = max(c0 + min(k, nh - 1), 0);
        const uint32_t* src = l < 16 ? (const uint32_t*)pidx + (size_t)c * 16 + l
                                     : (const uint32_t*)pw + (size_t)c * 16 + (l - 16);
        return *src;
    };
    const int m = (l >> 1) & 15;                  // the entry whose gain partial this lane pair writes
    // the gathers of chunk k (clamped to the half's last chunk, as load_v), and with EG its old partials
    auto issue = [&](int k, uint32_t V, HalfStage& st) {
        st.real = 0;
        unroll16<0>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            const int ij = half_lane<j>((int)V);
            st.real |= (ij >= 0 ? 1u : 0u) << j;
            st.x[j] = x[(size_t)max(ij, 0) * NL + l];
        });
        if (EG) st.eg = eg[(int64_t)max(c0 + min(k, nh - 1), 0) * 16 + m];
    };
    // the barriers keep the issue order load_v, gathers, fold (the waits depend on it)
    auto pin = [] { __asm__ volatile("" ::: "memory"); __builtin_amdgcn_sched_barrier(0); };
    float total = 0.0f;
    // chunk k from its gathered rows and its index/weight word V (the weights are read out only here)
    auto fold = [&](int k, HalfStage& b, uint32_t V) {
        float s = 0.0f;
        unroll16<0>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            const float w = __int_as_float(half_lane<16 + j>((int)V));
            if ((b.real >> j) & 1u) { const float p = w * b.x[j]; s = s + p; }
        });
        if (k < nh) total = total + s;
        if (EG) {
            // the products with the row's own values, in place of the gathered values; a pad's product (row 0's
            // values, finite) only enters the sum of its own entry, which is not written
#pragma unroll
            for (int j = 0; j < 16; ++j) b.x[j] = b.x[j] * own;
            const float e = half_sums16(b.x, l);
            if (k < nh && (l & 1) == 0 && ((b.real >> m) & 1u)) eg[(int64_t)(c0 + k) * 16 + m] = b.eg + e;
        }
    };
    // iteration k: the word of chunk k + 2 is loaded, chunk k + 1 gathered, chunk k folded.  Words rotate through
    // three registers and stages through two, so no register that a load is still writing is read or reused.
    HalfStage S0, S1;
    uint32_t V0 = load_v(0), V1 = load_v(1), V2;
    pin();
    issue(0, V0, S0);
    pin();
    auto step = [&](int k, uint32_t& vload, uint32_t vissue, HalfStage& sissue, HalfStage& sfold, uint32_t vfold) {
        vload = load_v(k + 2);
        pin();
        issue(k + 1, vissue, sissue);
        pin();
        fold(k, sfold, vfold);
    };
    for (int k = 0;;) {
        step(k, V2, V1, S1, S0, V0);
        if (++k >= nmax) break;
        step(k, V0, V2, S0, S1, V1);
        if (++k >= nmax) break;
        step(k, V1, V0, S1, S0, V2);
        if (++k >= nmax) break;
        step(k, V2, V1, S0, S1, V0);
        if (++k >= nmax) break;
        step(k, V0, V2, S1, S0, V1);
        if (++k >= nmax) break;
        step(k, V1, V0, S0, S1, V2);
        if (++k >= nmax) break;
    }
    return total;
}

// Per-unit coefficients and per-edge weights from the parameters.
__global__ __launch_bounds__(256) void k_adj_params(const float* __restrict__ tau, const float* __restrict__ bias,
                                                    const int* __restrict__ unit_type, float* __restrict__ a,
                                                    float* __restrict__ b, int n, float dt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int k = unit_type[i];
    a[i] = expf(-dt / tau[k]);
    b[i] = bias[k];
}

__global__ __launch_bounds__(256) void k_adj_weights(const float* __restrict__ gain, const int* __restrict__ pair_bin,
                                                     const float* __restrict__ w0, float* __restrict__ w, int64_t nnz) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    w[e] = gain[pair_bin[e]] * w0[e];
}

__global__ __launch_bounds__(256) void k_adj_dr(const float* __restrict__ v, const float* __restrict__ b,
                                                float* __restrict__ dr, int64_t total) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const float base = b[t / NL];
    float s = v[t] + base;
    dr[t] = fh_clip01(s) - base;
}

struct FwdArgs {
    const int *ptr, *idx, *sched;       // chunk-padded: ptr in chunks, idx / w 16 entries per chunk
    const float *w, *a, *b;
    const float *v, *adapt, *dr, *P;     // step t
    float *v1, *adapt1, *dr1, *inp;      // step t+1 (inp of step t, may be null)
    const int* readout;                  // (n) readout slot or -1
    const float* target;                 // (R, 32) targets of step t+1, or null
    float* loss;                         // (R) squared-error sums of step t+1 over the movies, or null
    int rows;
    float g_rr, g_ad, a_ad;
};

__global__ __launch_bounds__(256) void k_adj_forward(FwdArgs f) {
    const int l = threadIdx.x & (NL - 1);
    const int pairs = (f.rows + 1) / 2;
    const bool h = (threadIdx.x & 63) >= NL;
    for (int pr = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); pr < pairs; pr += gridDim.x * 4) {
        const int r0 = f.sched[2 * pr];
        const int r1 = 2 * pr + 1 < f.rows ? f.sched[2 * pr + 1] : -1;
        const int r = h ? r1 : r0;
        const float y = fold32_pair_padded<false>(f.idx, f.w, f.dr, f.ptr[r0],
                                                  f.ptr[r0 + 1], r1 < 0 ? 0 : f.ptr[r1], r1 < 0 ? 0 : f.ptr[r1 + 1],
                                                  l, h, 0.0f, nullptr);
        if (r < 0) continue;
        const int64_t q = (int64_t)r * NL + l;
        float inp = f.g_rr * y;
        inp = inp + f.P[q];
        float gad = f.g_ad * f.adapt[q];
        inp = inp - gad;
        const float v = f.v[q];
        float d = v - inp;
        d = d * f.a[r];
        const float vn = inp + d;
        const float old = f.dr[q];
        float da = f.adapt[q] - old;
        da = da * f.a_ad;
        f.v1[q] = vn;
        f.adapt1[q] = old + da;
        const float base = f.b[r];
        float s = vn + base;
        const float rn = fh_clip01(s);
        f.dr1[q] = rn - base;
        if (f.inp) f.inp[q] = inp;
        if (f.loss) {
            const int k = f.readout[r];
            if (k >= 0) {       // the whole half-wave takes this branch together (same row)
                float e = rn - f.target[(int64_t)k * NL + l];
                e = e * e;
                e = half_sum(e);
                if (l == 0) f.loss[k] = e;
            }
        }
    }
}

// Backward, part 1 (elementwise): ly = g_rr * lv * (1 - a); ga[i] += sum_b lv * (v_t - inp_t).
__global__ __launch_bounds__(256) void k_adj_back_ly(const float* __restrict__ lv, const float* __restrict__ v,
                                                     const float* __restrict__ inp, const float* __restrict__ a,
                                                     float* __restrict__ ly, float* __restrict__ ga, int rows,
                                                     float g_rr) {
    const int l = threadIdx.x & (NL - 1);
    const int r = blockIdx.x * (blockDim.x / NL) + (int)(threadIdx.x / NL);
    if (r >= rows) return;
    const int64_t q = (int64_t)r * NL + l;
    const float lam = lv[q];
    float oma = 1.0f - a[r];
    float li = lam * oma;
    ly[q] = g_rr * li;
    float dv = v[q] - inp[q];
    float c = lam * dv;
    c = half_sum(c);
    if (l == 0) ga[r] = ga[r] + c;
}

struct BackArgs {
    const int *ptrT, *idxT, *schedT;     // chunk-padded W^T: ptrT in chunks, idxT / wT 16 entries per chunk
    const float *wT, *a, *b;
    const float *ly, *v, *dr;            // step t
    float *lv, *la;                      // in: step t+1, out: step t (own rows only)
    const int* readout;
    const float* target;                 // (R, 32) targets of step t, or null (t = 0)
    float* gb;                           // (n) per-unit bias gradient accumulator
    float* eg;                           // per-edge gain partials in chunk-padded W^T order, or null
    float* ly_next;                      // ly of step t - 1 from this step's lv (own rows), or null
    const float* inp;                    // step t's inputs, for the tau gradient term, or null (no term)
    float* ga;                           // (n) per-unit tau gradient accumulator
    int rows;
    float g_ad, a_ad, loss_scale, g_rr;
};

// Backward, part 2 (transpose product + the rest of the step): for unit j
//   ldr = (W^T ly)_j + la_t+1 * (1 - a_ad)
//   linp = lv_t+1 * (1 - a_j);  ladapt_t = la_t+1 * a_ad - g_ad * linp
//   gate = [0 <= v_t + b_j <= 1] (torch clamp convention); lr = loss_scale * (r_t - target_t) on readout units
//   lv_t = lv_t+1 * a_j + (ldr + lr) * gate;  gb_j += sum_b ldr * (gate - 1) + lr * gate
// With EG the same gathers also give the per-edge gain partials: eg (chunk-padded W^T order) += sum over movies of
// ly[row, b] * dr_t[j, b] for each entry (row) of unit j's column.
template <bool EG>
__global__ __launch_bounds__(256) void k_adj_back_rows(BackArgs k) {
    const int l = threadIdx.x & (NL - 1);
    const int pairs = (k.rows + 1) / 2;
    const bool h = (threadIdx.x & 63) >= NL;
    for (int pr = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); pr < pairs; pr += gridDim.x * 4) {
        const int j0 = k.schedT[2 * pr];
        const int j1 = 2 * pr + 1 < k.rows ? k.schedT[2 * pr + 1] : -1;
        const int j = h ? j1 : j0;
        const float own = EG && j >= 0 ? k.dr[(int64_t)j * NL + l] : 0.0f;
        const float s = fold32_pair_padded<EG>(k.idxT, k.wT, k.ly, k.ptrT[j0],
                                               k.ptrT[j0 + 1], j1 < 0 ? 0 : k.ptrT[j1], j1 < 0 ? 0 : k.ptrT[j1 + 1],
                                               l, h, own, k.eg);
        if (j < 0) continue;
        const int64_t q = (int64_t)j * NL + l;
        const float la1 = k.la[q], lv1 = k.lv[q];
        float c = la1 * (1.0f - k.a_ad);
        const float ldr = s + c;
        float linp = lv1 * (1.0f - k.a[j]);
        float lad = la1 * k.a_ad;
        float gl = k.g_ad * linp;
        lad = lad - gl;
        const float u = k.v[q] + k.b[j];
        const float gate = (u >= 0.0f && u <= 1.0f) ? 1.0f : 0.0f;
        float lr = 0.0f;
        if (k.target) {
            const int ro = k.readout[j];
            if (ro >= 0) {
                float diff = fh_clip01(u) - k.target[(int64_t)ro * NL + l];
                lr = k.loss_scale * diff;
            }
        }
        float lvn = lv1 * k.a[j];
        float up = ldr + lr;
        up = up * gate;
        lvn = lvn + up;
        k.lv[q] = lvn;
        k.la[q] = lad;
        float gbb = ldr * (gate - 1.0f);
        float gr = lr * gate;
        gbb = gbb + gr;
        gbb = half_sum(gbb);
        if (l == 0) k.gb[j] = k.gb[j] + gbb;
        // k_adj_back_ly's work for the steps around this one, on the own row: step t's tau term from lv_t+1, and
        // ly of step t - 1 from lv_t (the same operations, so the same values)
        if (k.inp) {
            float dv = k.v[q] - k.inp[q];
            float cg = lv1 * dv;
            cg = half_sum(cg);
            if (l == 0) k.ga[j] = k.ga[j] + cg;
        }
        if (k.ly_next) {
            float oma = 1.0f - k.a[j];
            float li = lvn * oma;
            k.ly_next[q] = k.g_rr * li;
        }
    }
}

// out[k] = scale_k * sum over the segment [seg[k], seg[k+1]) of vals[perm[m]] * (mult ? mult[perm[m]] : 1):
// one wavefront per segment, lane-strided partial sums folded in lane order.
__global__ __launch_bounds__(256) void k_segment_sum(const float* __restrict__ vals, const float* __restrict__ mult,
                                                     const int* __restrict__ perm, const int* __restrict__ seg,
                                                     float* __restrict__ out, int nseg) {
    const int lane = threadIdx.x & 63;
    const int k = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (k >= nseg) return;
    float s = 0.0f;
    for (int m = seg[k] + lane; m < seg[k + 1]; m += 64) {
        const int e = perm[m];
        float v = vals[e];
        if (mult) v = v * mult[e];
        s = s + v;
    }
    float total = 0.0f;
    for (int j = 0; j < 64; ++j) total = total + fh_readlane(s, j);
    if (lane == 0) out[k] = total;
}

// loss = scale * sum of the (step, readout) partial sums: 1024 threads each fold a strided slice in order, then
// the 1024 slice sums are folded in thread order (a fixed order for a given n).
__global__ __launch_bounds__(1024) void k_loss_total(const float* __restrict__ parts, int64_t n, float scale,
                                                     float* __restrict__ out) {
    __shared__ float slice[1024];
    const int tid = threadIdx.x;
    float s = 0.0f;
    for (int64_t m = tid; m < n; m += 1024) s = s + parts[m];
    slice[tid] = s;
    __syncthreads();
    if (tid == 0) {
        float total = 0.0f;
        for (int j = 0; j < 1024; ++j) total = total + slice[j];
        out[0] = total * scale;
    }
}

inline unsigned nblk(int64_t total, int per = 256) { return (unsigned)((total + per - 1) / per); }

}  // namespace

// p = [tau (types), bias (types), unit_type (n), a (n), b (n)]
FH_EXPORT int launch_adj_params(void** p, int n, float dt, void* stream) {
    if (n < 0) return FH_ERR_ARGS;
    if (!n) return 0;
    hipLaunchKernelGGL(k_adj_params, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, (const float*)p[0],
                       (const float*)p[1], (const int*)p[2], (float*)p[3], (float*)p[4], n, dt);
    FH_LAUNCH_CHECK();
}

// p = [gain (types*types), pair_bin (nnz), w0 (nnz), w (nnz)]
FH_EXPORT int launch_adj_weights(void** p, int64_t nnz, void* stream) {
    if (nnz < 0) return FH_ERR_ARGS;
    if (!nnz) return 0;
    hipLaunchKernelGGL(k_adj_weights, dim3(nblk(nnz)), dim3(256), 0, (hipStream_t)stream, (const float*)p[0],
                       (const int*)p[1], (const float*)p[2], (float*)p[3], nnz);
    FH_LAUNCH_CHECK();
}

// p = [v (n, 32), b (n), dr (n, 32)]
FH_EXPORT int launch_adj_dr(void** p, int n, void* stream) {
    if (n < 0) return FH_ERR_ARGS;
    if (!n) return 0;
    hipLaunchKernelGGL(k_adj_dr, dim3(nblk((int64_t)n * NL)), dim3(256), 0, (hipStream_t)stream, (const float*)p[0],
                       (const float*)p[1], (float*)p[2], (int64_t)n * NL);
    FH_LAUNCH_CHECK();
}

// One forward step.  p = [chunk_ptr, pidx, sched, pw (chunk-padded W: 16 entries per chunk, pads idx -1), a, b, v,
// adapt, dr, P, v1, adapt1, dr1, inp|null, readout, target|null, loss|null]; f = [g_rr, g_ad, a_ad].
FH_EXPORT int launch_adj_forward(void** p, int rows, const float* f, int grid, void* stream) {
    if (rows < 0 || grid <= 0) return FH_ERR_ARGS;
    if (!rows) return 0;
    FwdArgs a;
    a.ptr = (const int*)p[0]; a.idx = (const int*)p[1]; a.sched = (const int*)p[2];
    a.w = (const float*)p[3]; a.a = (const float*)p[4]; a.b = (const float*)p[5];
    a.v = (const float*)p[6]; a.adapt = (const float*)p[7]; a.dr = (const float*)p[8]; a.P = (const float*)p[9];
    a.v1 = (float*)p[10]; a.adapt1 = (float*)p[11]; a.dr1 = (float*)p[12]; a.inp = (float*)p[13];
    a.readout = (const int*)p[14]; a.target = (const float*)p[15]; a.loss = (float*)p[16];
    a.rows = rows; a.g_rr = f[0]; a.g_ad = f[1]; a.a_ad = f[2];
    if (a.loss && !a.target) return FH_ERR_ARGS;
    hipLaunchKernelGGL(k_adj_forward, dim3(grid), dim3(256), 0, (hipStream_t)stream, a);
    FH_LAUNCH_CHECK();
}

// p = [lv, v, inp, a, ly, ga]
FH_EXPORT int launch_adj_back_ly(void** p, int rows, float g_rr, void* stream) {
    if (rows < 0) return FH_ERR_ARGS;
    if (!rows) return 0;
    hipLaunchKernelGGL(k_adj_back_ly, dim3(nblk((int64_t)rows * NL)), dim3(256), 0, (hipStream_t)stream,
                       (const float*)p[0], (const float*)p[1], (const float*)p[2], (const float*)p[3], (float*)p[4],
                       (float*)p[5], rows, g_rr);
    FH_LAUNCH_CHECK();
}

// p = [ptrT, idxT, schedT, wT (chunk-padded W^T as launch_adj_forward's W), a, b, ly, v, dr, lv, la, readout,
// target|null, gb, eg|null (16 per chunk of W^T), ly_next|null, inp|null, ga]; f = [g_ad, a_ad, loss_scale, g_rr].
// With inp the launch also adds step t's tau term to ga (launch_adj_back_ly's ga), and with ly_next it writes ly of
// step t - 1 from the new lv (launch_adj_back_ly's ly, a buffer other than ly).
FH_EXPORT int launch_adj_back_rows(void** p, int rows, const float* f, int grid, void* stream) {
    if (rows < 0 || grid <= 0) return FH_ERR_ARGS;
    if (!rows) return 0;
    BackArgs k;
    k.ptrT = (const int*)p[0]; k.idxT = (const int*)p[1]; k.schedT = (const int*)p[2];
    k.wT = (const float*)p[3]; k.a = (const float*)p[4]; k.b = (const float*)p[5];
    k.ly = (const float*)p[6]; k.v = (const float*)p[7]; k.dr = (const float*)p[8];
    k.lv = (float*)p[9]; k.la = (float*)p[10]; k.readout = (const int*)p[11]; k.target = (const float*)p[12];
    k.gb = (float*)p[13]; k.eg = (float*)p[14];
    k.ly_next = (float*)p[15]; k.inp = (const float*)p[16]; k.ga = (float*)p[17];
    k.rows = rows; k.g_ad = f[0]; k.a_ad = f[1]; k.loss_scale = f[2]; k.g_rr = f[3];
    if ((k.inp && !k.ga) || (k.ly_next && k.ly_next == k.ly)) return FH_ERR_ARGS;
    if (k.eg)
        hipLaunchKernelGGL(k_adj_back_rows<true>, dim3(grid), dim3(256), 0, (hipStream_t)stream, k);
    else
        hipLaunchKernelGGL(k_adj_back_rows<false>, dim3(grid), dim3(256), 0, (hipStream_t)stream, k);
    FH_LAUNCH_CHECK();
}

// p = [vals, mult|null, perm, seg (nseg+1), out]
FH_EXPORT int launch_segment_sum(void** p, int nseg, void* stream) {
    if (nseg < 0) return FH_ERR_ARGS;
    if (!nseg) return 0;
    hipLaunchKernelGGL(k_segment_sum, dim3(nblk(nseg, 4)), dim3(256), 0, (hipStream_t)stream, (const float*)p[0],
                       (const float*)p[1], (const int*)p[2], (const int*)p[3], (float*)p[4], nseg);
    FH_LAUNCH_CHECK();
}

// p = [parts, out]
FH_EXPORT int launch_loss_total(void** p, int64_t n, float scale, void* stream) {
    if (n < 0) return FH_ERR_ARGS;
    hipLaunchKernelGGL(k_loss_total, dim3(1), dim3(1024), 0, (hipStream_t)stream, (const float*)p[0], n, scale,
                       (float*)p[1]);
    FH_LAUNCH_CHECK();
}
