// Deterministic synaptic input for the plain (B, N) layout: no atomic adds of values, the reference CPU order.
//
// The reference's CPU path (flyverse/brain.py:679-692) adds w*x into every target in ascending presynaptic order
// after the decay g *= a_s (brain.py:832), skipping silent presynaptic cells; its CUDA event_scatter
// (flyverse/kernels/neural.cu:115-123) uses float atomicAdd and does not repeat run to run.  Here every target row
// is owned by one wavefront, which reads its in-edges (post-major CSR, columns ascending), ballots the ones whose
// presynaptic cell transmitted, and folds their products into the decayed g in lane (= presynaptic) order.
#include "common.h"

namespace {

// One wavefront per (fly, post).  ptr/idx: post-major CSR (int32), columns ascending; w: weights (float or half).
template <typename W>
__global__ __launch_bounds__(256) void k_pull(const int* __restrict__ ptr, const int* __restrict__ idx,
                                              const W* __restrict__ w, const float* __restrict__ x,
                                              float* __restrict__ g, int n, int batch, float a_s, int decay) {
    int64_t gw = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (gw >= (int64_t)n * batch) return;
    int lane = threadIdx.x & 63;
    int b = (int)(gw / n), row = (int)(gw % n);
    const float* xb = x + (int64_t)b * n;
    float acc = g[gw];
    if (decay) acc = acc * a_s;
    int start = ptr[row], end = ptr[row + 1];
    for (int base = start; base < end; base += kWave) {
        int e = base + lane;
        bool act = false;
        float c = 0.0f;
        if (e < end) {
            float xv = xb[idx[e]];
            if (xv != 0.0f) { act = true; c = (float)w[e] * xv; }
        }
        uint64_t m = __builtin_amdgcn_ballot_w64(act);
        while (m) {
            int j = __builtin_ctzll(m);
            acc = acc + fh_readlane(c, j);
            m &= m - 1;
        }
    }
    if (lane == 0) g[gw] = acc;
}

// The pull over packed spike bits (transmitted values 0/1): the spike vector of a fly is n bits, small enough to stay
// in the vector L1 cache, so the presynaptic lookups of all rows hit on-chip memory.  Rows are visited longest first
// over a persistent grid; rows without in-edges only decay, 64 per wavefront.
__global__ __launch_bounds__(256) void k_pull_bits(const int* __restrict__ ptr, const int* __restrict__ idx,
                                                   const float* __restrict__ w, const uint64_t* __restrict__ bits,
                                                   float* __restrict__ g, const int* __restrict__ sched, int n,
                                                   int n_full, int batch, float a_s, int decay) {
    const int lane = threadIdx.x & 63;
    const int nw = (n + 63) >> 6;
    const int64_t full_items = (int64_t)n_full * batch;
    const int empty_blocks = (n - n_full + 63) >> 6;
    const int64_t items = full_items + (int64_t)empty_blocks * batch;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t it = (int64_t)blockIdx.x * (blockDim.x >> 6) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
         it < items; it += waves) {
        if (it < full_items) {
            const int b = (int)(it / n_full), r = sched[it % n_full];
            const uint64_t* wb = bits + (int64_t)b * nw;
            float* gp = g + (int64_t)b * n + r;
            float acc = *gp;
            if (decay) acc = acc * a_s;
            const int e0 = ptr[r], e1 = ptr[r + 1];
            // 512 in-edges per round: all index loads first, then all bit lookups, then the folds, so a wavefront
            // waits for memory once per round instead of once per 64 edges (long rows are the critical path)
            constexpr int U = 8;
            for (int base = e0; base < e1; base += U * kWave) {
                int pre[U];
                bool act[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int e = base + u * kWave + lane;
                    pre[u] = e < e1 ? idx[e] : -1;
                }
#pragma unroll
                for (int u = 0; u < U; ++u)
                    act[u] = pre[u] >= 0 && ((wb[pre[u] >> 6] >> (pre[u] & 63)) & 1ull);
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    if (base + u * kWave >= e1) break;
                    const int e = base + u * kWave + lane;
                    const float c = act[u] ? w[e] : 0.0f;
                    uint64_t m = __builtin_amdgcn_ballot_w64(act[u]);
                    while (m) {
                        const int j = __builtin_ctzll(m);
                        acc = acc + fh_readlane(c, j);
                        m &= m - 1;
                    }
                }
            }
            if (lane == 0) *gp = acc;
        } else {
            const int64_t it2 = it - full_items;
            const int b = (int)(it2 / empty_blocks);
            const int k = n_full + (int)(it2 % empty_blocks) * 64 + lane;
            if (k < n && decay) {
                float* gp = g + (int64_t)b * n + sched[k];
                *gp = *gp * a_s;
            }
        }
    }
}

// CSC-compatible deterministic scatter: the contract of launch_events (neural.cu:124) without atomic adds to g.  One
// 1,024-thread workgroup per (fly, tile of kEvTile posts) holds g's tile in LDS.  It gathers the transmitting
// presynaptic cells of the walk (x != 0) in ascending order, up to kEvRound a round.  Each thread finds the in-tile
// out-edges of its cells of the round by binary searches over their ascending posts, and the round's contributions
// are numbered in edge order (ascending presynaptic cell, then the cell's own edge order: the CSC order).  They are
// applied kEvHits at a time: written out in number order (post and product), counted per post, placed by an LDS
// counting sort (the order within a post is whatever the atomics gave), each post's list sorted back into number
// order, and folded onto the tile in that order.  Rounds and parts go in number order, so every post receives its
// contributions in ascending presynaptic order, each product rounded first: the reference CPU order
// (flyverse/brain.py:679-692), bit for bit.
constexpr int kEvThreads = 1024;
constexpr int kEvWaves = kEvThreads / kWave;
constexpr int kEvTile = 4096;                     // posts per workgroup
constexpr int kEvPer = 4;                         // transmitting cells per thread per round
constexpr int kEvRound = kEvPer * kEvThreads;     // transmitting cells held per round
constexpr int kEvHits = 4096;                     // contributions applied per part
constexpr int kEvScan = 16;                       // walk entries per thread per block
constexpr int kEvBlock = kEvScan * kEvThreads;    // entry (u, t) of a block is walk entry blk + u * kEvThreads + t
constexpr int kEvSmall = 8;                       // a post's list ordered by its own thread up to this length

struct EvLds {
    float gt[kEvTile];
    int fpre[kEvRound];            // the round's transmitting cells; once searched, the placed contributions' numbers
    float fx[kEvRound];            // their x
    int fs[kEvRound];              // first in-tile out-edge
    int fb[kEvRound];              // number of the first contribution
    int bin[kEvTile + 1];          // per-post contributions of a part -> list bounds
    uint16_t hp[kEvHits];          // a part's contributions in number order: post within the tile
    float hv[kEvHits];             //                                         product
    float hc[kEvHits];             // placed contributions: product
    int tab[kEvScan * kEvWaves];   // a block's transmitting entries per (entry column, wavefront) -> their first position
    int wsum[kEvWaves];
    int total;
    int nlong;                     // a part's lists longer than kEvSmall, and their posts
    int longs[kEvHits / kEvSmall];
};

// Exclusive prefix of v over the workgroup's threads in thread order, and (in total) the sum; every thread calls it.
__device__ __forceinline__ int ev_scan(int v, int* wsum, int& total) {
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    int incl = v;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const int u = __shfl_up(incl, d, kWave);
        if (lane >= d) incl += u;
    }
    if (lane == kWave - 1) wsum[wave] = incl;
    __syncthreads();
    int base = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < kEvWaves; ++i) {
        const int s = wsum[i];
        base += i < wave ? s : 0;
        tot += s;
    }
    total = tot;
    __syncthreads();
    return base + incl - v;
}

// One round: the nf cells in L.fpre / L.fx (ascending); thread t searches cells t + j * kEvThreads.
template <typename W>
__device__ __forceinline__ void ev_round(EvLds& L, const int* __restrict__ ptr, const int* __restrict__ post,
                                         const W* __restrict__ w, int nf, int lo, int hi) {
    const int t = threadIdx.x, lane = t & (kWave - 1), wave = t / kWave;
    // the in-tile edges [s_j, s_j + k_j) of each searched cell: lower bounds of lo and of hi over its ascending
    // posts, the searches of a thread's cells in lockstep so their loads are in flight together
    int a[kEvPer], r[kEvPer], e[kEvPer];
#pragma unroll
    for (int j = 0; j < kEvPer; ++j) {
        const int f = t + j * kEvThreads;
        a[j] = 0;
        e[j] = 0;
        if (f < nf) {
            const int p = L.fpre[f];
            a[j] = ptr[p];
            e[j] = ptr[p + 1];
        }
        r[j] = e[j];
    }
    int sj[kEvPer], kj[kEvPer];
#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {
        const int bound = pass ? hi : lo;
        for (;;) {
            bool any = false;
            int mid[kEvPer], pv[kEvPer];
#pragma unroll
            for (int j = 0; j < kEvPer; ++j) {
                mid[j] = (a[j] + r[j]) >> 1;
                any = any || a[j] < r[j];
            }
            if (!any) break;
#pragma unroll
            for (int j = 0; j < kEvPer; ++j) pv[j] = a[j] < r[j] ? post[mid[j]] : 0;
#pragma unroll
            for (int j = 0; j < kEvPer; ++j) {
                if (a[j] < r[j]) {
                    if (pv[j] < bound) a[j] = mid[j] + 1;
                    else r[j] = mid[j];
                }
            }
        }
#pragma unroll
        for (int j = 0; j < kEvPer; ++j) {
            if (pass == 0) {
                sj[j] = a[j];
                r[j] = e[j];
            } else {
                kj[j] = a[j] - sj[j];
            }
        }
    }
    // numbers: cell f's contributions are fb[f] .. fb[f] + k - 1 (fb ascends with f)
    int running = 0;
#pragma unroll
    for (int j = 0; j < kEvPer; ++j) {
        int sum;
        const int base = running + ev_scan(kj[j], L.wsum, sum);
        running += sum;
        const int f = t + j * kEvThreads;
        if (f < nf) {
            L.fs[f] = sj[j];
            L.fb[f] = base;
        }
    }
    const int H = running;
    const int span = hi - lo;
    int* he = L.fpre;                                  // the cells are searched: their slots take the placed numbers
    __syncthreads();
    for (int h0 = 0; h0 < H; h0 += kEvHits) {
        const int m = min(kEvHits, H - h0);
        // the part's contributions in number order: the cell of number h is the last with fb <= h
        for (int i = t; i < m; i += kEvThreads) {
            const int h = h0 + i;
            int c0 = 0, c1 = nf;                       // fb[c0] <= h < fb[c1] (fb[nf] taken as H)
            while (c1 - c0 > 1) {
                const int cm = (c0 + c1) >> 1;
                if (L.fb[cm] <= h) c0 = cm;
                else c1 = cm;
            }
            const int edge = L.fs[c0] + (h - L.fb[c0]);
            L.hp[i] = (uint16_t)(post[edge] - lo);
            L.hv[i] = (float)w[edge] * L.fx[c0];
        }
        for (int i = t; i <= kEvTile; i += kEvThreads) L.bin[i] = 0;
        __syncthreads();
        for (int i = t; i < m; i += kEvThreads) atomicAdd(&L.bin[L.hp[i]], 1);
        __syncthreads();
        // list ends (inclusive prefix of the counts): kEvTile / kEvThreads posts per thread
        constexpr int kPosts = kEvTile / kEvThreads;
        int cnt[kPosts], own = 0;
#pragma unroll
        for (int u = 0; u < kPosts; ++u) {
            cnt[u] = L.bin[t * kPosts + u];
            own += cnt[u];
        }
        int total;
        int end = ev_scan(own, L.wsum, total);
#pragma unroll
        for (int u = 0; u < kPosts; ++u) {
            end += cnt[u];
            L.bin[t * kPosts + u] = end;
        }
        if (t == 0) L.bin[kEvTile] = total;
        __syncthreads();
        // placement from the list ends down: afterwards bin[p] is the start of post p's list and bin[p + 1] its end
        for (int i = t; i < m; i += kEvThreads) {
            const int slot = atomicSub(&L.bin[L.hp[i]], 1) - 1;
            he[slot] = i;
            L.hc[slot] = L.hv[i];
        }
        __syncthreads();
        // each post's contributions folded onto the tile in number order: a list of at most kEvSmall by its post's
        // thread, in registers; a longer one (a post of high in-degree) by a wavefront, each entry moved to its rank
        // among the list's numbers (hp / hv, spent, take the list in number order); a part has fewer than
        // kEvHits / kEvSmall long lists
        if (t == 0) L.nlong = 0;
        __syncthreads();
        for (int pt = t; pt < span; pt += kEvThreads) {
            const int s0 = L.bin[pt], k = L.bin[pt + 1] - s0;
            if (k == 0) continue;
            if (k > kEvSmall) {
                L.longs[atomicAdd(&L.nlong, 1)] = pt;
                continue;
            }
            // the list in registers; each entry's rank among the numbers, then the adds in rank order
            int key[kEvSmall];
            float val[kEvSmall];
#pragma unroll
            for (int q = 0; q < kEvSmall; ++q) {
                key[q] = q < k ? he[s0 + q] : INT_MAX;
                val[q] = q < k ? L.hc[s0 + q] : 0.0f;
            }
            int rk[kEvSmall];
#pragma unroll
            for (int i = 0; i < kEvSmall; ++i) {
                rk[i] = 0;
#pragma unroll
                for (int j = 0; j < kEvSmall; ++j) rk[i] += key[j] < key[i];
            }
            float acc = L.gt[pt];
#pragma unroll
            for (int r = 0; r < kEvSmall; ++r) {
                if (r < k) {
                    float v = 0.0f;
#pragma unroll
                    for (int i = 0; i < kEvSmall; ++i) v = rk[i] == r ? val[i] : v;
                    acc = acc + v;
                }
            }
            L.gt[pt] = acc;
        }
        __syncthreads();
        const int nlong = L.nlong;
        for (int li = wave; li < nlong; li += kEvWaves) {
            const int pt = L.longs[li];
            const int s0 = L.bin[pt], s1 = L.bin[pt + 1];
            for (int s = s0 + lane; s < s1; s += kWave) {
                const int key = he[s];
                int rank = 0;
                int q = s0;
                for (; q + 8 <= s1; q += 8) {          // eight reads in flight
                    int v[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) v[u] = he[q + u];
#pragma unroll
                    for (int u = 0; u < 8; ++u) rank += v[u] < key;
                }
                for (; q < s1; ++q) rank += he[q] < key;
                L.hp[s0 + rank] = (uint16_t)key;
                L.hv[s0 + rank] = L.hc[s];
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (lane == 0) {
                float acc = L.gt[pt];
                int s = s0;
                for (; s + 8 <= s1; s += 8) {          // eight reads in flight, then their adds in order
                    float v[8];
#pragma unroll
                    for (int q = 0; q < 8; ++q) v[q] = L.hv[s + q];
#pragma unroll
                    for (int q = 0; q < 8; ++q) acc = acc + v[q];
                }
                for (; s < s1; ++s) acc = acc + L.hv[s];
                L.gt[pt] = acc;
            }
        }
        __syncthreads();
    }
}

template <typename W>
__global__ __launch_bounds__(kEvThreads) void k_events(const int* __restrict__ ptr, const int* __restrict__ post,
                                                       const W* __restrict__ w, const float* __restrict__ x,
                                                       float* __restrict__ g, const int* __restrict__ pre_ids, int n,
                                                       int count) {
    extern __shared__ __align__(16) unsigned char ev_lds[];
    EvLds& L = *reinterpret_cast<EvLds*>(ev_lds);
    const int t = threadIdx.x, lane = t & (kWave - 1), wave = t / kWave;
    const int lo = blockIdx.x * kEvTile, hi = min(lo + kEvTile, n);
    float* gb = g + (int64_t)blockIdx.y * n;
    const float* xb = x + (int64_t)blockIdx.y * n;
    for (int i = t; i < hi - lo; i += kEvThreads) L.gt[i] = gb[lo + i];
    const int size = count < 0 ? n : count;
    int nf = 0;                                    // cells held for the round (the same in every thread)
    for (int blk = 0; blk < size; blk += kEvBlock) {
        // entry (u, t) = walk entry blk + u * kEvThreads + t: each wavefront's loads of one column are contiguous.
        // The block's transmitting entries in walk order get positions: column-major over (column, wavefront),
        // then lane order within the wavefront.
        auto entry = [&](int u, int& pre) -> float {
            const int k = blk + u * kEvThreads + t;
            if (k >= size) return 0.0f;
            pre = count < 0 ? k : pre_ids[k];
            return xb[pre];
        };
#pragma unroll
        for (int u = 0; u < kEvScan; ++u) {
            int pre;
            const uint64_t bal = __builtin_amdgcn_ballot_w64(entry(u, pre) != 0.0f);
            if (lane == 0) L.tab[u * kEvWaves + wave] = __builtin_popcountll(bal);
        }
        __syncthreads();
        if (wave == 0) {                           // exclusive prefix of the kEvScan * kEvWaves counts
            constexpr int kPer = kEvScan * kEvWaves / kWave;
            int c[kPer], own = 0;
#pragma unroll
            for (int q = 0; q < kPer; ++q) {
                c[q] = L.tab[lane * kPer + q];
                own += c[q];
            }
            int incl = own;
#pragma unroll
            for (int d = 1; d < kWave; d <<= 1) {
                const int v = __shfl_up(incl, d, kWave);
                if (lane >= d) incl += v;
            }
            int run = incl - own;
#pragma unroll
            for (int q = 0; q < kPer; ++q) {
                L.tab[lane * kPer + q] = run;
                run += c[q];
            }
            if (lane == kWave - 1) L.total = incl;
        }
        __syncthreads();
        const int total = L.total;
        for (int done = 0; done < total;) {
            const int take = min(total - done, kEvRound - nf);   // block positions [done, done + take) go next
#pragma unroll
            for (int u = 0; u < kEvScan; ++u) {
                int pre = 0;
                const float xv = entry(u, pre);
                const bool fired = xv != 0.0f;
                const uint64_t bal = __builtin_amdgcn_ballot_w64(fired);
                const int q = L.tab[u * kEvWaves + wave] +
                              (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                             __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
                if (fired && q >= done && q < done + take) {
                    L.fpre[nf + q - done] = pre;
                    L.fx[nf + q - done] = xv;
                }
            }
            nf += take;
            done += take;
            __syncthreads();
            if (nf == kEvRound) {
                ev_round(L, ptr, post, w, nf, lo, hi);
                nf = 0;
            }
        }
    }
    if (nf) ev_round(L, ptr, post, w, nf, lo, hi);
    __syncthreads();
    for (int i = t; i < hi - lo; i += kEvThreads) gb[lo + i] = L.gt[i];
}

// The synaptic input of the clocked step (flyverse/brain.py:903-921): per (fly, row) wavefront, g = g * a_s[row]
// (the row's decay this phase, 1 when its clock does not tick); then, when any clock class forms its product this
// step, g = g + y with y the row's product if its class ticks (its in-edges' w * x_in[pre] for x_in != 0, folded from
// +0 in ascending presynaptic order, x_in the class's input: the delayed spikes for clock 1, the class's
// accumulated spikes otherwise) and +0 for the other rows (the zero rows of the other classes' products).
constexpr int kMaxClockClasses = 8;
struct ClockInputs {
    const float* x[kMaxClockClasses];
};

__global__ __launch_bounds__(256) void k_pull_clocked(const int* __restrict__ ptr, const int* __restrict__ idx,
                                                      const float* __restrict__ w, ClockInputs xin,
                                                      const unsigned char* __restrict__ row_class,
                                                      const float* __restrict__ a_s, float* __restrict__ g, int n,
                                                      int batch, int tick_mask) {
    const int64_t gw = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (gw >= (int64_t)n * batch) return;
    const int lane = threadIdx.x & 63;
    const int b = (int)(gw / n), row = (int)(gw % n);
    float acc = g[gw] * a_s[row];
    if (tick_mask) {
        const int cls = row_class[row];
        float y = 0.0f;
        if ((tick_mask >> cls) & 1) {
            const float* xb = xin.x[cls] + (int64_t)b * n;
            const int start = ptr[row], end = ptr[row + 1];
            for (int base = start; base < end; base += kWave) {
                const int e = base + lane;
                bool act = false;
                float c = 0.0f;
                if (e < end) {
                    const float xv = xb[idx[e]];
                    if (xv != 0.0f) { act = true; c = w[e] * xv; }
                }
                uint64_t m = __builtin_amdgcn_ballot_w64(act);
                while (m) {
                    const int j = __builtin_ctzll(m);
                    y = y + fh_readlane(c, j);
                    m &= m - 1;
                }
            }
        }
        acc = acc + y;
    }
    if (lane == 0) g[gw] = acc;
}

// The clock accumulators (flyverse/brain.py:914-918): acc_j += x for the classes in add_mask, then acc_j = 0 for
// those in clear_mask.
__global__ __launch_bounds__(256) void k_clock_acc(const float* __restrict__ x, ClockInputs acc, int64_t count,
                                                   int nacc, int add_mask, int clear_mask) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    for (int j = 0; j < nacc; ++j) {
        float* a = const_cast<float*>(acc.x[j]);
        if ((add_mask >> j) & 1) a[t] = a[t] + x[t];
        if ((clear_mask >> j) & 1) a[t] = 0.0f;
    }
}

}  // namespace

// g[b, post] = g[b, post] * a_s (if flags & 2 is clear) + sum over in-edges in ascending presynaptic order of
// w * x[b, pre] for x != 0.  p = [ptr int32 (n+1), idx int32 (nnz), w (float, or half if flags & 1),
// x float (B, n), g float (B, n)].
FH_EXPORT int launch_pull(void** p, int n, int batch, float a_s, int flags, void* stream) {
    if (n < 0 || batch < 0) return FH_ERR_ARGS;
    int64_t waves = (int64_t)n * batch;
    if (waves == 0) return 0;
    if (waves * 64 >= (1ll << 40)) return FH_ERR_SIZE;
    unsigned blocks = (unsigned)((waves + 3) / 4);
    int decay = (flags & 2) ? 0 : 1;
    if (flags & 1)
        hipLaunchKernelGGL(k_pull<__half>, dim3(blocks), dim3(256), 0, (hipStream_t)stream, (const int*)p[0],
                           (const int*)p[1], (const __half*)p[2], (const float*)p[3], (float*)p[4], n, batch, a_s, decay);
    else
        hipLaunchKernelGGL(k_pull<float>, dim3(blocks), dim3(256), 0, (hipStream_t)stream, (const int*)p[0],
                           (const int*)p[1], (const float*)p[2], (const float*)p[3], (float*)p[4], n, batch, a_s, decay);
    FH_LAUNCH_CHECK();
}

// launch_pull for 0/1 spikes given as packed bits.  p = [ptr (n+1), idx, w (float), bits uint64 (batch,
// ceil(n/64)), g (batch, n), sched (n: the rows with in-edges first, longest first, then the rest)];
// n_full = rows with in-edges; grid = workgroups of the persistent launch.
FH_EXPORT int launch_pull_bits(void** p, int n, int n_full, int batch, float a_s, int flags, int grid, void* stream) {
    if (n < 0 || batch < 0 || n_full < 0 || n_full > n || grid <= 0) return FH_ERR_ARGS;
    if ((int64_t)n * batch == 0) return 0;
    hipLaunchKernelGGL(k_pull_bits, dim3(grid), dim3(256), 0, (hipStream_t)stream, (const int*)p[0], (const int*)p[1],
                       (const float*)p[2], (const uint64_t*)p[3], (float*)p[4], (const int*)p[5], n, n_full, batch,
                       a_s, (flags & 2) ? 0 : 1);
    FH_LAUNCH_CHECK();
}

// Same contract as flyverse/kernels/neural.cu:124 (launch_events): p = [ptr, post, w, x, g, pre_ids] with the
// pre-major CSC of Brain.set_weights (flyverse/brain.py:530-540); count < 0 walks every presynaptic cell, else the
// `count` ascending entries of pre_ids.  Deterministic: the result equals the reference CPU order bit for bit.
FH_EXPORT int launch_events(void** p, int n, int batch, int half, int count, void* stream) {
    if (n < 0 || batch < 0) return FH_ERR_ARGS;
    if ((int64_t)n * batch == 0) return 0;
    if (batch > 65535) return FH_ERR_ARGS;
    const dim3 grid((unsigned)((n + kEvTile - 1) / kEvTile), (unsigned)batch);
    const size_t lds = sizeof(EvLds);
    hipError_t e = hipFuncSetAttribute(half ? (const void*)k_events<__half> : (const void*)k_events<float>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return (int)e;
    if (half)
        hipLaunchKernelGGL(k_events<__half>, grid, dim3(kEvThreads), lds, (hipStream_t)stream, (const int*)p[0],
                           (const int*)p[1], (const __half*)p[2], (const float*)p[3], (float*)p[4], (const int*)p[5], n,
                           count);
    else
        hipLaunchKernelGGL(k_events<float>, grid, dim3(kEvThreads), lds, (hipStream_t)stream, (const int*)p[0],
                           (const int*)p[1], (const float*)p[2], (const float*)p[3], (float*)p[4], (const int*)p[5], n,
                           count);
    FH_LAUNCH_CHECK();
}

// The clocked synaptic input.  p = [ptr (n+1), idx, w (float), g (batch, n), row_class uint8 (n: the row's clock
// class), a_s (n: this phase's decay per row), x_0 .. x_{ncls-1} (batch, n): each class's input].  tick_mask: bit j
// set when class j forms its product this step (0: decay only).  Replaces the g.mul_(c["a_s"]) and the _matmul_add
// calls of Brain._step_clocked, flyverse/brain.py:905-921, with a deterministic product order.
FH_EXPORT int launch_pull_clocked(void** p, int n, int batch, int ncls, int tick_mask, void* stream) {
    if (!p || n < 0 || batch < 0 || ncls < 1 || ncls > kMaxClockClasses || (tick_mask >> ncls)) return FH_ERR_ARGS;
    const int64_t waves = (int64_t)n * batch;
    if (waves == 0) return 0;
    if (waves * 64 >= (1ll << 40)) return FH_ERR_SIZE;
    ClockInputs xin{};
    for (int j = 0; j < ncls; ++j) {
        xin.x[j] = (const float*)p[6 + j];
        if (((tick_mask >> j) & 1) && !xin.x[j]) return FH_ERR_ARGS;
    }
    hipLaunchKernelGGL(k_pull_clocked, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, (hipStream_t)stream,
                       (const int*)p[0], (const int*)p[1], (const float*)p[2], xin, (const unsigned char*)p[4],
                       (const float*)p[5], (float*)p[3], n, batch, tick_mask);
    FH_LAUNCH_CHECK();
}

// The clock accumulators: p = [x (count), acc_0 .. acc_{nacc-1} (count each)]; acc_j += x for j in add_mask, then
// acc_j = 0 for j in clear_mask (flyverse/brain.py:914-918).
FH_EXPORT int launch_clock_acc(void** p, int64_t count, int nacc, int add_mask, int clear_mask, void* stream) {
    if (!p || count < 0 || nacc < 0 || nacc > kMaxClockClasses || (add_mask >> nacc) || (clear_mask >> nacc))
        return FH_ERR_ARGS;
    if (count == 0 || nacc == 0) return 0;
    ClockInputs acc{};
    for (int j = 0; j < nacc; ++j) {
        acc.x[j] = (const float*)p[1 + j];
        if (!acc.x[j]) return FH_ERR_ARGS;
    }
    hipLaunchKernelGGL(k_clock_acc, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       (const float*)p[0], acc, count, nacc, add_mask, clear_mask);
    FH_LAUNCH_CHECK();
}
