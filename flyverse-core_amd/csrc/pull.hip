// Deterministic synaptic input for the plain (B, N) layout: no atomics, the reference CPU order.
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

// CSC-compatible deterministic scatter: the contract of launch_events (neural.cu:124) without atomics.  One
// wavefront per (fly, tile of kTile posts) holds g's tile in LDS and walks the transmitting presynaptic cells in
// ascending order; a cell's out-edges have distinct posts, so one pass over them needs no ordering, and the walk
// itself gives every post its contributions in ascending presynaptic order.  Work grows with the number of tiles,
// so this is the small-graph path; Brain uses the pull above.
constexpr int kTile = 2048;

template <typename W>
__global__ __launch_bounds__(64) void k_events(const int* __restrict__ ptr, const int* __restrict__ post,
                                               const W* __restrict__ w, const float* __restrict__ x,
                                               float* __restrict__ g, const int* __restrict__ pre_ids, int n,
                                               int count) {
    __shared__ float tile[kTile];
    int lane = threadIdx.x;
    int b = blockIdx.y;
    int lo = blockIdx.x * kTile, hi = min(lo + kTile, n);
    float* gb = g + (int64_t)b * n;
    const float* xb = x + (int64_t)b * n;
    for (int k = lane; k < hi - lo; k += kWave) tile[k] = gb[lo + k];
    __syncthreads();
    int size = count < 0 ? n : count;
    for (int base = 0; base < size; base += kWave) {
        int k = base + lane;
        int pre = 0;
        float xv = 0.0f;
        if (k < size) { pre = count < 0 ? k : pre_ids[k]; xv = xb[pre]; }
        uint64_t m = __builtin_amdgcn_ballot_w64(xv != 0.0f);
        while (m) {
            int j = __builtin_ctzll(m);
            m &= m - 1;
            int p = __builtin_amdgcn_readlane(pre, j);
            float xp = fh_readlane(xv, j);
            int s = ptr[p], e = ptr[p + 1];
            for (int q0 = s; q0 < e; q0 += kWave) {
                int q = q0 + lane;
                int t = q < e ? post[q] : n;
                if (t >= lo && t < hi) {
                    float c = (float)w[q] * xp;
                    tile[t - lo] = tile[t - lo] + c;
                }
                // posts are ascending within the column: stop once this round has passed the tile
                if (__builtin_amdgcn_readlane(t, kWave - 1) >= hi) break;
            }
        }
    }
    __syncthreads();
    for (int k = lane; k < hi - lo; k += kWave) gb[lo + k] = tile[k];
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
    dim3 grid((unsigned)((n + kTile - 1) / kTile), (unsigned)batch);
    if (half)
        hipLaunchKernelGGL(k_events<__half>, grid, dim3(64), 0, (hipStream_t)stream, (const int*)p[0], (const int*)p[1],
                           (const __half*)p[2], (const float*)p[3], (float*)p[4], (const int*)p[5], n, count);
    else
        hipLaunchKernelGGL(k_events<float>, grid, dim3(64), 0, (hipStream_t)stream, (const int*)p[0], (const int*)p[1],
                           (const float*)p[2], (const float*)p[3], (float*)p[4], (const int*)p[5], n, count);
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
