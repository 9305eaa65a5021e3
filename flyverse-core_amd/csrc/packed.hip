// Packed uneven subgraphs: many independent LIF graphs (one fly each, sizes uneven) are packed block-diagonally into
// one CSR with offset tables and stepped together.  An addition: the reference runs every Connectome.subset as its
// own Brain (flyverse/connectome.py:469-481).
//
// Per step and target row, the synaptic input keeps the reference CPU order (flyverse/brain.py:832-833, 679-692):
// g = g*a_s, then the in-edges whose presynaptic cell fired n_delay steps ago, in ascending presynaptic order; the
// LIF update is flyverse/brain.py:838-868 with the Shiu settings.  Spikes are bits in a ring of 64 slots (one bit per
// cell per step).  Rows are independent for n_delay consecutive steps (their input comes from spikes at least
// n_delay steps old), so the grid synchronises only once per window of <= n_delay steps: a wavefront takes a row,
// folds its in-edges (lanes over edges: ballot of the fired bits, then the fired weights added in edge order) and
// updates it for every step of the window with its state in registers.  Each window also clears the ring slots the
// next window writes; those slots are older than anything the current window reads (64 > 2 * n_delay + window).
//
// The Poisson uniform of a cell is Philox keyed by its instance's seed with counter (local cell index, step, 0, 0):
// every instance draws exactly what it would draw run alone as a one-fly brain.
//
// Two drivers share the window body: launch_packed_window runs one window per launch (the kernel boundary is the
// synchronisation; a rollout is a captured graph of such launches with the step base on the device), and
// launch_packed_run runs a whole rollout in one cooperative launch with a grid barrier between windows.
#include "common.h"

#include <type_traits>

namespace {

constexpr int kSlots = 64;

struct Packed {
    const int* ptr;          // (n + 1) post-major CSR over all packed cells
    const int* idx;          // (nnz) presynaptic cell (packed index)
    const float* w;          // (nnz)
    const int* sched;        // (n) rows, longest first
    float *v, *g, *refrac, *rate, *counts;
    const float* drive;      // (n)
    const float* pois;       // (n) per-step Poisson probability
    const int* inst;         // (n) instance of each cell
    const int* local;        // (n) index of the cell within its instance
    const uint64_t* seeds;   // (instances) Philox keys
    uint64_t* ring;          // (kSlots, words) fired bits
    unsigned* bar_count;     // grid barrier arrivals
    unsigned* bar_gen;       // grid barrier generation
    int* error;              // (4) set when a barrier wait exceeds its bound: flag, arrivals seen, grid, window
    const float* P;          // (12) v_rest, v_reset, v_th, a_m, dt, t_ref, -, -, -, a_r, rate_gain, a_s
    int n, words, steps, window, delay, t0;
    uint64_t spin_ticks;     // bound of one barrier wait, in wall-clock ticks
};

// Grid-wide barrier over a cooperative launch.  Release fence before arriving (this window's ring bits and state are
// visible device-wide), acquire fence after leaving (no stale vector-L1 lines of the ring survive into the next
// window).  The wait is bounded in wall-clock time; once any wait has timed out, every later barrier returns at once,
// so a grid that is not fully resident ends the kernel with the error flag set instead of hanging.
__device__ __forceinline__ void grid_barrier(const Packed& a, unsigned nblocks, int T0) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const unsigned g0 = __hip_atomic_load(a.bar_gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (atomicAdd(a.bar_count, 1u) == nblocks - 1) {
            __hip_atomic_store(a.bar_count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            atomicAdd(a.bar_gen, 1u);
        } else if (__hip_atomic_load(a.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
            const uint64_t start = wall_clock64();
            while (__hip_atomic_load(a.bar_gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g0) {
                if (wall_clock64() - start > a.spin_ticks) {
                    if (atomicExch(a.error, 1) == 0) {
                        a.error[1] = (int)__hip_atomic_load(a.bar_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        a.error[2] = (int)nblocks;
                        a.error[3] = T0;
                    }
                    break;
                }
                __builtin_amdgcn_s_sleep(8);
            }
        }
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// One window: steps T0 .. T0 + K - 1 of every row.  Also clears the ring slots of the window that follows it.
__device__ __forceinline__ void packed_window(const Packed& a, int T0, int K) {
    const int lane = threadIdx.x & 63;
    const int wave = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nwaves = gridDim.x * 4;
    const int gthread = blockIdx.x * blockDim.x + threadIdx.x;
    const int nthreads = gridDim.x * blockDim.x;
    const float v_rest = a.P[0], v_reset = a.P[1], v_th = a.P[2], am = a.P[3], dt = a.P[4], t_ref = a.P[5];
    const float ar = a.P[9], rg = a.P[10], a_s = a.P[11];
    // clear the slots the next window writes (steps T0 + K .. T0 + K + window - 1)
    for (int64_t q = gthread; q < (int64_t)a.window * a.words; q += nthreads) {
        const int s = (T0 + K + (int)(q / a.words)) & (kSlots - 1);
        a.ring[(int64_t)s * a.words + q % a.words] = 0ull;
    }
    for (int si = wave; si < a.n; si += nwaves) {
        const int r = a.sched[si];
        const int e0 = a.ptr[r], e1 = a.ptr[r + 1];
        float v = a.v[r], g = a.g[r], ref = a.refrac[r], rate = a.rate[r], cnt = a.counts[r];
        const float drive = a.drive[r], p = a.pois[r];
        const uint64_t seed = a.seeds[a.inst[r]];
        const uint32_t loc = (uint32_t)a.local[r];
        // LIF update of step k (uniform across the wavefront: every lane computes the same row)
        auto lif_step = [&](int k, float acc) {
            const int t = T0 + k;
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
            if (p > 0.0f && fh_uniform(seed, (uint32_t)t, loc, 0u, 0u) < p) spike = 1.0f;
            const bool fired = spike > 0.0f;
            if (fired) { vn = v_reset; rn = t_ref; }
            float rr = rate * ar;
            float rs = spike * rg;
            g = acc; v = vn; ref = rn; rate = rr + rs; cnt = cnt + spike;
            if (fired && lane == 0)
                atomicOr((unsigned long long*)(a.ring + (int64_t)(t & (kSlots - 1)) * a.words + (r >> 6)),
                         1ull << (r & 63));
        };
        auto slot_of = [&](int k) {
            return a.ring + (int64_t)((T0 + k - a.delay) & (kSlots - 1)) * a.words;
        };
        // at most NC * 64 in-edges: presynaptic indices and weights stay in registers for the window, and the bit
        // words of step k + 1 are gathered before step k is folded (two buffers alternating, no copies)
        auto cached = [&](auto nc_tag) {
            constexpr int NC = decltype(nc_tag)::value;
            int pre[NC];
            float wc[NC];
            bool val[NC];
            const int last = max(e1 - 1, e0);
#pragma unroll
            for (int u = 0; u < NC; ++u) {
                const int e = e0 + u * kWave + lane;
                val[u] = e < e1;
                pre[u] = e1 > e0 ? a.idx[min(e, last)] : 0;
                wc[u] = e1 > e0 ? a.w[min(e, last)] : 0.0f;
            }
            const int nch = (e1 - e0 + kWave - 1) / kWave;
            uint64_t wa[NC], wb[NC];
            auto look = [&](int k, uint64_t* wd) {
                const uint64_t* slot = slot_of(k);
#pragma unroll
                for (int u = 0; u < NC; ++u) wd[u] = slot[pre[u] >> 6];
            };
            auto step = [&](int k, const uint64_t* wd) {
                float acc = g * a_s;
#pragma unroll
                for (int u = 0; u < NC; ++u) {
                    if (u < nch) {
                        const bool act = val[u] && ((wd[u] >> (pre[u] & 63)) & 1ull);
                        uint64_t m = __builtin_amdgcn_ballot_w64(act);
                        while (m) {
                            const int j = __builtin_ctzll(m);
                            acc = acc + fh_readlane(wc[u], j);
                            m &= m - 1;
                        }
                    }
                }
                lif_step(k, acc);
            };
            look(0, wa);
            for (int k = 0; k < K; k += 2) {
                look(min(k + 1, K - 1), wb);
                step(k, wa);
                look(min(k + 2, K - 1), wa);
                if (k + 1 < K) step(k + 1, wb);
            }
        };
        if (e1 - e0 <= 4 * kWave) {
            cached(std::integral_constant<int, 4>{});
        } else if (e1 - e0 <= 8 * kWave) {
            cached(std::integral_constant<int, 8>{});
        } else {
            for (int k = 0; k < K; ++k) {
                const uint64_t* slot = slot_of(k);
                float acc = g * a_s;
                for (int base = e0; base < e1; base += 4 * kWave) {
                    int pre[4];
                    bool act[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int e = base + u * kWave + lane;
                        pre[u] = e < e1 ? a.idx[e] : -1;
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        act[u] = pre[u] >= 0 && ((slot[pre[u] >> 6] >> (pre[u] & 63)) & 1ull);
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
                lif_step(k, acc);
            }
        }
        if (lane == 0) { a.v[r] = v; a.g[r] = g; a.refrac[r] = ref; a.rate[r] = rate; a.counts[r] = cnt; }
    }
}

__global__ __launch_bounds__(256) void k_packed_run(Packed a) {
    for (int T0 = a.t0; T0 < a.t0 + a.steps; T0 += a.window) {
        packed_window(a, T0, min(a.window, a.t0 + a.steps - T0));
        grid_barrier(a, gridDim.x, T0);
    }
}

// One window per launch; the window's first step is *t_base + off.
__global__ __launch_bounds__(256) void k_packed_window(Packed a, const int* t_base, int off, int K) {
    packed_window(a, *t_base + off, K);
}

__global__ void k_advance(int* t_base, int steps) {
    if (threadIdx.x == 0) *t_base += steps;
}

static void fill_packed(Packed& a, void** p) {
    a.ptr = (const int*)p[0]; a.idx = (const int*)p[1]; a.w = (const float*)p[2]; a.sched = (const int*)p[3];
    a.v = (float*)p[4]; a.g = (float*)p[5]; a.refrac = (float*)p[6]; a.rate = (float*)p[7]; a.counts = (float*)p[8];
    a.drive = (const float*)p[9]; a.pois = (const float*)p[10]; a.inst = (const int*)p[11];
    a.local = (const int*)p[12]; a.seeds = (const uint64_t*)p[13]; a.ring = (uint64_t*)p[14];
    a.bar_count = (unsigned*)p[15]; a.bar_gen = (unsigned*)p[15] + 1; a.error = (int*)p[16];
    a.P = (const float*)p[17];
}

// Workgroups of k_packed_run that are resident together: one per CU fewer than the occupancy query reports (on
// MI355X it reports 7 per CU for this kernel, yet at most 6 per CU were ever resident at once; the bounded barrier
// wait reports such a shortfall as an error instead of hanging).
static int resident_grid(int per_cu, int cus) { return (per_cu > 1 ? per_cu - 1 : per_cu) * cus; }

static bool packed_shape_ok(const Packed& a) {
    return a.n >= 0 && a.steps >= 0 && a.window >= 1 && a.window <= a.delay && 2 * a.delay + a.window <= kSlots &&
           2 * a.window + a.delay <= kSlots && a.t0 >= 0;
}

}  // namespace

// Run `steps` LIF steps (from global step t0) of the packed graph in one cooperative launch.
// p = [ptr, idx, w, sched, v, g, refrac, rate, counts, drive, pois, inst, local, seeds, ring (64, words),
//      bar (2 uint32: count, generation), error (4 int32), P (12)];
// dims = [n, steps, window, delay, grid, t0, barrier bound in ms].  Requires window <= delay and
// 2 * delay + window <= 64; a run that continues an earlier one (t0 > 0) must use the same window, whose last window
// cleared the next slots.
FH_EXPORT int launch_packed_run(void** p, const int* dims, void* stream) {
    Packed a;
    fill_packed(a, p);
    a.n = dims[0]; a.steps = dims[1]; a.window = dims[2]; a.delay = dims[3];
    const int grid = dims[4];
    a.t0 = dims[5];
    a.words = (a.n + 63) / 64;
    if (!packed_shape_ok(a) || grid <= 0 || dims[6] <= 0) return FH_ERR_ARGS;
    if (a.n == 0 || a.steps == 0) return 0;
    int per_cu = 0, cus = 0, dev = 0, khz = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    e = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev);
    if (e != hipSuccess) return (int)e;
    a.spin_ticks = (uint64_t)khz * (uint64_t)dims[6];
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_packed_run, 256, 0);
    if (e != hipSuccess) return (int)e;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return (int)e;
    if (grid > resident_grid(per_cu, cus)) return FH_ERR_SIZE;   // every block must be resident for the barrier
    void* args[] = {&a};
    e = hipLaunchCooperativeKernel((const void*)k_packed_run, dim3(grid), dim3(256), args, 0, (hipStream_t)stream);
    if (e != hipSuccess) return (int)e;
    FH_LAUNCH_CHECK();
}

// One window of K steps starting at step *t_base + off (t_base: int32 on the device, read when the kernel runs, so a
// captured graph of windows replays at advancing steps).  p as launch_packed_run (bar and error unused);
// dims = [n, K, window, delay, grid, off].  The kernel boundary orders the window against the next one.
FH_EXPORT int launch_packed_window(void** p, const int* dims, const int* t_base, void* stream) {
    Packed a;
    fill_packed(a, p);
    a.n = dims[0]; a.steps = dims[1]; a.window = dims[2]; a.delay = dims[3];
    const int grid = dims[4], off = dims[5];
    a.t0 = 0;
    a.words = (a.n + 63) / 64;
    a.spin_ticks = 0;
    if (!packed_shape_ok(a) || a.steps > a.window || grid <= 0 || off < 0 || t_base == nullptr) return FH_ERR_ARGS;
    if (a.n == 0 || a.steps == 0) return 0;
    hipLaunchKernelGGL(k_packed_window, dim3(grid), dim3(256), 0, (hipStream_t)stream, a, t_base, off, a.steps);
    FH_LAUNCH_CHECK();
}

// *t_base += steps on the stream (the last node of a captured rollout).
FH_EXPORT int launch_packed_advance(int* t_base, int steps, void* stream) {
    if (t_base == nullptr) return FH_ERR_ARGS;
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, (hipStream_t)stream, t_base, steps);
    FH_LAUNCH_CHECK();
}

// The largest grid launch_packed_run accepts (workgroups resident together).
FH_EXPORT int packed_max_grid(int* out) {
    int per_cu = 0, cus = 0, dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_packed_run, 256, 0);
    if (e != hipSuccess) return (int)e;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return (int)e;
    *out = resident_grid(per_cu, cus);
    return 0;
}
