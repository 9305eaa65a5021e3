// Packed uneven subgraphs: many independent LIF graphs (one fly each, sizes uneven) are packed block-diagonally into
// one CSR with offset tables and stepped together.  An addition: the reference runs every Connectome.subset as its
// own Brain (flyverse/connectome.py:469-481).
//
// Per step and target row, the synaptic input keeps the reference CPU order (flyverse/brain.py:832-833, 679-692):
// g = g*a_s, then the in-edges whose presynaptic cell fired n_delay steps ago, in ascending presynaptic order; the
// LIF update is flyverse/brain.py:838-868 with the Shiu settings.  Spikes are kept as one 64-bit history word per
// cell: bit (t mod 64) says whether the cell fired at step t, so a word holds the cell's last 64 steps.  Rows are
// independent for n_delay consecutive steps (their input comes from spikes at least n_delay steps old), so the grid
// synchronises only once per window of <= n_delay steps; within a window a wavefront owns a block of 64 cells (see
// packed_window).  A window rewrites the bits of its own steps in its cells' words and reads only bits of steps at
// least n_delay older (64 >= 2 * n_delay + window keeps the two sets of bit positions apart), so a word that is read
// while its owner rewrites it gives the reader the same bits either way.
//
// The Poisson uniform of a cell is Philox keyed by its instance's seed with counter (local cell index, step, 0, 0):
// every instance draws exactly what it would draw run alone as a one-fly brain.
//
// Two drivers share the window body: launch_packed_window runs one window per launch (the kernel boundary is the
// synchronisation; a rollout is a captured graph of such launches with the step base on the device), and
// launch_packed_run runs a whole rollout in one cooperative launch with a grid barrier between windows.
#include "chain.h"

#include <type_traits>

namespace {

using fh_chain::kSlots;

struct Packed {
    const int* ptr;          // (n + 1) post-major CSR over all packed cells
    const int* idx;          // (nnz) presynaptic cell (packed index)
    const float* w;          // (nnz)
    const int* sched;        // (words) blocks of 64 cells, most in-edges first
    float *v, *g, *refrac, *rate, *counts;
    const float* drive;      // (n)
    const float* pois;       // (n) per-step Poisson probability
    const int* inst;         // (n) instance of each cell
    const int* local;        // (n) index of the cell within its instance
    const uint64_t* seeds;   // (instances) Philox keys
    uint64_t* ring;          // (n) spike history words: bit (t mod 64) of word c = cell c fired at step t
    unsigned* bar_count;     // grid barrier arrivals
    unsigned* bar_gen;       // grid barrier generation
    int* error;              // (4) set when a barrier wait exceeds its bound: flag, arrivals seen, grid, window
    const float* P;          // (12) v_rest, v_reset, v_th, a_m, dt, t_ref, -, -, -, a_r, rate_gain, a_s
    const int* long_rows;    // (n_long) cells with more than 512 in-edges, most first
    const int* long_id;      // (n) index of a long cell in long_rows, -1 for the others
    float* long_g;           // (n_long, kMaxWindow) the long cells' synaptic chains of the current window
    int n, words, steps, window, delay, t0, n_long, n_vlong;
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

using fh_chain::Edges;
using fh_chain::kMaxWindow;
using fh_chain::kShortEdges;
using fh_chain::Window;
constexpr int kVlongChunks = 36;        // chunks per wavefront in k_packed_vlong (cells of up to 9216 in-edges)

__device__ __forceinline__ Edges edges_of(const Packed& a) { return Edges{a.idx, a.w, a.ring}; }
__device__ __forceinline__ Window window_of(const Packed& a, int t0, int k) {
    return Window((uint32_t)t0, k, a.delay, a.P[11]);
}

// Phase A of a block of 64 consecutive cells c0 .. c0 + nc - 1, lanes over each cell's in-edges in turn: the
// synaptic chain of the window (g_chain_*) of every cell, handed to out(i, gk) (g_k in lane k of gk); it needs
// nothing of the membrane, so it runs ahead of the LIF update.  With LONG_BUF the chains of the long cells (more than
// 512 in-edges) were computed before, by k_packed_long / k_packed_vlong: SKIP_LONG leaves them out, otherwise they
// are read from long_g.
template <bool LONG_BUF, bool SKIP_LONG, class Out>
__device__ __forceinline__ void block_chains(const Packed& a, const Window& W, const Edges& E, int c0, int nc,
                                             int lane, Out out) {
    const int pl0 = a.ptr[c0 + min(lane, nc)], pl1 = a.ptr[c0 + min(lane + 1, nc)];
    const float gl = lane < nc ? a.g[c0 + lane] : 0.0f;
    for (int i = 0; i < nc; ++i) {
        const int e0 = __builtin_amdgcn_readlane(pl0, i), e1 = __builtin_amdgcn_readlane(pl1, i);
        const float g0 = fh_readlane(gl, i);
        float gk;
        if (e1 - e0 <= kWave) {
            gk = fh_chain::g_chain_cached<1>(E, W, e0, e1, g0, lane);
        } else if (e1 - e0 <= 2 * kWave) {
            gk = fh_chain::g_chain_cached<2>(E, W, e0, e1, g0, lane);
        } else if (e1 - e0 <= 4 * kWave) {
            gk = fh_chain::g_chain_cached<4>(E, W, e0, e1, g0, lane);
        } else if (e1 - e0 <= kShortEdges) {
            gk = fh_chain::g_chain_cached<8>(E, W, e0, e1, g0, lane);
        } else {
            const int li = LONG_BUF && a.n_long > 0 ? a.long_id[c0 + i] : -1;
            if (li >= 0) {
                if (SKIP_LONG) continue;
                gk = lane < W.K ? a.long_g[(int64_t)li * kMaxWindow + lane] : 0.0f;
            } else {
                gk = fh_chain::g_chain_stream(E, W, e0, e1, g0, lane);
            }
        }
        out(i, gk);
    }
}

// Phase B of the block, lane = cell c0 + lane: the LIF update of flyverse/brain.py:838-868 (Shiu settings) for every
// step of the window with g_k = gk(k); the lane then rewrites the window's bits of its own cell's history word with
// one plain store (the block's cells are this wavefront's alone, so no atomics are needed).
template <class G>
__device__ __forceinline__ void block_lif(const Packed& a, int T0, int K, int c0, int nc, int lane, G gk) {
    const float v_rest = a.P[0], v_reset = a.P[1], v_th = a.P[2], am = a.P[3], dt = a.P[4], t_ref = a.P[5];
    const float ar = a.P[9], rg = a.P[10];
    const bool ok = lane < nc;
    const int r = c0 + (ok ? lane : 0);
    float v = a.v[r], ref = a.refrac[r], rate = a.rate[r], cnt = a.counts[r], g = a.g[r];
    const float drive = a.drive[r], p = ok ? a.pois[r] : 0.0f;
    const uint64_t seed = a.seeds[a.inst[r]];
    const uint32_t loc = (uint32_t)a.local[r];
    const uint64_t old_hist = a.ring[r];
    uint64_t mine = 0, fired_bits = 0;    // the window's bit positions, and those of the steps the cell fired
    for (int k = 0; k < K; ++k) {
        const int t = T0 + k;
        const float acc = gk(k);
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
        const uint64_t bit = 1ull << (t & (kSlots - 1));
        mine |= bit;
        if (fired) fired_bits |= bit;
    }
    if (ok) {
        a.v[r] = v; a.g[r] = g; a.refrac[r] = ref; a.rate[r] = rate; a.counts[r] = cnt;
        a.ring[r] = (old_hist & ~mine) | fired_bits;
    }
}

// One window: steps T0 .. T0 + K - 1 of every cell.  A wavefront takes a block of 64 consecutive cells (blocks in
// the order of `sched`): phase A, the chains of its cells into LDS; phase B, their LIF updates.
template <bool LONG_BUF>
__device__ __forceinline__ void packed_window(const Packed& a, int T0, int K) {
    __shared__ float gbuf[4][kWave * kMaxWindow];
    const int lane = threadIdx.x & 63;
    const int wave = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nwaves = gridDim.x * 4;
    float* const gb = gbuf[threadIdx.x >> 6];
    const Window W = window_of(a, T0, K);
    const Edges E = edges_of(a);
    for (int bi = wave; bi < a.words; bi += nwaves) {
        const int c0 = a.sched[bi] * 64;
        const int nc = min(64, a.n - c0);
        block_chains<LONG_BUF, false>(a, W, E, c0, nc, lane, [&](int i, float gk) {
            if (lane < K) gb[i * kMaxWindow + lane] = gk;
        });
        __builtin_amdgcn_wave_barrier();
        block_lif(a, T0, K, c0, nc, lane, [&](int k) { return gb[lane * kMaxWindow + k]; });
        __builtin_amdgcn_wave_barrier();
    }
}

__global__ __launch_bounds__(256) void k_packed_run(Packed a) {
    for (int T0 = a.t0; T0 < a.t0 + a.steps; T0 += a.window) {
        packed_window<false>(a, T0, min(a.window, a.t0 + a.steps - T0));
        grid_barrier(a, gridDim.x, T0);
    }
}

// One window per launch; the window's first step is *t_base + off.
__global__ __launch_bounds__(256) void k_packed_window(Packed a, const int* t_base, int off, int K) {
    packed_window<true>(a, *t_base + off, K);
}

// The synaptic chains of the long cells for the window that k_packed_window runs next, one wavefront per cell
// (long_rows, most in-edges first): edges and history words held in registers up to 2048 in-edges, streamed per
// step beyond.  g_k of long cell i goes to long_g[i * kMaxWindow + k].
__global__ __launch_bounds__(256) void k_packed_long(Packed a, const int* t_base, int off, int K) {
    const int lane = threadIdx.x & 63;
    const int wave = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const Window W = window_of(a, *t_base + off, K);
    const Edges E = edges_of(a);
    for (int i = a.n_vlong + wave; i < a.n_long; i += gridDim.x * 4) {
        const int r = a.long_rows[i];
        const int e0 = a.ptr[r], e1 = a.ptr[r + 1];
        const float g0 = a.g[r];
        float gk;
        if (e1 - e0 <= 16 * kWave)
            gk = fh_chain::g_chain_cached<16>(E, W, e0, e1, g0, lane);
        else if (e1 - e0 <= 32 * kWave)
            gk = fh_chain::g_chain_cached<32>(E, W, e0, e1, g0, lane);
        else
            gk = fh_chain::g_chain_stream(E, W, e0, e1, g0, lane);
        if (lane < K) a.long_g[(int64_t)i * kMaxWindow + lane] = gk;
    }
}

// The synaptic chains of the very long cells (the first n_vlong of long_rows, more than 2048 in-edges), one
// workgroup per cell: wavefront w holds chunks [w * NCW, (w + 1) * NCW) of the cell's in-edges in registers, and for
// each step the running sum passes from wavefront to wavefront in edge order through LDS (one barrier per segment).
// Cells longer than 4 * NCW chunks are streamed by wavefront 0.
template <int NCW>
__global__ __launch_bounds__(256) void k_packed_vlong(Packed a, const int* t_base, int off, int K) {
    __shared__ float lists[4 * (NCW * kWave + kWave)];
    __shared__ int counts[4];
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const Window W = window_of(a, *t_base + off, K);
    const Edges E = edges_of(a);
    for (int i = blockIdx.x; i < a.n_vlong; i += gridDim.x) {
        const int r = a.long_rows[i];
        const int e0 = a.ptr[r], e1 = a.ptr[r + 1];
        float gk;
        if (e1 - e0 <= 4 * NCW * kWave)
            gk = fh_chain::g_chain_compact<NCW>(E, W, e0, e1, a.g[r], lane, wv, lists, counts);
        else
            gk = wv == 0 ? fh_chain::g_chain_stream(E, W, e0, e1, a.g[r], lane) : 0.0f;
        if (wv == 0 && lane < K) a.long_g[(int64_t)i * kMaxWindow + lane] = gk;
        __syncthreads();
    }
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
    a.long_rows = nullptr; a.long_id = nullptr; a.long_g = nullptr; a.n_long = 0; a.n_vlong = 0;
}

// Workgroups of k_packed_run that are resident together: one per CU fewer than the occupancy query reports (on
// MI355X it reports 7 per CU for this kernel, yet at most 6 per CU were ever resident at once; the bounded barrier
// wait reports such a shortfall as an error instead of hanging).
static int resident_grid(int per_cu, int cus) { return (per_cu > 1 ? per_cu - 1 : per_cu) * cus; }

static bool packed_shape_ok(const Packed& a) {
    return a.n >= 0 && a.steps >= 0 && a.window >= 1 && a.window <= a.delay && a.window <= kMaxWindow &&
           2 * a.delay + a.window <= kSlots && 2 * a.window + a.delay <= kSlots && a.t0 >= 0;
}

}  // namespace

// Run `steps` LIF steps (from global step t0) of the packed graph in one cooperative launch.
// p = [ptr, idx, w, sched (ceil(n/64) blocks of 64 cells, in visiting order), v, g, refrac, rate, counts, drive,
//      pois, inst, local, seeds, ring (n spike history words: bit t mod 64 of word c = cell c fired at step t),
//      bar (2 uint32: count, generation), error (4 int32), P (12)];
// dims = [n, steps, window, delay, grid, t0, barrier bound in ms].  Requires window <= delay and
// 2 * delay + window <= 64 (so window <= 21).
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
// captured graph of windows replays at advancing steps).  p as launch_packed_run (bar and error unused) followed by
// [long_rows (n_long), long_id (n), long_g (n_long, 21) scratch]: the cells with more than 512 in-edges, most first,
// whose synaptic chains first kernels compute (the first n_vlong, more than 2048 in-edges, one workgroup per cell;
// the rest one wavefront per cell; long_id = -1 elsewhere; n_long = 0 streams them inside the window kernel instead);
// dims = [n, K, window, delay, grid, off, n_long, n_vlong].  The kernel boundaries order the window against the next.
static int packed_window_parts(void** p, const int* dims, const int* t_base, int parts, void* stream) {
    Packed a;
    fill_packed(a, p);
    a.n = dims[0]; a.steps = dims[1]; a.window = dims[2]; a.delay = dims[3];
    const int grid = dims[4], off = dims[5];
    a.n_long = dims[6];
    a.n_vlong = dims[7];
    a.t0 = 0;
    a.words = (a.n + 63) / 64;
    a.spin_ticks = 0;
    if (!packed_shape_ok(a) || a.steps > a.window || grid <= 0 || off < 0 || t_base == nullptr || a.n_long < 0 ||
        a.n_long > a.n || a.n_vlong < 0 || a.n_vlong > a.n_long)
        return FH_ERR_ARGS;
    if (a.n_long > 0) {
        a.long_rows = (const int*)p[18]; a.long_id = (const int*)p[19]; a.long_g = (float*)p[20];
        if (!a.long_rows || !a.long_id || !a.long_g) return FH_ERR_ARGS;
    }
    if (parts & ~7) return FH_ERR_ARGS;
    if (a.n == 0 || a.steps == 0) return 0;
    if ((parts & 1) && a.n_vlong > 0)
        hipLaunchKernelGGL(k_packed_vlong<kVlongChunks>, dim3(min(a.n_vlong, 65535)), dim3(256), 0,
                           (hipStream_t)stream, a, t_base, off, a.steps);
    if ((parts & 2) && a.n_long > a.n_vlong) {
        const int lgrid = min(grid, (a.n_long - a.n_vlong + 3) / 4);
        hipLaunchKernelGGL(k_packed_long, dim3(lgrid), dim3(256), 0, (hipStream_t)stream, a, t_base, off, a.steps);
    }
    if (parts & 4)
        hipLaunchKernelGGL(k_packed_window, dim3(grid), dim3(256), 0, (hipStream_t)stream, a, t_base, off, a.steps);
    FH_LAUNCH_CHECK();
}

FH_EXPORT int launch_packed_window(void** p, const int* dims, const int* t_base, void* stream) {
    return packed_window_parts(p, dims, t_base, 7, stream);
}

// Parts of launch_packed_window, same arguments: parts bit 1 = the very long cells' chains, bit 2 = the long cells'
// chains, bit 4 = the window kernel.  The two chain kernels are independent of each other, so a caller may run them on
// two streams; the window kernel reads what both wrote and must follow them.
FH_EXPORT int launch_packed_window_parts(void** p, const int* dims, const int* t_base, int parts, void* stream) {
    return packed_window_parts(p, dims, t_base, parts, stream);
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
