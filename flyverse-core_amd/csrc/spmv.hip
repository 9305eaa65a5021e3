// CSR products for the optic lobe (flyverse/optic.py:161-165; the warp kernel flyverse/kernels/neural.cu:91-112).
//
// Summation order (defined by this build, reproduced bit for bit by the CPU oracle): a row's stored entries are cut
// into chunks of kChunk consecutive entries; each chunk is a sequential fp32 fold of its products from +0 (each
// product rounded first), and the row total is a sequential fold of the chunk sums from +0.  The reference's CPU
// order (torch sparse mm) is not pinned by anything it ships; agreement with it is to 1e-5 relative.
#include "common.h"
#include "groups.h"
#include "panel.h"

#include <mutex>

namespace {

constexpr int kChunk = 16;

// One wavefront per (fly, window of kWinRows consecutive rows) over the (B, cols) fly-major x.  With half weights x
// is rounded to half first (neural.cu:99).
// Each 16-lane DPP row folds one chunk: fifteen rounds of a one-lane shift within the DPP row (row_shr:1, lane 0 of
// the row taking +0) and an add, so lane j holds (((0 + c_0) + c_1) + ... + c_j) once round j is done; lanes past a
// row's end hold -0.0, which leaves every sum unchanged.  When no row of the window has more than kLaneRow entries,
// the window's chunks run as one stream, four a round, and each chunk sum is folded into its row's total in chunk
// order.  Otherwise each row of the window runs alone: its entries on the lanes, 64 a round, up to kLaneRow entries,
// one chunk per lane (folded by the lane) beyond.
constexpr int kLaneRow = 4 * kWave;
constexpr int kWinRows = 4;
constexpr int kRowShr1 = 0x111;

// the chunk fold of one 64-entry round: the sum of chunk q of the round ends in lane 16 q + 15
__device__ __forceinline__ float fold_chunks16(float c) {
    float s = 0.0f + c;
#pragma unroll
    for (int i = 1; i < kChunk; ++i) {
        const float t = __int_as_float(
            __builtin_amdgcn_update_dpp(__float_as_int(0.0f), __float_as_int(s), kRowShr1, 0xf, 0xf, false));
        s = t + c;
    }
    return s;
}

template <typename W>
__device__ __forceinline__ float csr_product(const W* __restrict__ w, float xv, int e) {
    if (sizeof(W) == 2) xv = __half2float(__float2half_rn(xv));
    return (float)w[e] * xv;
}

// The total of one row [start, end) (uniform), every lane returning it.
template <typename W>
__device__ __forceinline__ float csr_row(const int* __restrict__ idx, const W* __restrict__ w,
                                         const float* __restrict__ xb, int start, int end, int lane) {
    float total = 0.0f;
    if (end - start <= kLaneRow) {
        for (int base = start; base < end; base += kWave) {
            const int e = base + lane;
            const float c = e < end ? csr_product(w, xb[idx[e]], e) : -0.0f;
            const float s = fold_chunks16(c);
            const int nch = min(kWave / kChunk, (end - base + kChunk - 1) / kChunk);
            for (int q = 0; q < nch; ++q) total = total + fh_readlane(s, q * kChunk + kChunk - 1);
        }
    } else {
        for (int cb = start; cb < end; cb += kWave * kChunk) {
            int cs = cb + lane * kChunk, ce = min(cs + kChunk, end);
            float s = 0.0f;
            for (int e = cs; e < ce; ++e) s = s + csr_product(w, xb[idx[e]], e);
            int nch = min(kWave, (end - cb + kChunk - 1) / kChunk);
            for (int j = 0; j < nch; ++j) total = total + fh_readlane(s, j);
        }
    }
    return total;
}

template <typename W>
__global__ __launch_bounds__(256) void k_csr(const int* __restrict__ ptr, const int* __restrict__ idx,
                                             const W* __restrict__ w, const float* __restrict__ x,
                                             float* __restrict__ y, int rows, int cols, int batch) {
    const int nwin = (rows + kWinRows - 1) / kWinRows;
    const int64_t gw = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (gw >= (int64_t)nwin * batch) return;
    const int lane = threadIdx.x & 63;
    const int b = (int)(gw / nwin), r0 = (int)(gw % nwin) * kWinRows;
    const float* xb = x + (int64_t)b * cols;
    float* yb = y + (int64_t)b * rows;
    const int nr = min(kWinRows, rows - r0);
    // the window's rows [st_r, en_r), absent rows empty
    const int p0 = ptr[r0];
    const int p1 = nr > 0 ? ptr[r0 + 1] : p0;
    const int p2 = nr > 1 ? ptr[r0 + 2] : p1;
    const int p3 = nr > 2 ? ptr[r0 + 3] : p2;
    const int p4 = nr > 3 ? ptr[r0 + 4] : p3;
    const int longest = max(max(p1 - p0, p2 - p1), max(p3 - p2, p4 - p3));
    if (longest > kLaneRow) {
        const int pr[kWinRows + 1] = {p0, p1, p2, p3, p4};
#pragma unroll
        for (int r = 0; r < kWinRows; ++r) {
            if (r < nr) {
                const float total = csr_row(idx, w, xb, pr[r], pr[r + 1], lane);
                if (lane == 0) yb[r0 + r] = total;
            }
        }
        return;
    }
    // chunk stream: the rows' chunk counts and their running offsets
    const int k1 = (p1 - p0 + kChunk - 1) / kChunk;
    const int k2 = k1 + (p2 - p1 + kChunk - 1) / kChunk, k3 = k2 + (p3 - p2 + kChunk - 1) / kChunk;
    const int T = k3 + (p4 - p3 + kChunk - 1) / kChunk;
    float t0 = 0.0f, t1 = 0.0f, t2 = 0.0f, t3 = 0.0f;
    for (int g0 = 0; g0 < T; g0 += kWave / kChunk) {
        const int g = g0 + (lane >> 4);
        const int r = (g >= k1) + (g >= k2) + (g >= k3);
        const int rs = r == 0 ? p0 : r == 1 ? p1 : r == 2 ? p2 : p3;
        const int re = r == 0 ? p1 : r == 1 ? p2 : r == 2 ? p3 : p4;
        const int kr = r == 0 ? 0 : r == 1 ? k1 : r == 2 ? k2 : k3;
        const int e = rs + (g - kr) * kChunk + (lane & (kChunk - 1));
        const float c = (g < T && e < re) ? csr_product(w, xb[idx[e]], e) : -0.0f;
        const float s = fold_chunks16(c);
        const int nq = min(kWave / kChunk, T - g0);
        for (int q = 0; q < nq; ++q) {
            const int gq = g0 + q;
            const float v = fh_readlane(s, q * kChunk + kChunk - 1);
            if (gq < k1) t0 = t0 + v;
            else if (gq < k2) t1 = t1 + v;
            else if (gq < k3) t2 = t2 + v;
            else t3 = t3 + v;
        }
    }
    if (lane == 0) {
        yb[r0] = t0;
        if (nr > 1) yb[r0 + 1] = t1;
        if (nr > 2) yb[r0 + 2] = t2;
        if (nr > 3) yb[r0 + 3] = t3;
    }
}

// Binned variant for the (B, cols) layout: one lane per chunk over a precomputed chunk table, then one lane per row
// folds its chunk sums.  chunk_row is not needed: chunk c of the matrix covers entries [c0[r] + k*kChunk, ...) of
// row r where c0 = chunk_ptr.  Pass 1 writes partial[b, c]; pass 2 folds them per row in chunk order.
__global__ __launch_bounds__(256) void k_chunk_sums(const int* __restrict__ ptr, const int* __restrict__ chunk_row,
                                                    const int* __restrict__ chunk_ptr, const int* __restrict__ idx,
                                                    const float* __restrict__ w, const float* __restrict__ x,
                                                    float* __restrict__ partial, int nchunks, int cols, int batch) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nchunks * batch) return;
    int b = (int)(t / nchunks), c = (int)(t % nchunks);
    int r = chunk_row[c];
    int cs = ptr[r] + (c - chunk_ptr[r]) * kChunk, ce = min(cs + kChunk, ptr[r + 1]);
    const float* xb = x + (int64_t)b * cols;
    float s = 0.0f;
#pragma unroll 4
    for (int e = cs; e < ce; ++e) {
        float cc = w[e] * xb[idx[e]];
        s = s + cc;
    }
    partial[t] = s;
}

__global__ __launch_bounds__(256) void k_chunk_fold(const int* __restrict__ chunk_ptr, const float* __restrict__ partial,
                                                    float* __restrict__ y, int rows, int nchunks, int batch,
                                                    float scale, int accumulate) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)rows * batch) return;
    int b = (int)(t / rows), r = (int)(t % rows);
    const float* pb = partial + (int64_t)b * nchunks;
    float total = 0.0f;
    for (int c = chunk_ptr[r]; c < chunk_ptr[r + 1]; ++c) total = total + pb[c];
    if (scale != 1.0f) total = total * scale;
    if (accumulate) total = y[t] + total;
    y[t] = total;
}

// Single-pass grouped product over chunk-padded storage: every row's entries are stored padded to whole chunks
// (pad entries have idx -1 and are skipped, so they add nothing, not even 0 * x), chunk c of the matrix occupying
// entries [16 c, 16 c + 16).  Consecutive rows are grouped on the host so a group has at most 64 chunks and 64 rows;
// one wavefront per (group, fly): lane j forms the sum of the group's chunk j from two 64-byte vector loads, the sums
// go through LDS, and lane r folds row r's chunk sums in order.  A row of more than 64 chunks is a group of its own
// and its wavefront folds the chunk sums of each round of 64 in order.  Same additions, same order as k_csr.
__global__ __launch_bounds__(256) void k_spmv_groups(const int* __restrict__ group_row, const int* __restrict__ chunk_ptr,
                                                     const int4* __restrict__ pidx, const float4* __restrict__ pw,
                                                     const float* __restrict__ x, float* __restrict__ y, int ngroups,
                                                     int rows, int cols, float scale, int accumulate) {
    __shared__ float sums[4][kWave];
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int g = blockIdx.x * 4 + wv;
    if (g >= ngroups) return;
    const int b = blockIdx.y;
    const float* xb = x + (int64_t)b * cols;
    const fh_groups::Grouped m{group_row, chunk_ptr, pidx, pw, ngroups};
    fh_groups::group_product(m, g, lane, sums[wv], [&](int i) { return xb[i]; }, [&](int row, float total) {
        if (scale != 1.0f) total = total * scale;
        float* yp = y + (int64_t)b * rows + row;
        if (accumulate) total = *yp + total;
        *yp = total;
    });
}

// LDS-staged product over chunk-padded storage: the gathers of x go to LDS instead of the vector caches (a 4-byte
// gather through the vector L1 moves a whole cache line from L2).  One workgroup of 1024 threads per tile of at
// most 1024 consecutive chunks (whole rows, host-built); thread i holds chunk i of the tile in registers.  x is
// staged range by range (each range as many columns as the LDS buffer holds); within a range a thread adds, in
// entry order, the entries of its chunk whose column lies in the range, so over the ranges every entry is added
// once, in column order, which is the chunk's entry order.  Then thread i folds the chunk sums of the tile's row i
// in chunk order.  Same additions, same order as k_csr.
__global__ __launch_bounds__(1024) void k_spmv_lds(const int* __restrict__ tile_row, const int* __restrict__ chunk_ptr,
                                                  const int4* __restrict__ pidx, const float4* __restrict__ pw,
                                                  const float* __restrict__ x, float* __restrict__ y, int rows,
                                                  int cols, int range_cols, float scale, int accumulate) {
    extern __shared__ float xs[];                  // range_cols floats (reused for the chunk sums)
    const int b = blockIdx.y;
    const float* xb = x + (int64_t)b * cols;
    const int r0 = tile_row[blockIdx.x], r1 = tile_row[blockIdx.x + 1];
    const int c0 = chunk_ptr[r0], c1 = chunk_ptr[r1];
    const int t = threadIdx.x;
    const bool mine = c0 + t < c1;
    int iv[16];
    float wv[16];
    if (mine) {
        const int4* ip = pidx + (int64_t)(c0 + t) * 4;
        const float4* wp = pw + (int64_t)(c0 + t) * 4;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int4 i4 = ip[q];
            const float4 w4 = wp[q];
            iv[4 * q] = i4.x; iv[4 * q + 1] = i4.y; iv[4 * q + 2] = i4.z; iv[4 * q + 3] = i4.w;
            wv[4 * q] = w4.x; wv[4 * q + 1] = w4.y; wv[4 * q + 2] = w4.z; wv[4 * q + 3] = w4.w;
        }
    } else {
#pragma unroll
        for (int j = 0; j < 16; ++j) { iv[j] = -1; wv[j] = 0.0f; }
    }
    // the next range's values are loaded into registers while the current range is folded (at most 16 per thread)
    float xr[16];
    auto load_range = [&](int lo) {
        const int n = min(range_cols, cols - lo);
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int k = q * 1024 + t;
            xr[q] = k < n ? xb[lo + k] : 0.0f;
        }
    };
    float s = 0.0f;
    load_range(0);
    for (int lo = 0; lo < cols; lo += range_cols) {
        const int n = min(range_cols, cols - lo);
        __syncthreads();                           // the previous range is no longer read
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int k = q * 1024 + t;
            if (k < n) xs[k] = xr[q];
        }
        __syncthreads();
        if (lo + range_cols < cols) load_range(lo + range_cols);
        const unsigned un = (unsigned)n;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const unsigned off = (unsigned)(iv[j] - lo);   // in range iff off < n (pads: idx -1 never)
            if (iv[j] >= 0 && off < un) {
                const float c = wv[j] * xs[off];
                s = s + c;
            }
        }
    }
    __syncthreads();
    xs[t] = s;
    __syncthreads();
    if (t < r1 - r0) {
        const int row = r0 + t;
        float total = 0.0f;
        const int ca = chunk_ptr[row] - c0, cb = chunk_ptr[row + 1] - c0;
        int c = ca;
        for (; c + 8 <= cb; c += 8) {              // eight reads in flight, then their adds in order
            float v[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) v[q] = xs[c + q];
#pragma unroll
            for (int q = 0; q < 8; ++q) total = total + v[q];
        }
        for (; c < cb; ++c) total = total + xs[c];
        if (scale != 1.0f) total = total * scale;
        float* yp = y + (int64_t)b * rows + row;
        if (accumulate) total = *yp + total;
        *yp = total;
    }
}

// Persistent panel product (panel.h): one 1024-thread workgroup per (row block, fly), the block's chunks in
// registers, x staged in LDS one panel of columns at a time.  Same additions, same order as k_csr.
__global__ __launch_bounds__(1024) void k_spmv_panel(fh_panel::Panel m, const float* __restrict__ x,
                                                     float* __restrict__ y, int rows, float scale, int accumulate) {
    extern __shared__ float xs[];
    const int b = blockIdx.y;
    fh_panel::block_product(m, blockIdx.x, x + (int64_t)b * m.cols, xs, [&](int row, float total) {
        if (scale != 1.0f) total = total * scale;
        float* yp = y + (int64_t)b * rows + row;
        if (accumulate) total = *yp + total;
        *yp = total;
    });
}

}  // namespace

// Same contract as flyverse/kernels/neural.cu:105 (launch_csr): p = [ptr int32, idx int32, values (float or half),
// x float (B, cols), y float (B, rows)].
FH_EXPORT int launch_csr(void** p, int rows, int cols, int batch, int half, void* stream) {
    if (rows < 0 || cols < 0 || batch < 0) return FH_ERR_ARGS;
    const int64_t waves = (int64_t)((rows + kWinRows - 1) / kWinRows) * batch;
    if (waves == 0) return 0;
    const unsigned blocks = (unsigned)((waves + 3) / 4);
    if (half)
        hipLaunchKernelGGL(k_csr<__half>, dim3(blocks), dim3(256), 0, (hipStream_t)stream, (const int*)p[0],
                           (const int*)p[1], (const __half*)p[2], (const float*)p[3], (float*)p[4], rows, cols, batch);
    else
        hipLaunchKernelGGL(k_csr<float>, dim3(blocks), dim3(256), 0, (hipStream_t)stream, (const int*)p[0],
                           (const int*)p[1], (const float*)p[2], (const float*)p[3], (float*)p[4], rows, cols, batch);
    FH_LAUNCH_CHECK();
}

// Single-pass grouped product, same order as launch_csr, over chunk-padded storage: p = [group_row (ngroups+1):
// first row of each group, chunk_ptr (rows+1), pidx int32 (16 * nchunks, pad = -1), pw float (16 * nchunks),
// x (B, cols), y (B, rows)].  y = scale * (W x) (+ y if accumulate).  Groups hold at most 64 chunks and 64 rows,
// or one row of more than 64 chunks.
FH_EXPORT int launch_spmv_grouped(void** p, int rows, int cols, int ngroups, int batch, float scale, int accumulate,
                                  void* stream) {
    if (rows < 0 || cols < 0 || batch < 0 || ngroups < 0 || batch > 65535) return FH_ERR_ARGS;
    if ((int64_t)rows * batch == 0 || ngroups == 0) return 0;
    hipLaunchKernelGGL(k_spmv_groups, dim3((unsigned)((ngroups + 3) / 4), (unsigned)batch), dim3(256), 0,
                       (hipStream_t)stream, (const int*)p[0], (const int*)p[1], (const int4*)p[2], (const float4*)p[3],
                       (const float*)p[4], (float*)p[5], ngroups, rows, cols, scale, accumulate);
    FH_LAUNCH_CHECK();
}

// LDS-staged product, same order as launch_csr, over the chunk-padded storage of launch_spmv_grouped:
// p = [tile_row (ntiles+1): first row of each tile (at most 1024 chunks, at most 1024 rows), chunk_ptr (rows+1),
// pidx (16*nchunks), pw (16*nchunks), x (B, cols), y (B, rows)]; range_cols = columns of x staged at a time
// (0: as many as the LDS holds).  y = scale * W x (+ y if accumulate).
FH_EXPORT int launch_spmv_lds(void** p, int rows, int cols, int ntiles, int batch, int range_cols, float scale,
                              int accumulate, void* stream) {
    if (rows < 0 || cols < 0 || batch < 0 || ntiles < 0 || batch > 65535 || range_cols < 0) return FH_ERR_ARGS;
    if ((int64_t)rows * batch == 0 || ntiles == 0) return 0;
    int dev = 0, lds = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    e = hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    if (e != hipSuccess) return (int)e;
    int rc = min(lds / 4, 16 * 1024);                  // a thread stages at most 16 values of a range
    if (range_cols > 0) rc = min(rc, range_cols);
    rc = max(rc, 1024);                                // the chunk sums reuse the buffer
    if ((int64_t)rc * 4 > lds) return FH_ERR_SIZE;
    e = hipFuncSetAttribute((const void*)k_spmv_lds, hipFuncAttributeMaxDynamicSharedMemorySize, rc * 4);
    if (e != hipSuccess) return (int)e;
    hipLaunchKernelGGL(k_spmv_lds, dim3((unsigned)ntiles, (unsigned)batch), dim3(1024), (size_t)rc * 4,
                       (hipStream_t)stream, (const int*)p[0], (const int*)p[1], (const int4*)p[2], (const float4*)p[3],
                       (const float*)p[4], (float*)p[5], rows, cols, rc, scale, accumulate);
    FH_LAUNCH_CHECK();
}

// Panel product, same order as launch_csr, over the chunk-padded storage of launch_spmv_grouped: p = [blk_row
// (nblk+1): first row of each block (at most 3 * 1024 chunks), chunk_ptr (rows+1), pidx (16*nchunks), pw
// (16*nchunks), x (B, cols), y (B, rows), nat (nchunks: the chunk each slot of pidx / pw holds, a permutation within
// each block; null for the natural order)]; panel = columns of x staged at a time (0: as many as fit, balanced).
// y = scale * W x (+ y if accumulate).
FH_EXPORT int launch_spmv_panel(void** p, int rows, int cols, int nblk, int batch, int panel, float scale,
                                int accumulate, void* stream) {
    if (rows < 0 || cols < 0 || batch < 0 || nblk < 0 || batch > 65535 || panel < 0) return FH_ERR_ARGS;
    if ((int64_t)rows * batch == 0 || nblk == 0) return 0;
    const int pc = fh_panel::panel_columns(cols, panel);
    const int lds_floats = fh_panel::lds_floats(pc);
    hipError_t e = hipFuncSetAttribute((const void*)k_spmv_panel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       lds_floats * 4);
    if (e != hipSuccess) return (int)e;
    const fh_panel::Panel m{(const int*)p[0], (const int*)p[1], (const int4*)p[2], (const float4*)p[3], (const int*)p[6],
                              cols, pc};
    hipLaunchKernelGGL(k_spmv_panel, dim3((unsigned)nblk, (unsigned)batch), dim3(fh_panel::kThreads),
                       (size_t)lds_floats * 4, (hipStream_t)stream, m, (const float*)p[4], (float*)p[5], rows, scale,
                       accumulate);
    FH_LAUNCH_CHECK();
}

// Chunk-binned product, same order as launch_csr: p = [ptr, chunk_row (nchunks), chunk_ptr (rows+1), idx, w,
// x (B, cols), y (B, rows), partial (B, nchunks) scratch].  y = scale * (W x) (+ y if accumulate).
FH_EXPORT int launch_spmv_chunked(void** p, int rows, int cols, int nchunks, int batch, float scale, int accumulate,
                                  void* stream) {
    if (rows < 0 || cols < 0 || batch < 0 || nchunks < 0) return FH_ERR_ARGS;
    if ((int64_t)rows * batch == 0) return 0;
    int64_t n1 = (int64_t)nchunks * batch, n2 = (int64_t)rows * batch;
    if (n1)
        hipLaunchKernelGGL(k_chunk_sums, dim3((unsigned)((n1 + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                           (const int*)p[0], (const int*)p[1], (const int*)p[2], (const int*)p[3], (const float*)p[4],
                           (const float*)p[5], (float*)p[7], nchunks, cols, batch);
    hipLaunchKernelGGL(k_chunk_fold, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       (const int*)p[2], (const float*)p[7], (float*)p[6], rows, nchunks, batch, scale, accumulate);
    FH_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ allocation-free variant
// launch_spmv_chunked keeps its chunk partial sums (batch * nchunks floats) in a workspace the caller passes.  For
// callers that cannot allocate (code under HIP graph capture, tight serving loops), launch_spmv_chunked_noalloc keeps
// them in `scratch`, a file-scope device array of at most kNoAllocFloats floats: each GPU the library runs on holds
// its own copy until the process ends.  Its device address on a GPU comes from hipGetSymbolAddress, taken once per
// device, with that device current, on the entry point's first call there, and kept in g_scratch; host code passes
// only that address to kernels.  Every call on a GPU shares that GPU's array: calls that overlap on one GPU (on
// different streams) race on it.  The chunk kernel writes every partial the fold then reads, so nothing in the array
// needs clearing between calls.
constexpr int64_t kNoAllocFloats = int64_t(1) << 22;    // 4,194,304 floats (16 MiB) per GPU
constexpr int kMaxDevices = 64;

static __device__ __align__(256) float scratch[kNoAllocFloats];

namespace {
std::mutex g_scratch_lock;
float* g_scratch[kMaxDevices];          // device address of `scratch` on each GPU

// The address of `scratch` on the current device (hipGetSymbolAddress on the first call there).
int scratch_address(float** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    if (dev < 0 || dev >= kMaxDevices) return FH_ERR_ARGS;
    std::lock_guard<std::mutex> guard(g_scratch_lock);
    if (!g_scratch[dev]) {
        void* ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return (int)e;
        if (!ptr) return FH_ERR_ARGS;
        g_scratch[dev] = (float*)ptr;
    }
    *out = g_scratch[dev];
    return 0;
}
}  // namespace

// launch_spmv_chunked without the workspace: p = [ptr, chunk_row(nchunks), chunk_ptr(rows+1), idx, w, x (batch, cols),
// y (batch, rows)]; the partial sums go to the library's device scratch.  Returns FH_ERR_SIZE when batch * nchunks
// exceeds kNoAllocFloats.  The stream must belong to the current device.  Allocates nothing.
FH_EXPORT int launch_spmv_chunked_noalloc(void** p, int rows, int cols, int nchunks, int batch, float scale,
                                          int accumulate, void* stream) {
    if (!p || rows < 0 || cols < 0 || batch < 0 || nchunks < 0) return FH_ERR_ARGS;
    if ((int64_t)nchunks * batch > kNoAllocFloats) return FH_ERR_SIZE;
    if ((int64_t)rows * batch == 0) return 0;
    float* partial = nullptr;
    const int err = scratch_address(&partial);
    if (err) return err;
    void* q[8] = {p[0], p[1], p[2], p[3], p[4], p[5], p[6], partial};
    return launch_spmv_chunked(q, rows, cols, nchunks, batch, scale, accumulate, stream);
}
