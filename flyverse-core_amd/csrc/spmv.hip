// CSR products for the optic lobe (flyverse/optic.py:161-165; the warp kernel flyverse/kernels/neural.cu:91-112).
//
// Summation order (defined by this build, reproduced bit for bit by the CPU oracle): the columns are cut into
// kBlocks blocks of width ceil(cols / kBlocks); a row's stored entries (ascending column) in one block form a
// segment; a segment is cut into chunks of kChunk consecutive entries, each chunk a sequential fp32 fold of its
// products from +0 (each product rounded first); the segment partial is a sequential fold of its chunk sums from +0;
// the row total is a sequential fold of the partials of its non-empty segments from +0, in block order.  The blocks
// let a product be split by column block (each block's slice of x kept in one cache) with the partials combined in
// order afterwards.  The reference's CPU order (torch sparse mm) is not pinned by anything it ships; agreement with
// it is to 1e-5 relative.
#include "common.h"

namespace {

constexpr int kChunk = 16;
constexpr int kBlocks = 8;

__device__ __forceinline__ int block_width(int cols) { return max(1, (cols + kBlocks - 1) / kBlocks); }

// One wavefront per (fly, row) over the (B, cols) fly-major x: the row's segments in order; within a segment lane k
// folds chunk k of each round of 64 chunks, then the wave folds the chunk sums in order.  With half weights x is
// rounded to half first (neural.cu:99).
template <typename W>
__global__ __launch_bounds__(256) void k_csr(const int* __restrict__ ptr, const int* __restrict__ idx,
                                             const W* __restrict__ w, const float* __restrict__ x,
                                             float* __restrict__ y, int rows, int cols, int batch) {
    int64_t gw = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (gw >= (int64_t)rows * batch) return;
    int lane = threadIdx.x & 63;
    int b = (int)(gw / rows), row = (int)(gw % rows);
    const float* xb = x + (int64_t)b * cols;
    const int width = block_width(cols);
    const int start = ptr[row], end = ptr[row + 1];
    float total = 0.0f;
    int e = start;
    while (e < end) {
        // the segment [e, se): entries whose column lies in the block of idx[e] (columns ascend, so a prefix)
        const int bend = (idx[e] / width + 1) * width;
        int se = e;
        for (;;) {
            const int k = se + lane;
            const uint64_t bal = __builtin_amdgcn_ballot_w64(k < end && idx[k] < bend);
            const int n = __builtin_popcountll(bal);
            se += n;
            if (n < kWave) break;
        }
        float part = 0.0f;
        for (int cb = e; cb < se; cb += kWave * kChunk) {
            const int cs = cb + lane * kChunk, ce = min(cs + kChunk, se);
            float s = 0.0f;
            for (int q = cs; q < ce; ++q) {
                float xv = xb[idx[q]];
                if (sizeof(W) == 2) xv = __half2float(__float2half_rn(xv));
                float c = (float)w[q] * xv;
                s = s + c;
            }
            const int nch = min(kWave, (se - cb + kChunk - 1) / kChunk);
            for (int j = 0; j < nch; ++j) part = part + fh_readlane(s, j);
        }
        total = total + part;
        e = se;
    }
    if (lane == 0) y[gw] = total;
}

// Single-pass grouped product over segment-padded storage: every segment's entries are stored padded to whole chunks
// (pad entries have idx -1 and are skipped, so they add nothing, not even 0 * x), chunk c of the matrix occupying
// entries [16 c, 16 c + 16); a chunk's block is that of its first entry.  Consecutive rows are grouped on the host so
// a group has at most 64 chunks and 64 rows; one wavefront per (group, fly): lane j forms the sum of the group's
// chunk j from two 64-byte vector loads, the sums and blocks go through LDS, and lane r folds row r's chunk sums into
// segment partials and the partials into the row total, in order.  A row of more than 64 chunks is a group of its
// own and its wavefront folds the chunks of each round of 64 in order.
__global__ __launch_bounds__(256) void k_spmv_groups(const int* __restrict__ group_row, const int* __restrict__ chunk_ptr,
                                                     const int4* __restrict__ pidx, const float4* __restrict__ pw,
                                                     const float* __restrict__ x, float* __restrict__ y, int ngroups,
                                                     int rows, int cols, float scale, int accumulate) {
    __shared__ float sums[4][kWave];
    __shared__ int blks[4][kWave];
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int g = blockIdx.x * 4 + wv;
    if (g >= ngroups) return;
    const int b = blockIdx.y;
    const float* xb = x + (int64_t)b * cols;
    const int width = block_width(cols);
    const int r0 = group_row[g], r1 = group_row[g + 1];
    const int c0 = chunk_ptr[r0], c1 = chunk_ptr[r1];
    // the sum of chunk c and its block
    auto chunk_sum = [&](int c, int& blk) {
        const int4* ip = pidx + (int64_t)c * 4;
        const float4* wp = pw + (int64_t)c * 4;
        int4 iv[4];
        float4 wv4[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) { iv[q] = ip[q]; wv4[q] = wp[q]; }
        blk = iv[0].x / width;
        float xv[16];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            xv[4 * q + 0] = xb[max(iv[q].x, 0)];
            xv[4 * q + 1] = xb[max(iv[q].y, 0)];
            xv[4 * q + 2] = xb[max(iv[q].z, 0)];
            xv[4 * q + 3] = xb[max(iv[q].w, 0)];
        }
        float s = 0.0f;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            float c_;
            c_ = wv4[q].x * xv[4 * q + 0]; if (iv[q].x >= 0) s = s + c_;
            c_ = wv4[q].y * xv[4 * q + 1]; if (iv[q].y >= 0) s = s + c_;
            c_ = wv4[q].z * xv[4 * q + 2]; if (iv[q].z >= 0) s = s + c_;
            c_ = wv4[q].w * xv[4 * q + 3]; if (iv[q].w >= 0) s = s + c_;
        }
        return s;
    };
    auto finish = [&](float total, int row) {
        if (scale != 1.0f) total = total * scale;
        float* yp = y + (int64_t)b * rows + row;
        if (accumulate) total = *yp + total;
        *yp = total;
    };
    if (c1 - c0 <= kWave) {
        int blk = -1;
        const float s = lane < c1 - c0 ? chunk_sum(c0 + lane, blk) : 0.0f;
        sums[wv][lane] = s;
        blks[wv][lane] = blk;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (lane < r1 - r0) {
            const int row = r0 + lane;
            float total = 0.0f, part = 0.0f;
            int prev = -1;
            for (int c = chunk_ptr[row] - c0; c < chunk_ptr[row + 1] - c0; ++c) {
                const int bc = blks[wv][c];
                if (bc != prev) {
                    if (prev >= 0) total = total + part;
                    part = 0.0f;
                    prev = bc;
                }
                part = part + sums[wv][c];
            }
            if (prev >= 0) total = total + part;
            finish(total, row);
        }
    } else {                                       // one long row
        float total = 0.0f, part = 0.0f;
        int prev = -1;
        for (int cb = c0; cb < c1; cb += kWave) {
            int blk = -1;
            const float s = cb + lane < c1 ? chunk_sum(cb + lane, blk) : 0.0f;
            const int n = min(kWave, c1 - cb);
            for (int j = 0; j < n; ++j) {
                const int bc = __builtin_amdgcn_readlane(blk, j);
                if (bc != prev) {
                    if (prev >= 0) total = total + part;
                    part = 0.0f;
                    prev = bc;
                }
                part = part + fh_readlane(s, j);
            }
        }
        if (prev >= 0) total = total + part;
        if (lane == 0) finish(total, r0);
    }
}

}  // namespace

// Same contract as flyverse/kernels/neural.cu:105 (launch_csr): p = [ptr int32, idx int32, values (float or half),
// x float (B, cols), y float (B, rows)].
FH_EXPORT int launch_csr(void** p, int rows, int cols, int batch, int half, void* stream) {
    if (rows < 0 || cols < 0 || batch < 0) return FH_ERR_ARGS;
    int64_t waves = (int64_t)rows * batch;
    if (waves == 0) return 0;
    unsigned blocks = (unsigned)((waves + 3) / 4);
    if (half)
        hipLaunchKernelGGL(k_csr<__half>, dim3(blocks), dim3(256), 0, (hipStream_t)stream, (const int*)p[0],
                           (const int*)p[1], (const __half*)p[2], (const float*)p[3], (float*)p[4], rows, cols, batch);
    else
        hipLaunchKernelGGL(k_csr<float>, dim3(blocks), dim3(256), 0, (hipStream_t)stream, (const int*)p[0],
                           (const int*)p[1], (const float*)p[2], (const float*)p[3], (float*)p[4], rows, cols, batch);
    FH_LAUNCH_CHECK();
}

// Single-pass grouped product, same order as launch_csr, over segment-padded storage: p = [group_row (ngroups+1):
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
