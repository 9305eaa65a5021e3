// CSR products for the optic lobe (flyverse/optic.py:161-165; the warp kernel flyverse/kernels/neural.cu:91-112).
//
// Summation order (defined by this build, reproduced bit for bit by the CPU oracle): a row's stored entries are cut
// into chunks of kChunk consecutive entries; each chunk is a sequential fp32 fold of its products from +0 (each
// product rounded first), and the row total is a sequential fold of the chunk sums from +0.  The reference's CPU
// order (torch sparse mm) is not pinned by anything it ships; agreement with it is to 1e-5 relative.
#include "common.h"

namespace {

constexpr int kChunk = 16;

// One wavefront per (fly, row) over the (B, cols) fly-major x: lane k folds chunk k of each round of 64 chunks,
// then the wave folds the chunk sums in order.  With half weights x is rounded to half first (neural.cu:99).
template <typename W>
__global__ __launch_bounds__(256) void k_csr(const int* __restrict__ ptr, const int* __restrict__ idx,
                                             const W* __restrict__ w, const float* __restrict__ x,
                                             float* __restrict__ y, int rows, int cols, int batch) {
    int64_t gw = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (gw >= (int64_t)rows * batch) return;
    int lane = threadIdx.x & 63;
    int b = (int)(gw / rows), row = (int)(gw % rows);
    const float* xb = x + (int64_t)b * cols;
    int start = ptr[row], end = ptr[row + 1];
    float total = 0.0f;
    for (int cb = start; cb < end; cb += kWave * kChunk) {
        int cs = cb + lane * kChunk, ce = min(cs + kChunk, end);
        float s = 0.0f;
        for (int e = cs; e < ce; ++e) {
            float xv = xb[idx[e]];
            if (sizeof(W) == 2) xv = __half2float(__float2half_rn(xv));
            float c = (float)w[e] * xv;
            s = s + c;
        }
        int nch = min(kWave, (end - cb + kChunk - 1) / kChunk);
        for (int j = 0; j < nch; ++j) total = total + fh_readlane(s, j);
    }
    if (lane == 0) y[gw] = total;
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
