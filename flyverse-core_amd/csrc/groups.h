// Grouped chunked CSR product of one wavefront, shared by the product kernel and the per-frame optic kernels.
//
// Storage: chunk-padded CSR, chunk c of the matrix is entries [16 c, 16 c + 16) of pidx / pw (int4 / float4 views),
// row r owns chunks [chunk_ptr[r], chunk_ptr[r + 1]) and pads have index -1; they are skipped, so they add nothing,
// not even 0 * x.  Consecutive rows are grouped on the host so a group has at most 64 chunks and 64 rows, or is one
// row of more than 64 chunks.  Order: each chunk is a sequential fold of its products from +0 (each product rounded
// first) and the row total a sequential fold of the chunk sums from +0.
#pragma once
#include "common.h"

namespace fh_groups {

struct Grouped {
    const int* group_row;   // first row of each group, ngroups + 1
    const int* chunk_ptr;   // rows + 1
    const int4* pidx;
    const float4* pw;
    int ngroups;
};

// Sum of chunk c, the 16 gathers x(i) issued before the fold.
template <class X>
__device__ __forceinline__ float chunk_sum(const Grouped& m, int c, X x) {
    const int4* ip = m.pidx + (int64_t)c * 4;
    const float4* wp = m.pw + (int64_t)c * 4;
    int4 iv[4];
    float4 wv4[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) { iv[q] = ip[q]; wv4[q] = wp[q]; }
    float xv[16];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        xv[4 * q + 0] = x(max(iv[q].x, 0));
        xv[4 * q + 1] = x(max(iv[q].y, 0));
        xv[4 * q + 2] = x(max(iv[q].z, 0));
        xv[4 * q + 3] = x(max(iv[q].w, 0));
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
}

// Group g of m for one fly: lane j forms the sum of the group's chunk j, the sums go through the wave's 64 floats of
// LDS, and lane r folds row r's chunk sums in order; a one-row group of more than 64 chunks folds the chunk sums of
// each round of 64 in order on every lane.  finish(row, total) runs once per row (on lane 0 for a long row).
template <class X, class Finish>
__device__ __forceinline__ void group_product(const Grouped& m, int g, int lane, float* sums, X x, Finish finish) {
    const int r0 = m.group_row[g], r1 = m.group_row[g + 1];
    const int c0 = m.chunk_ptr[r0], c1 = m.chunk_ptr[r1];
    if (c1 - c0 <= kWave) {
        const float s = lane < c1 - c0 ? chunk_sum(m, c0 + lane, x) : 0.0f;
        sums[lane] = s;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (lane < r1 - r0) {
            const int row = r0 + lane;
            float total = 0.0f;
            for (int c = m.chunk_ptr[row] - c0; c < m.chunk_ptr[row + 1] - c0; ++c) total = total + sums[c];
            finish(row, total);
        }
    } else {                                       // one long row
        float total = 0.0f;
        for (int cb = c0; cb < c1; cb += kWave) {
            const float s = cb + lane < c1 ? chunk_sum(m, cb + lane, x) : 0.0f;
            const int n = min(kWave, c1 - cb);
            for (int j = 0; j < n; ++j) total = total + fh_readlane(s, j);
        }
        if (lane == 0) finish(r0, total);
    }
}

}  // namespace fh_groups
