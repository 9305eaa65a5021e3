// Panel product of one fly's row block, shared by the B = 1 product kernel and the fused optic substep.
//
// Storage: the chunk-padded CSR of groups.h (chunk c = entries [16 c, 16 c + 16) of pidx / pw, pads idx -1, entries
// of a row in ascending column order), cut on the host into blocks of consecutive rows of at most kChunks chunks
// per thread; within a block the chunks may sit in any slot order (the host orders them by first column, so the
// chunks of a wavefront cover nearby columns), nat giving each slot's chunk.  One 1024-thread workgroup works on
// one (block, fly) for the whole launch: thread t holds the indices and weights of the block's slots t, t + 1024,
// t + 2048 in registers.  x is staged in LDS one panel of
// columns at a time; in each pass a thread adds, in entry order, the products of its chunks' entries whose column
// lies in the panel.  The entries of a chunk ascend in column and the panels are visited in ascending order, so
// each chunk is still one sequential fold of its products from +0, in entry order.  The chunk sums then go through
// LDS and one thread per row folds them in chunk order: the same additions in the same order as the grouped
// product (groups.h) and launch_csr.
//
// The point of the panels: a 4-byte gather through the vector caches moves a whole 128-byte line from L2 into the
// L1, one line per entry; from LDS it moves 4 bytes.  Each workgroup reads x once per launch, so the HBM stream of
// the CSR itself (8 bytes per entry) is what is left.
//
// The first panel's loads are issued before the chunks' loads, so they overlap the CSR stream.  (Two half-size panel
// buffers, the next panel's loads in flight while the current one is added, measured slower at the full-scale W_rr:
// each pass is bound by the latency of its panel's loads, and half-size panels double the passes.)
#pragma once
#include "common.h"

namespace fh_panel {

constexpr int kThreads = 1024;        // threads per workgroup
constexpr int kChunks = 3;            // chunks per thread
constexpr int kMaxPanel = 40000;      // floats of x per panel (160,000 bytes of LDS)

// Floats of x per panel: `panel` if given (at most kMaxPanel), else the columns cut into as few panels as fit, of
// equal size rounded up to a multiple of 64.
inline int panel_columns(int cols, int panel) {
    if (panel > 0) return min(panel, kMaxPanel);
    const int n = max(1, (cols + kMaxPanel - 1) / kMaxPanel);
    const int pc = ((cols + n - 1) / n + 63) / 64 * 64;
    return max(64, min(pc, kMaxPanel));
}

// Floats of dynamic LDS for panels of `pc` columns: the panel, and at least the block's chunk sums.
inline int lds_floats(int pc) { return max(pc, kChunks * kThreads); }

struct Panel {
    const int* blk_row;    // first row of each block, nblk + 1
    const int* chunk_ptr;  // rows + 1
    const int4* pidx;      // the chunk-padded storage, each block's chunks in slot order
    const float4* pw;
    const int* nat;        // slot -> chunk in the natural order (where its sum goes); null: the same
    int cols, panel;       // columns of x, floats per panel
};

// Direct-to-LDS loads of src[0, n) into dst (global_load_lds: lane i of a wave writes its 16 or 4 bytes at the wave's
// LDS base + 16 i or 4 i), all in flight at once without registers: 16-byte loads when src is 16-byte aligned, the
// last n mod 4 floats by 4-byte loads of wavefront 0; 4-byte loads otherwise.
__device__ __forceinline__ void stage(const float* __restrict__ src, int n, float* dst, int wave, int lane) {
    if ((reinterpret_cast<uintptr_t>(src) & 15) == 0) {
        const int n4 = n & ~3;
        for (int base = wave * 256; base < n4; base += kThreads * 4) {
            if (base + 4 * lane < n4)
                __builtin_amdgcn_global_load_lds((const void*)(src + base + 4 * lane),
                                                 (__attribute__((address_space(3))) void*)(dst + base), 16, 0, 0);
        }
        if (wave == 0 && lane < n - n4)
            __builtin_amdgcn_global_load_lds((const void*)(src + n4 + lane),
                                             (__attribute__((address_space(3))) void*)(dst + n4), 4, 0, 0);
    } else {
        for (int base = wave * 64; base < n; base += kThreads) {
            if (base + lane < n)
                __builtin_amdgcn_global_load_lds((const void*)(src + base + lane),
                                                 (__attribute__((address_space(3))) void*)(dst + base), 4, 0, 0);
        }
    }
}

// s_waitcnt vmcnt(0) (expcnt and lgkmcnt left at their maxima) through the builtin, so the compiler's own wait
// placement knows every vector load issued so far has completed.
__device__ __forceinline__ void wait_vm_all() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// y-row results of block `blk` for the fly whose x row is xb; xs is the dynamic LDS buffer (lds_floats(panel)
// floats).  finish(row, total) runs once per row of the block, on some thread.
template <class Finish>
__device__ __forceinline__ void block_product(const Panel& m, int blk, const float* __restrict__ xb, float* xs,
                                              Finish finish) {
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
    const int r0 = m.blk_row[blk], r1 = m.blk_row[blk + 1];
    const int cb = m.chunk_ptr[r0], ce = m.chunk_ptr[r1];
    const int np = (m.cols + m.panel - 1) / m.panel;
    if (np > 0) stage(xb, min(m.panel, m.cols), xs, wave, lane);
    // the chunks' indices and weights; as unsigned numbers a chunk's columns ascend to its end (pads, -1, are the
    // largest)
    uint32_t iv[kChunks][16];
    float wv[kChunks][16];
    int nat[kChunks];                                  // where each held chunk's sum goes (-1: no chunk)
#pragma unroll
    for (int u = 0; u < kChunks; ++u) {
        const int c = cb + t + u * kThreads;
        nat[u] = c < ce ? (m.nat ? m.nat[c] : c) : -1;
        if (c < ce) {
            const int4* ip = m.pidx + (int64_t)c * 4;
            const float4* wp = m.pw + (int64_t)c * 4;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int4 i4 = ip[q];
                const float4 w4 = wp[q];
                iv[u][4 * q] = (uint32_t)i4.x; iv[u][4 * q + 1] = (uint32_t)i4.y;
                iv[u][4 * q + 2] = (uint32_t)i4.z; iv[u][4 * q + 3] = (uint32_t)i4.w;
                wv[u][4 * q] = w4.x; wv[u][4 * q + 1] = w4.y; wv[u][4 * q + 2] = w4.z; wv[u][4 * q + 3] = w4.w;
            }
        } else {
#pragma unroll
            for (int j = 0; j < 16; ++j) { iv[u][j] = ~0u; wv[u][j] = 0.0f; }
        }
    }
    float s[kChunks];
#pragma unroll
    for (int u = 0; u < kChunks; ++u) s[u] = 0.0f;
    for (int p = 0; p < np; ++p) {
        const int lo = p * m.panel;
        const int n = min(m.panel, m.cols - lo);
        const int hi = lo + n;
        const uint32_t ulo = (uint32_t)lo, uhi = (uint32_t)hi;
        if (p > 0) {
            __syncthreads();                           // panel p - 1 is no longer read
            stage(xb + lo, n, xs, wave, lane);
        }
        wait_vm_all();                                 // this wavefront's loads of panel p (and of the chunks)
        __syncthreads();                               // everyone's
        // a chunk, then a group of R entries, that no lane of the wave has in this panel is skipped (its columns
        // ascend, so its first and last tell); the reads of a group's R entries are issued before their adds (an
        // entry outside the panel reads xs[0] and is not added), so the LDS latency is paid once per R entries
        constexpr int R = 4;
        const unsigned un = (unsigned)n;
#pragma unroll
        for (int u = 0; u < kChunks; ++u) {
            if (!__builtin_amdgcn_ballot_w64(iv[u][0] < uhi && iv[u][15] >= ulo)) continue;
#pragma unroll
            for (int j0 = 0; j0 < 16; j0 += R) {
                if (!__builtin_amdgcn_ballot_w64(iv[u][j0] < uhi && iv[u][j0 + R - 1] >= ulo)) continue;
                float xv[R];
                bool ok[R];
                unsigned off[R];
#pragma unroll
                for (int j = 0; j < R; ++j) {
                    off[j] = iv[u][j0 + j] - ulo;              // in the panel iff off < n; pads never
                    ok[j] = off[j] < un;
                }
#pragma unroll
                for (int j = 0; j < R; ++j) xv[j] = xs[ok[j] ? off[j] : 0u];
#pragma unroll
                for (int j = 0; j < R; ++j) {
                    const float c = wv[u][j0 + j] * xv[j];
                    const float a = s[u] + c;
                    s[u] = ok[j] ? a : s[u];
                }
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kChunks; ++u)
        if (nat[u] >= 0) xs[nat[u] - cb] = s[u];       // chunk cb + i's sum at xs[i]
    __syncthreads();
    for (int row = r0 + t; row < r1; row += kThreads) {
        const int ca = m.chunk_ptr[row] - cb, cz = m.chunk_ptr[row + 1] - cb;
        float total = 0.0f;
        int c = ca;
        for (; c + 16 <= cz; c += 16) {                // sixteen reads in flight, then their adds in order
            float v[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) v[q] = xs[c + q];
#pragma unroll
            for (int q = 0; q < 16; ++q) total = total + v[q];
        }
        for (; c < cz; ++c) total = total + xs[c];
        finish(row, total);
    }
}

}  // namespace fh_panel
