// The bandwidth probe bench.py reports next to the rooflines: a streaming device copy, 16 bytes per lane, with
// non-temporal loads and stores (the data is touched once), over a grid-stride loop.  Not part of the brain step.
#include "common.h"

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_copy_stream(const v4f* __restrict__ x, v4f* __restrict__ y, int64_t n4) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n4; i += (int64_t)gridDim.x * 256)
        __builtin_nontemporal_store(__builtin_nontemporal_load(&x[i]), &y[i]);
}

}  // namespace

// y[0, 4 n4) = x[0, 4 n4) (floats), `grid` workgroups of 256 threads; x and y 16-byte aligned and not overlapping.
FH_EXPORT int launch_copy_stream(const void* x, void* y, int64_t n4, int grid, void* stream) {
    if (n4 < 0 || grid <= 0 || ((reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(y)) & 15)) return FH_ERR_ARGS;
    if (!n4) return 0;
    hipLaunchKernelGGL(k_copy_stream, dim3(grid), dim3(256), 0, (hipStream_t)stream, (const v4f*)x, (v4f*)y, n4);
    FH_LAUNCH_CHECK();
}
