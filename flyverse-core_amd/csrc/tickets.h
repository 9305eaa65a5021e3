// Work tickets for a launch in two phases, whose phase-2 jobs read what its phase-1 jobs wrote.
//
// A workgroup takes its job from a device counter (take), not from blockIdx, so tickets go out in the order the
// workgroups start running: when a workgroup holds a phase-2 ticket, every phase-1 ticket is already held by a
// running workgroup, and phase-1 jobs wait on nothing, so the wait of a phase-2 job always ends.  The wait is also
// bounded (kSpinLimit sleeps, far beyond any phase-1 duration); past it the job sets the error word and goes on, so
// every wave exits whatever happens.  The last workgroup to leave puts the counters back to zero for the next launch
// (no memset, so the launch can sit in a captured graph).
//
// Counter words (uint32, zero before the first launch): [0] next ticket, [1] phase-1 jobs done, [2] workgroups
// finished, [3] error word (bit 0: a wait hit the spin limit; left set for the host to read and clear).
#pragma once
#include "common.h"

namespace fh_tickets {

constexpr uint32_t kSpinLimit = 1u << 22;

__device__ __forceinline__ uint32_t take(uint32_t* c) {
    __shared__ uint32_t ticket;
    if (threadIdx.x == 0) ticket = __hip_atomic_fetch_add(&c[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    return ticket;
}

// After a phase-1 job's stores: every wave makes its stores visible device-wide, then one thread counts the job.
__device__ __forceinline__ void publish(uint32_t* c) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_fetch_add(&c[1], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Before a phase-2 job's loads: wait until `need` phase-1 jobs are done.
__device__ __forceinline__ void wait(uint32_t* c, uint32_t need) {
    if (threadIdx.x == 0) {
        uint32_t spins = 0;
        while (__hip_atomic_load(&c[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < need) {
            if (++spins > kSpinLimit) {
                __hip_atomic_fetch_or(&c[3], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// End of a workgroup's job; true in the last workgroup of the launch (of `total`), after the counters are reset.
__device__ __forceinline__ bool leave(uint32_t* c, uint32_t total) {
    __shared__ uint32_t last;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t done = __hip_atomic_fetch_add(&c[2], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        last = done + 1 == total;
        if (last) {
            __hip_atomic_store(&c[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&c[1], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&c[2], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    __syncthreads();
    if (last) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return last;
}

}  // namespace fh_tickets
