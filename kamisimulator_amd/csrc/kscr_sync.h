// kscr_sync.h -- workgroup synchronisation shared by the kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace kscr {

// Workgroup barrier that orders LDS only: it waits for the wave's LDS
// accesses, not for its global stores and atomics (which __syncthreads waits
// for).  For kernels that never read in the kernel what they write to global
// memory.
__device__ inline void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

}  // namespace kscr
