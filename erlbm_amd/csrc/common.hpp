// common.hpp -- shared host/device helpers for the ERLBM HIP library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "../../include/erlbm.h"

namespace erlbm {

// Per-run device status words (unsigned long long each).
//   STOP_STEP : first step index that must NOT run (ULLONG_MAX = none).  Set
//               with atomicMin by a density violation at step k (k+1: the
//               violating step itself completes, as the reference's sweep
//               does) or by an energy check before step k (k).
//   REASON    : bit 0 density violation, bit 1 energy violation.
enum StatusWord { ST_STOP_STEP = 0, ST_REASON = 1, ST_WORDS = 4 };

constexpr unsigned long long NO_STOP = ~0ull;

#define ERLBM_HIP_TRY(expr)                                                       \
    do {                                                                          \
        hipError_t _e = (expr);                                                   \
        if (_e != hipSuccess) {                                                   \
            erlbm::set_last_error(#expr, _e);                                     \
            return ERLBM_ERR_HIP;                                                 \
        }                                                                         \
    } while (0)

void set_last_error(const char *what, hipError_t e);
void set_last_error_msg(const char *msg);

// Wavefront (64-lane) sum via shuffles; lane order fixed, so reproducible.
// Workgroup barrier that orders LDS only: waits for this wave's LDS
// accesses (lgkmcnt) and meets the other waves, but leaves global loads and
// stores in flight -- __syncthreads() also waits for those (its release
// fence covers global memory), which would drain a prefetch at every
// barrier.  The "memory" clobber keeps the compiler from moving memory
// accesses across it.
__device__ __forceinline__ void lds_barrier()
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
    return v;
}

}  // namespace erlbm
