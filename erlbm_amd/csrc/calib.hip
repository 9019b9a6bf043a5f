// calib.hip -- a plain device-to-device copy kernel of known size, used to
// calibrate HBM traffic counters (FETCH_SIZE / WRITE_SIZE) and as the
// achievable-bandwidth reference beside the collide-stream kernel.
#include <hip/hip_runtime.h>

#include <cstdlib>

#include "common.hpp"

namespace {

typedef double dbl2 __attribute__((ext_vector_type(2)));

__global__ __launch_bounds__(256) void k_hbm_copy(const dbl2 *__restrict__ src, dbl2 *__restrict__ dst, long long n)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const dbl2 v = __builtin_nontemporal_load(&src[i]);
        __builtin_nontemporal_store(v, &dst[i]);
    }
}

// One pass, U double2 per thread at stride blockDim (coalesced), no loop.
template <int U>
__global__ __launch_bounds__(256) void k_hbm_copy_unrolled(const dbl2 *__restrict__ src, dbl2 *__restrict__ dst,
                                                           long long n)
{
    const long long base = (long long)blockIdx.x * 256 * U + threadIdx.x;
    dbl2 v[U];
#pragma unroll
    for (int u = 0; u < U; u++)
        if (base + u * 256 < n) v[u] = __builtin_nontemporal_load(&src[base + u * 256]);
#pragma unroll
    for (int u = 0; u < U; u++)
        if (base + u * 256 < n) __builtin_nontemporal_store(v[u], &dst[base + u * 256]);
}

}  // namespace

extern "C" int erlbm_hbm_copy_bench(long long bytes, int iters, float *ms_per_iter)
{
    if (bytes < 16 || iters < 1 || !ms_per_iter) return ERLBM_ERR_ARG;
    const long long n = bytes / 16;
    dbl2 *a = nullptr, *b = nullptr;
    ERLBM_HIP_TRY(hipMalloc(&a, n * 16));
    if (hipMalloc(&b, n * 16) != hipSuccess) {
        (void)hipFree(a);
        erlbm::set_last_error_msg("erlbm_hbm_copy_bench: hipMalloc");
        return ERLBM_ERR_NOMEM;
    }
    int rc = ERLBM_OK;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipMemset(a, 0, n * 16);
    int dev = 0, ncu = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    // ERLBM_COPY_VARIANT: 0 grid-stride (8 blocks per CU), 1 grid-stride (32
    // blocks per CU), 2 one pass with 4 double2 per thread, 3 with 8.
    const char *ev = getenv("ERLBM_COPY_VARIANT");
    const int variant = ev ? atoi(ev) : 0;
    auto launch = [&]() {
        if (variant == 2) {
            hipLaunchKernelGGL(k_hbm_copy_unrolled<4>, dim3((unsigned)((n + 1023) / 1024)), dim3(256), 0, nullptr, a, b, n);
        } else if (variant == 3) {
            hipLaunchKernelGGL(k_hbm_copy_unrolled<8>, dim3((unsigned)((n + 2047) / 2048)), dim3(256), 0, nullptr, a, b, n);
        } else {
            const int grid = ncu * (variant == 1 ? 32 : 8);
            hipLaunchKernelGGL(k_hbm_copy, dim3(grid), dim3(256), 0, nullptr, a, b, n);
        }
    };
    launch();  // warm-up
    (void)hipEventRecord(e0, nullptr);
    for (int it = 0; it < iters; it++) launch();
    (void)hipEventRecord(e1, nullptr);
    if (hipEventSynchronize(e1) != hipSuccess) rc = ERLBM_ERR_HIP;
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    *ms_per_iter = ms / iters;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return rc;
}
