// Streaming-bandwidth probe of the device's HBM (measurement only): read-only, write-only and
// copy kernels moving 16 bytes per lane, for the attainable rate the step's passes are compared
// with.
#include "axg_internal.h"

namespace axg {

constexpr int PROBE_NT = 256, PROBE_UNROLL = 4;
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));  // 16 bytes per lane

__global__ void __launch_bounds__(PROBE_NT) probe_read_kernel(const u32x4* __restrict__ a, int64_t n,
                                                             u32x4* __restrict__ sink)
{
    const int64_t base = (int64_t)blockIdx.x * PROBE_NT * PROBE_UNROLL + threadIdx.x;
    uint32_t acc = 0;
#pragma unroll
    for (int j = 0; j < PROBE_UNROLL; j++) {
        const int64_t i = base + (int64_t)j * PROBE_NT;
        if (i < n) {
            const u32x4 v = a[i];
            acc ^= v.x ^ v.y ^ v.z ^ v.w;
        }
    }
    // never true for the zero-filled buffer; keeps the loads alive
    if (acc == 0x9E3779B9u) sink[threadIdx.x] = u32x4(acc);
}

__global__ void __launch_bounds__(PROBE_NT) probe_write_kernel(u32x4* __restrict__ b, int64_t n, uint32_t val)
{
    const int64_t base = (int64_t)blockIdx.x * PROBE_NT * PROBE_UNROLL + threadIdx.x;
#pragma unroll
    for (int j = 0; j < PROBE_UNROLL; j++) {
        const int64_t i = base + (int64_t)j * PROBE_NT;
        if (i < n) b[i] = u32x4{val, val, val, (uint32_t)i};
    }
}

template <bool NT>
__global__ void __launch_bounds__(PROBE_NT) probe_copy_kernel(const u32x4* __restrict__ a, u32x4* __restrict__ b,
                                                             int64_t n)
{
    const int64_t base = (int64_t)blockIdx.x * PROBE_NT * PROBE_UNROLL + threadIdx.x;
    u32x4 v[PROBE_UNROLL];
#pragma unroll
    for (int j = 0; j < PROBE_UNROLL; j++) {
        const int64_t i = base + (int64_t)j * PROBE_NT;
        if (i < n) {
            if constexpr (NT) v[j] = __builtin_nontemporal_load(a + i);
            else v[j] = a[i];
        }
    }
#pragma unroll
    for (int j = 0; j < PROBE_UNROLL; j++) {
        const int64_t i = base + (int64_t)j * PROBE_NT;
        if (i < n) {
            if constexpr (NT) __builtin_nontemporal_store(v[j], b + i);
            else b[i] = v[j];
        }
    }
}

}  // namespace axg

using namespace axg;

extern "C" int axg_hbm_probe(int32_t device, int64_t bytes, int32_t reps, double gbs[4])
{
    if (!gbs || bytes < (1 << 20) || reps < 1) return AXG_EINVAL;
    void* a = nullptr;
    void* b = nullptr;
    void* sink = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = AXG_OK;
    try {
        AXG_HIP(hipSetDevice(device));
        const int64_t n = bytes / 16;
        AXG_HIP(hipMalloc(&a, n * 16));
        AXG_HIP(hipMalloc(&b, n * 16));
        AXG_HIP(hipMalloc(&sink, PROBE_NT * 16));
        AXG_HIP(hipMemset(a, 0, n * 16));
        AXG_HIP(hipMemset(b, 0, n * 16));
        AXG_HIP(hipEventCreate(&e0));
        AXG_HIP(hipEventCreate(&e1));
        const int64_t blocks = (n + PROBE_NT * PROBE_UNROLL - 1) / (PROBE_NT * PROBE_UNROLL);
        const dim3 grid((unsigned)blocks), blk(PROBE_NT);
        for (int kind = 0; kind < 4; kind++) {
            auto launch = [&] {
                switch (kind) {
                case 0: hipLaunchKernelGGL(probe_read_kernel, grid, blk, 0, 0, (const u32x4*)a, n, (u32x4*)sink); break;
                case 1: hipLaunchKernelGGL(probe_write_kernel, grid, blk, 0, 0, (u32x4*)b, n, 7u); break;
                case 2:
                    hipLaunchKernelGGL(probe_copy_kernel<false>, grid, blk, 0, 0, (const u32x4*)a, (u32x4*)b, n);
                    break;
                default:
                    hipLaunchKernelGGL(probe_copy_kernel<true>, grid, blk, 0, 0, (const u32x4*)a, (u32x4*)b, n);
                    break;
                }
            };
            launch();  // warm-up
            AXG_HIP(hipGetLastError());
            AXG_HIP(hipEventRecord(e0, 0));
            for (int r = 0; r < reps; r++) launch();
            AXG_HIP(hipEventRecord(e1, 0));
            AXG_HIP(hipEventSynchronize(e1));
            float ms = 0;
            AXG_HIP(hipEventElapsedTime(&ms, e0, e1));
            const double moved = (double)n * 16 * (kind >= 2 ? 2 : 1) * reps;
            gbs[kind] = moved / (ms * 1e-3) / 1e9;
        }
    } catch (const Error& e) {
        set_handleless_error(e.msg);
        rc = e.code;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (sink) (void)hipFree(sink);
    return rc;
}
