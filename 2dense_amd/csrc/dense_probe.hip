// Roofline calibration on the device the library runs on (no counterpart in
// the reference): the FP64 vector FMA rate and the HBM read and copy rates
// that the stage kernels' achieved rates are compared with.  The published
// MI355X figures are 78.6 TFLOP/s FP64 and 8 TB/s; a roofline uses the lower
// of the published and the measured value.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>

#include "../../include/dense_capi.h"

namespace {

constexpr int FMA_CHAINS = 16;  // independent FMA chains per lane: hide the FMA latency

// x <- x*a + b on FMA_CHAINS registers, iters times: 2*FMA_CHAINS*iters flops per lane.
__global__ void __launch_bounds__(256) fma_probe_kernel(double* out, int iters, double a, double b)
{
    double x[FMA_CHAINS];
#pragma unroll
    for (int k = 0; k < FMA_CHAINS; k++) x[k] = (double)(threadIdx.x + k);
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int k = 0; k < FMA_CHAINS; k++) x[k] = fma(x[k], a, b);
    }
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < FMA_CHAINS; k++) s += x[k];
    out[(size_t)blockIdx.x * 256 + threadIdx.x] = s;
}

// Streaming read of n doubles (16-byte loads), one partial sum per lane.
__global__ void __launch_bounds__(256) read_probe_kernel(const double2* __restrict__ in, size_t n2, double* out)
{
    double s = 0.0;
    const size_t stride = (size_t)gridDim.x * 256;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n2; i += stride) {
        const double2 v = in[i];
        s += v.x + v.y;
    }
    out[(size_t)blockIdx.x * 256 + threadIdx.x] = s;
}

// Streaming copy of n doubles (16-byte loads and stores).
__global__ void __launch_bounds__(256) copy_probe_kernel(const double2* __restrict__ in, double2* __restrict__ outp,
                                                         size_t n2)
{
    const size_t stride = (size_t)gridDim.x * 256;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n2; i += stride) outp[i] = in[i];
}

int grid_for_device()
{
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                                                 hipSuccess)
        return 2048;
    return std::max(cus, 1) * 8;  // 8 workgroups of 256 lanes per CU
}

// Best of reps timed launches (ms) of a launcher on stream s.
template <class F>
hipError_t best_ms(hipStream_t s, int reps, F launch, float* best)
{
    hipEvent_t e0, e1;
    hipError_t e = hipEventCreate(&e0);
    if (e != hipSuccess) return e;
    e = hipEventCreate(&e1);
    if (e != hipSuccess) {
        (void)hipEventDestroy(e0);
        return e;
    }
    *best = 1e30f;
    launch();  // warm-up
    for (int r = 0; r < reps && e == hipSuccess; r++) {
        e = hipEventRecord(e0, s);
        launch();
        if (e == hipSuccess) e = hipGetLastError();
        if (e == hipSuccess) e = hipEventRecord(e1, s);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e == hipSuccess) *best = std::min(*best, ms);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return e;
}

}  // namespace

extern "C" int dense_probe_fp64(double* tflops)
{
    if (!tflops) return DENSE_ERR_ARG;
    const int grid = grid_for_device();
    const int iters = 4096;
    double* out = nullptr;
    if (hipMalloc(&out, sizeof(double) * (size_t)grid * 256) != hipSuccess) return DENSE_ERR_NOMEM;
    hipStream_t s = nullptr;
    float ms = 0.f;
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess)
        e = best_ms(s, 5, [&] {
            hipLaunchKernelGGL(fma_probe_kernel, dim3(grid), dim3(256), 0, s, out, iters, 0.9999999, 1e-7);
        }, &ms);
    if (s) (void)hipStreamDestroy(s);
    (void)hipFree(out);
    if (e != hipSuccess) return DENSE_ERR_HIP;
    *tflops = 2.0 * FMA_CHAINS * (double)iters * grid * 256 / (ms * 1e-3) / 1e12;
    return DENSE_OK;
}

extern "C" int dense_probe_hbm(size_t bytes, double* read_gbs, double* copy_gbs)
{
    if (!read_gbs || !copy_gbs || bytes < (1u << 20)) return DENSE_ERR_ARG;
    const size_t n2 = bytes / 16;
    const int grid = grid_for_device();
    double2 *a = nullptr, *b = nullptr;
    double* out = nullptr;
    if (hipMalloc(&a, n2 * 16) != hipSuccess) return DENSE_ERR_NOMEM;
    if (hipMalloc(&b, n2 * 16) != hipSuccess || hipMalloc(&out, sizeof(double) * (size_t)grid * 256) != hipSuccess) {
        (void)hipFree(a);
        if (b) (void)hipFree(b);
        return DENSE_ERR_NOMEM;
    }
    hipStream_t s = nullptr;
    float ms_read = 0.f, ms_copy = 0.f;
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMemsetAsync(a, 0, n2 * 16, s);
    if (e == hipSuccess)
        e = best_ms(s, 5, [&] { hipLaunchKernelGGL(read_probe_kernel, dim3(grid), dim3(256), 0, s, a, n2, out); },
                    &ms_read);
    if (e == hipSuccess)
        e = best_ms(s, 5, [&] { hipLaunchKernelGGL(copy_probe_kernel, dim3(grid), dim3(256), 0, s, a, b, n2); },
                    &ms_copy);
    if (s) (void)hipStreamDestroy(s);
    (void)hipFree(a);
    (void)hipFree(b);
    (void)hipFree(out);
    if (e != hipSuccess) return DENSE_ERR_HIP;
    *read_gbs = (double)n2 * 16 / (ms_read * 1e-3) / 1e9;
    *copy_gbs = 2.0 * (double)n2 * 16 / (ms_copy * 1e-3) / 1e9;
    return DENSE_OK;
}
