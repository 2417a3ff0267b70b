// Plain device-to-device copy: the known-size kernel used to calibrate HBM byte counters.
#include <chrono>
#include "scene.h"

namespace nr {

namespace {
__global__ void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
    for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}
__global__ void k_fill(float4* __restrict__ dst, size_t n) {
    for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        dst[i] = make_float4(float(i & 1023), 1.0f, 2.0f, 3.0f);
}
}  // namespace

double bench_copy(size_t bytes, int reps) {
    const size_t n = bytes / sizeof(float4);
    if (n == 0 || reps <= 0) throw std::runtime_error("bench_copy: empty copy");
    DevBuf a, b;
    a.ensure(n * sizeof(float4));
    b.ensure(n * sizeof(float4));
    int cus = 0, dev = 0;
    check_hip(hipGetDevice(&dev), "hipGetDevice");
    check_hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "CU count");
    const dim3 grid(cus * 16), block(256);
    hipLaunchKernelGGL(k_fill, grid, block, 0, nullptr, a.as<float4>(), n);
    hipLaunchKernelGGL(k_copy, grid, block, 0, nullptr, a.as<float4>(), b.as<float4>(), n);  // warm-up
    check_hip(hipDeviceSynchronize(), "copy warm-up");
    hipEvent_t e0, e1;
    check_hip(hipEventCreate(&e0), "event");
    check_hip(hipEventCreate(&e1), "event");
    check_hip(hipEventRecord(e0, nullptr), "record");
    for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL(k_copy, grid, block, 0, nullptr, a.as<float4>(), b.as<float4>(), n);
    check_hip(hipEventRecord(e1, nullptr), "record");
    check_hip(hipEventSynchronize(e1), "sync");
    float ms = 0;
    check_hip(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return 2.0 * double(n * sizeof(float4)) * reps / (double(ms) * 1e-3) / 1e9;
}

}  // namespace nr
