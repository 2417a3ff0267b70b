// Plain device-to-device copy: the known-size kernel used to calibrate HBM byte counters.  And the
// rate of device-scope atomics on counters that share a 128-byte line or not (the wavefront's queue
// counters and ray claims are laid out from it).
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
// The shade step's append pattern: every block iteration, one thread adds to its block's counter
// (counter blockIdx % n_cnt, `stride` ints apart) and the block waits for the old value.
__global__ void k_atomics(int* __restrict__ ctr, int iters, int n_cnt, int stride, int* __restrict__ out) {
    __shared__ int base;
    int acc = 0;
    int* const c = ctr + (blockIdx.x % n_cnt) * stride;
    for (int i = 0; i < iters; ++i) {
        if (threadIdx.x == 0) base = atomicAdd(c, 256);
        __syncthreads();
        acc += base;
        __syncthreads();
    }
    out[blockIdx.x * blockDim.x + threadIdx.x] = acc;
}
}  // namespace

double bench_atomics(int n_cnt, int stride, int iters) {
    if (n_cnt < 1 || stride < 1 || iters < 1 || size_t(n_cnt) * stride > (size_t(1) << 24))
        throw std::runtime_error("bench_atomics: counters must fit 2^24 ints");
    int cus = 0, dev = 0;
    check_hip(hipGetDevice(&dev), "hipGetDevice");
    check_hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "CU count");
    const int blocks = cus * 5, threads = 256;
    DevBuf ctr, out;
    ctr.ensure(sizeof(int) * size_t(n_cnt) * stride);
    out.ensure(sizeof(int) * size_t(blocks) * threads);
    check_hip(hipMemset(ctr.as<int>(), 0, sizeof(int) * size_t(n_cnt) * stride), "memset");
    hipLaunchKernelGGL(k_atomics, dim3(blocks), dim3(threads), 0, nullptr, ctr.as<int>(), iters, n_cnt, stride,
                       out.as<int>());  // warm-up
    check_hip(hipDeviceSynchronize(), "atomics warm-up");
    hipEvent_t e0, e1;
    check_hip(hipEventCreate(&e0), "event");
    check_hip(hipEventCreate(&e1), "event");
    check_hip(hipEventRecord(e0, nullptr), "record");
    hipLaunchKernelGGL(k_atomics, dim3(blocks), dim3(threads), 0, nullptr, ctr.as<int>(), iters, n_cnt, stride,
                       out.as<int>());
    check_hip(hipEventRecord(e1, nullptr), "record");
    check_hip(hipEventSynchronize(e1), "sync");
    float ms = 0;
    check_hip(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return double(blocks) * iters / (double(ms) * 1e-3) / 1e6;
}

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
