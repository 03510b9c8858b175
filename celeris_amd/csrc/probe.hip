// probe.hip -- HBM bandwidth probes: streaming read, write and copy kernels over device
// buffers larger than the last-level cache.  They give the device's achievable streaming
// rates on the same GPU as the step's kernels (the ceiling the roofline fractions are read
// against) and the known-size launches that calibrate the HBM byte counters.
#include <hip/hip_runtime.h>

#include <string>

#include "engine_api.hpp"

namespace cel {

namespace {

constexpr int kProbeThreads = 256;
constexpr int kProbeUnroll = 4;                                   // float4 per thread
constexpr size_t kProbeChunk = (size_t)kProbeThreads * kProbeUnroll;   // float4 per block

// every thread reads kProbeUnroll float4 of its block's chunk (consecutive lanes on
// consecutive float4); the sum is stored only if it takes an impossible value, so the loads
// stay live without a store per thread
__global__ __launch_bounds__(kProbeThreads) void probe_read(const float4 *__restrict__ src, float *sink)
{
    const size_t base = (size_t)blockIdx.x * kProbeChunk + threadIdx.x;
    float4 v[kProbeUnroll];
#pragma unroll
    for (int j = 0; j < kProbeUnroll; ++j) v[j] = src[base + (size_t)j * kProbeThreads];
    float acc = 0.f;
#pragma unroll
    for (int j = 0; j < kProbeUnroll; ++j) acc += (v[j].x + v[j].y) + (v[j].z + v[j].w);
    if (acc == -1234.5f) sink[threadIdx.x] = acc;
}

__global__ __launch_bounds__(kProbeThreads) void probe_write(float4 *__restrict__ dst, float value)
{
    const size_t base = (size_t)blockIdx.x * kProbeChunk + threadIdx.x;
#pragma unroll
    for (int j = 0; j < kProbeUnroll; ++j) dst[base + (size_t)j * kProbeThreads] = make_float4(value, value, value, value);
}

__global__ __launch_bounds__(kProbeThreads) void probe_copy(const float4 *__restrict__ src, float4 *__restrict__ dst)
{
    const size_t base = (size_t)blockIdx.x * kProbeChunk + threadIdx.x;
    float4 v[kProbeUnroll];
#pragma unroll
    for (int j = 0; j < kProbeUnroll; ++j) v[j] = src[base + (size_t)j * kProbeThreads];
#pragma unroll
    for (int j = 0; j < kProbeUnroll; ++j) dst[base + (size_t)j * kProbeThreads] = v[j];
}

void check(hipError_t e, const char *what)
{
    if (e != hipSuccess)
        throw Error(e == hipErrorOutOfMemory ? CEL_ENOMEM : CEL_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace

// kind 0: read `bytes`, 1: write `bytes`, 2: copy `bytes` (read + write); `launches` timed
// launches after one warm-up; returns the average launch time in ms.
float probe_bandwidth(int kind, size_t bytes, int launches)
{
    if (kind < 0 || kind > 2) throw Error(CEL_EINVAL, "probe kind must be 0 (read), 1 (write) or 2 (copy)");
    if (launches < 1 || launches > 10000) throw Error(CEL_EINVAL, "launches must lie in [1, 10000]");
    const size_t chunk_bytes = kProbeChunk * sizeof(float4);
    if (bytes < chunk_bytes || bytes % chunk_bytes != 0 || bytes > ((size_t)16 << 30))
        throw Error(CEL_EINVAL, "probe size must be a multiple of 16 KiB, at most 16 GiB");
    const unsigned nblocks = (unsigned)(bytes / chunk_bytes);
    void *a = nullptr, *b = nullptr, *sink = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float ms = 0.f;
    auto release = [&]() {
        if (st) (void)hipStreamSynchronize(st);
        if (a) (void)hipFree(a);
        if (b) (void)hipFree(b);
        if (sink) (void)hipFree(sink);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (st) (void)hipStreamDestroy(st);
    };
    try {
        check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
        check(hipEventCreate(&e0), "hipEventCreate");
        check(hipEventCreate(&e1), "hipEventCreate");
        check(hipMalloc(&a, bytes), "hipMalloc(probe)");
        if (kind == 2) check(hipMalloc(&b, bytes), "hipMalloc(probe)");
        check(hipMalloc(&sink, kProbeThreads * sizeof(float)), "hipMalloc(probe)");
        check(hipMemsetAsync(a, 0, bytes, st), "hipMemset(probe)");
        auto launch = [&]() {
            if (kind == 0)
                hipLaunchKernelGGL(probe_read, dim3(nblocks), dim3(kProbeThreads), 0, st,
                                   static_cast<const float4 *>(a), static_cast<float *>(sink));
            else if (kind == 1)
                hipLaunchKernelGGL(probe_write, dim3(nblocks), dim3(kProbeThreads), 0, st, static_cast<float4 *>(a),
                                   1.0f);
            else
                hipLaunchKernelGGL(probe_copy, dim3(nblocks), dim3(kProbeThreads), 0, st,
                                   static_cast<const float4 *>(a), static_cast<float4 *>(b));
            check(hipGetLastError(), "probe launch");
        };
        launch();
        check(hipEventRecord(e0, st), "event");
        for (int i = 0; i < launches; ++i) launch();
        check(hipEventRecord(e1, st), "event");
        check(hipEventSynchronize(e1), "event sync");
        check(hipEventElapsedTime(&ms, e0, e1), "event");
    } catch (...) {
        release();
        throw;
    }
    release();
    return ms / (float)launches;
}

}  // namespace cel
