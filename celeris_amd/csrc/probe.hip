// probe.hip -- HBM bandwidth probes: streaming read, write and copy kernels over device
// buffers larger than the last-level cache.  They give the device's achievable streaming
// rates on the same GPU as the step's kernels (the ceiling the roofline fractions are read
// against) and the known-size launches that calibrate the HBM byte counters.  A further
// probe replays the explicit stencil's memory-access pattern without its arithmetic: the
// time the stencil's data movement alone takes in its layout and in two others.
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

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

// ---- the explicit stencil's memory-access pattern, without its arithmetic
//
// One wave per (strip, band) tile as in explicit_step (XCD-contiguous tile order): per row
// the same loads (w, hu, BY, BA of row y+2; BX and the ten Adams-Bashforth history planes of
// row y; hv from the band's transposed staging) and stores (w, D_x, three dU/dt and two
// (F*, G*) planes of row y; D_y transposed after the band), of values that only keep the
// loads live.  Layouts: 0 the stencil's (60-column strips: 64 lanes with two halo lanes per
// side); 1 the same with hv read and D_y written row-major (no transposed accesses); 2
// 64-column strips aligned to 128-byte lines, every lane storing.  Four waves per SIMD, as
// the stencil.
struct PatternPlanes {
    const float *w, *hu, *BY, *BA, *BX, *hvT, *hvR;
    const float *hist[10];
    float *out[7];
    float *DyT, *DyR, *sink;
    int nx, ny, P, PT, nstrips, nbands, stride;
};

// LDS per wave padded to 10 KB, so that a CU holds 16 of them: the stencil's own occupancy
// (4 waves per SIMD, set by its registers)
// (layout 1 keeps no D_y staging)
template <int R, int LAYOUT>
constexpr int kPatternPad = (10240 - (LAYOUT == 1 ? R + 4 : 2 * R + 4) * 64 * (int)sizeof(float)) / (int)sizeof(float);

template <int R, int LAYOUT>
__global__ __launch_bounds__(64, 4) void probe_pattern(PatternPlanes p)
{
    __shared__ float hv[R + 4][64];
    __shared__ float dy[R][64];
    __shared__ float pad[kPatternPad<R, LAYOUT>];
    const int ntiles = p.nstrips * p.nbands, per = (ntiles + 7) / 8;
    const int tile = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (tile >= ntiles) return;
    const int X = p.nx + 4, lane = threadIdx.x, strip = tile % p.nstrips, band = tile / p.nstrips;
    const int x = p.stride * strip + lane;
    const int xc = x < X ? x : X - 1;
    const bool out = (LAYOUT == 2 ? x >= 2 : (lane >= 2 && lane < 62)) && x < p.nx + 2;
    const int ys = 2 + band * R, ye = min(ys + R, p.ny + 2);
    const size_t P = (size_t)p.P;
    if (LAYOUT == 1) {
        for (int r = 0; r < ye - ys + 4; ++r) hv[r][lane] = p.hvR[(size_t)(ys - 2 + r) * P + xc];
    } else {
        const float4 *src = reinterpret_cast<const float4 *>(p.hvT + (size_t)xc * p.PT + ys - 2);
        for (int r = 0; r < ye - ys + 4; r += 4) {
            const float4 v = src[r / 4];
            if (r < R + 4) hv[r][lane] = v.x;
            if (r + 1 < R + 4) hv[r + 1][lane] = v.y;
            if (r + 2 < R + 4) hv[r + 2][lane] = v.z;
            if (r + 3 < R + 4) hv[r + 3][lane] = v.w;
        }
    }
    __syncthreads();
    auto ld = [&](const float *b, int row) { return b[(size_t)row * P + xc]; };
    float acc = 0.f;
    for (int j = 0; j < 4; ++j) acc += ld(p.w, ys - 2 + j) + ld(p.BA, ys - 2 + j) + ld(p.hu, ys - 2 + j) + ld(p.BY, ys - 2 + j);
    float nw = ld(p.w, ys + 2), nhu = ld(p.hu, ys + 2), nBY = ld(p.BY, ys + 2), nBA = ld(p.BA, ys + 2), nBX = ld(p.BX, ys);
    float nh[10];
#pragma unroll
    for (int k = 0; k < 10; ++k) nh[k] = ld(p.hist[k], ys);
    for (int y = ys; y < ye; ++y) {
        const float v = nw + nhu + nBY + nBA + nBX + hv[y + 2 - (ys - 2)][lane];
        float hs = 0.f;
#pragma unroll
        for (int k = 0; k < 10; ++k) hs += nh[k];
        const int yn = min(y + 3, p.ny + 3), yc = min(y + 1, ye - 1);
        nw = ld(p.w, yn); nhu = ld(p.hu, yn); nBY = ld(p.BY, yn); nBA = ld(p.BA, yn); nBX = ld(p.BX, yc);
#pragma unroll
        for (int k = 0; k < 10; ++k) nh[k] = ld(p.hist[k], yc);
        if (out) {
            const size_t i = (size_t)y * P + x;
            p.out[0][i] = v;
            p.out[1][LAYOUT == 2 ? i : i - 2] = hs;
#pragma unroll
            for (int k = 2; k < 7; ++k) p.out[k][i] = v + (float)k * hs;
            if (LAYOUT == 1) p.DyR[i] = v * hs;
            else dy[y - ys][lane] = v * hs;
        }
        acc += v;
    }
    if (LAYOUT != 1 && out) {
        float *dst = p.DyT + (size_t)x * p.PT + ys - 2;
        if (ye - ys == R) {
            for (int r = 0; r < R; r += 4)
                reinterpret_cast<float4 *>(dst)[r / 4] = make_float4(dy[r][lane], dy[r + 1][lane], dy[r + 2][lane], dy[r + 3][lane]);
        } else {
            for (int r = 0; r < ye - ys; ++r) dst[r] = dy[r][lane];
        }
    }
    if (acc == -1234.5f) {
        pad[lane] = acc;
        p.sink[lane] = pad[63 - lane];
    }
}

void check(hipError_t e, const char *what)
{
    if (e != hipSuccess)
        throw Error(e == hipErrorOutOfMemory ? CEL_ENOMEM : CEL_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace

// The stencil's access pattern (probe_pattern, layout 0..2) on an nx x ny grid with R-row
// bands (8, 12 or 16): `launches` timed launches after one warm-up; returns the average
// launch time in ms.  Allocates its own planes (26 of (nx + 4) rows, about 450 MB at 2048^2).
float probe_stencil_pattern(int nx, int ny, int rows, int layout, int launches)
{
    if (nx < 4 || ny < 4 || nx > 8192 || ny > 8192) throw Error(CEL_EINVAL, "grid must lie in [4, 8192] per side");
    if (rows != 8 && rows != 12 && rows != 16) throw Error(CEL_EINVAL, "band rows must be 8, 12 or 16");
    if (layout < 0 || layout > 2) throw Error(CEL_EINVAL, "layout must be 0, 1 or 2");
    if (launches < 1 || launches > 10000) throw Error(CEL_EINVAL, "launches must lie in [1, 10000]");
    PatternPlanes q{};
    q.nx = nx; q.ny = ny;
    q.P = (nx + 4 + 16 + 63) / 64 * 64;
    q.PT = (ny + 4 + 16 + 63) / 64 * 64;
    q.stride = layout == 2 ? 64 : 60;
    q.nstrips = layout == 2 ? (nx + 2 + 63) / 64 : (nx + 59) / 60;
    q.nbands = (ny + rows - 1) / rows;
    const size_t plane = (size_t)q.P * (ny + 4) * sizeof(float), tplane = (size_t)q.PT * (nx + 4) * sizeof(float);
    std::vector<void *> bufs;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float ms = 0.f;
    auto release = [&]() {
        if (st) (void)hipStreamSynchronize(st);
        for (void *b : bufs) (void)hipFree(b);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (st) (void)hipStreamDestroy(st);
    };
    auto alloc = [&](size_t bytes) {
        void *m = nullptr;
        check(hipMalloc(&m, bytes), "hipMalloc(pattern probe)");
        bufs.push_back(m);
        check(hipMemsetAsync(m, 0, bytes, st), "hipMemset(pattern probe)");
        return static_cast<float *>(m);
    };
    try {
        check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
        check(hipEventCreate(&e0), "hipEventCreate");
        check(hipEventCreate(&e1), "hipEventCreate");
        q.w = alloc(plane); q.hu = alloc(plane); q.BY = alloc(plane); q.BA = alloc(plane); q.BX = alloc(plane);
        q.hvT = alloc(tplane); q.hvR = alloc(plane);
        for (auto &h : q.hist) h = alloc(plane);
        for (auto &o : q.out) o = alloc(plane);
        q.DyT = alloc(tplane); q.DyR = alloc(plane); q.sink = alloc(64 * sizeof(float));
        const dim3 grid(8 * ((q.nstrips * q.nbands + 7) / 8)), block(64);
        auto launch = [&]() {
#define CEL_PP(RR)                                                                                    \
    do {                                                                                              \
        if (layout == 0) hipLaunchKernelGGL((probe_pattern<RR, 0>), grid, block, 0, st, q);          \
        else if (layout == 1) hipLaunchKernelGGL((probe_pattern<RR, 1>), grid, block, 0, st, q);     \
        else hipLaunchKernelGGL((probe_pattern<RR, 2>), grid, block, 0, st, q);                      \
    } while (0)
            if (rows == 8) CEL_PP(8);
            else if (rows == 12) CEL_PP(12);
            else CEL_PP(16);
#undef CEL_PP
            check(hipGetLastError(), "pattern probe launch");
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
