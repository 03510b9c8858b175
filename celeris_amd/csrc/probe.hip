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
// the same loads (w, hu, BY, BA of row y+2; BX and the two Adams-Bashforth history slots of
// row y, each a quad plane and a G* plane; hv from the band's transposed staging) and stores
// (w, D_x and the new history slot of row y; D_y transposed after the band), of values that
// only keep the loads live.  Layouts: 0 the stencil's (60-column strips: 64 lanes with two halo lanes per
// side); 1 the same with hv read and D_y written row-major (no transposed accesses); 2
// 64-column strips aligned to 128-byte lines, every lane storing.  Four waves per SIMD, as
// the stencil.
struct PatternPlanes {
    const float *w, *hu, *BY, *BA, *BX, *hvT, *hvR;
    const float4 *hq[2];     // history slots of steps n-1, n-2: quad planes
    const float *h1[2];      // and their G* planes
    float *wn, *Dx;
    float4 *oq;              // new history slot
    float *o1;
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
    float4 nq[2];
    float n1[2];
    auto ldh = [&](int row) {
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            nq[k] = p.hq[k][(size_t)row * P + xc];
            n1[k] = ld(p.h1[k], row);
        }
    };
    ldh(ys);
    for (int y = ys; y < ye; ++y) {
        const float v = nw + nhu + nBY + nBA + nBX + hv[y + 2 - (ys - 2)][lane];
        float hs = 0.f;
#pragma unroll
        for (int k = 0; k < 2; ++k) hs += (nq[k].x + nq[k].y) + (nq[k].z + nq[k].w) + n1[k];
        const int yn = min(y + 3, p.ny + 3), yc = min(y + 1, ye - 1);
        nw = ld(p.w, yn); nhu = ld(p.hu, yn); nBY = ld(p.BY, yn); nBA = ld(p.BA, yn); nBX = ld(p.BX, yc);
        ldh(yc);
        if (out) {
            const size_t i = (size_t)y * P + x;
            p.wn[i] = v;
            p.Dx[LAYOUT == 2 ? i : i - 2] = hs;
            p.oq[i] = make_float4(v + 2.f * hs, v + 3.f * hs, v + 4.f * hs, v + 5.f * hs);
            p.o1[i] = v + 6.f * hs;
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

// ---- the line solves' structure, without their coefficient arithmetic
//
// One 256-thread block per line of 2048 unknowns (8 per thread, as the (8, 4) shape of
// solve_lines_block), lines of both directions in one launch: 16-byte loads of the right-hand
// side and the bed, the three chunk levels (one LDS exchange and barrier each), eight
// cross-thread reduction and nine substitution levels in LDS (a barrier each), the chunk
// substitution and 8-byte stores.  ELIM 0 combines equations with additions only; 1 with the
// reference's eliminations (two quotients per reduced equation, one per substituted unknown)
// on a diagonally dominant system.
struct SolvePatternPlanes {
    const float *D[2], *B[2];
    float *out[2];
    int lines, pitch;
};

struct alignas(16) PEq { float a, b, c, d; };

template <int ELIM>
__device__ __forceinline__ PEq probe_reduce(const PEq &L, const PEq &H, const PEq &R)
{
    if (ELIM == 0) return PEq{L.a + H.a, H.b + R.b, R.c + H.c, H.d + L.d + R.d};
    const float k1 = H.a / L.b, k2 = H.c / R.b;
    return PEq{-L.a * k1, H.b - L.c * k1 - R.a * k2, -R.c * k2, H.d - L.d * k1 - R.d * k2};
}

template <int ELIM>
__global__ __launch_bounds__(256) void probe_solve(SolvePatternPlanes p)
{
    constexpr int C = 8, NT = 256;
    __shared__ float ta[NT + 8], tb[NT + 8], tc[NT + 8], td[NT + 8];
    __shared__ PEq xch[2][NT];
    const int dir = blockIdx.x >= p.lines ? 1 : 0;
    const int line = (int)blockIdx.x - dir * p.lines + 2;
    const int t = threadIdx.x, base = t * C;
    const float *D = p.D[dir] + (size_t)line * p.pitch, *B = p.B[dir] + (size_t)line * p.pitch;
    float *out = p.out[dir] + (size_t)line * p.pitch + 2;
    PEq q[C];
    {
        const float4 d0 = *reinterpret_cast<const float4 *>(D + base), d1 = *reinterpret_cast<const float4 *>(D + base + 4);
        const float4 b0 = *reinterpret_cast<const float4 *>(B + base), b1 = *reinterpret_cast<const float4 *>(B + base + 4),
                     b2 = *reinterpret_cast<const float4 *>(B + base + 8);
        const float dv[C] = {d0.x, d0.y, d0.z, d0.w, d1.x, d1.y, d1.z, d1.w};
        const float bv[C + 4] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w, b2.x, b2.y, b2.z, b2.w};
#pragma unroll
        for (int i = 0; i < C; ++i) q[i] = PEq{-bv[i], 1.f + bv[i + 2], -bv[i + 4], dv[i]};
    }
#pragma unroll
    for (int k = 1; k <= 3; ++k) {
        const int s = 1 << (k - 1);
        xch[k & 1][t] = q[s - 1];
        __syncthreads();
        const PEq nr = xch[k & 1][t + 1 < NT ? t + 1 : t];
#pragma unroll
        for (int j = 0; j < (C >> k); ++j) {
            const int pp = ((j + 1) << k) - 1;
            q[pp] = probe_reduce<ELIM>(q[pp - s], q[pp], pp + s < C ? q[pp + s] : nr);
        }
    }
    auto at = [](int i) { return i + (i >> 5); };
    ta[at(t)] = q[C - 1].a; tb[at(t)] = q[C - 1].b; tc[at(t)] = q[C - 1].c; td[at(t)] = q[C - 1].d;
    __syncthreads();
#pragma unroll
    for (int j = 1; j <= 8; ++j) {
        const int s = 1 << (j - 1), i = ((t + 1) << j) - 1;
        if (i < NT) {
            const int l = at(i - s), m = at(i), r = i + s < NT ? at(i + s) : m;
            const PEq o = probe_reduce<ELIM>(PEq{ta[l], tb[l], tc[l], td[l]}, PEq{ta[m], tb[m], tc[m], td[m]},
                                             PEq{ta[r], tb[r], tc[r], td[r]});
            ta[m] = o.a; tb[m] = o.b; tc[m] = o.c; td[m] = o.d;
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 8; j >= 0; --j) {
        const int s = 1 << j, i = ((2 * t + 1) << j) - 1;
        if (i < NT) {
            const float xl = i - s >= 0 ? td[at(i - s)] : 0.f, xr = i + s < NT ? td[at(i + s)] : 0.f;
            const int m = at(i);
            td[m] = ELIM == 0 ? td[m] + ta[m] * xl + tc[m] * xr : (td[m] - ta[m] * xl - tc[m] * xr) / tb[m];
        }
        __syncthreads();
    }
    q[C - 1].d = td[at(t)];
    const float prev = t > 0 ? td[at(t - 1)] : 0.f;
#pragma unroll
    for (int k = 2; k >= 0; --k) {
        const int s = 1 << k;
#pragma unroll
        for (int qq = 0; qq < (C >> (k + 1)); ++qq) {
            const int pp = s * (2 * qq + 1) - 1;
            const float xl = pp - s >= 0 ? q[pp - s].d : prev, xr = q[pp + s].d;
            q[pp].d = ELIM == 0 ? q[pp].d + q[pp].a * xl + q[pp].c * xr : (q[pp].d - q[pp].a * xl - q[pp].c * xr) / q[pp].b;
        }
    }
#pragma unroll
    for (int i = 0; i < C; i += 2) *reinterpret_cast<float2 *>(out + base + i) = make_float2(q[i].d, q[i + 1].d);
}

void check(hipError_t e, const char *what)
{
    if (e != hipSuccess)
        throw Error(e == hipErrorOutOfMemory ? CEL_ENOMEM : CEL_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace

// The stencil's access pattern (probe_pattern, layout 0..2) on an nx x ny grid with R-row
// bands (8, 12 or 16): `launches` timed launches after one warm-up; returns the average
// launch time in ms.  Allocates its own planes (the equivalent of 26 of (nx + 4) rows,
// about 450 MB at 2048^2).
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
        for (int k = 0; k < 2; ++k) {
            q.hq[k] = reinterpret_cast<const float4 *>(alloc(4 * plane));
            q.h1[k] = alloc(plane);
        }
        q.wn = alloc(plane); q.Dx = alloc(plane);
        q.oq = reinterpret_cast<float4 *>(alloc(4 * plane)); q.o1 = alloc(plane);
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

// The line solves' structure (probe_solve) over `lines` x-lines and as many y-lines of 2048
// unknowns (lines <= 8192): `launches` timed launches after one warm-up; returns the average
// launch time in ms.  elim 0: additions only; 1: the reference's eliminations.  Allocates its
// own planes.
float probe_solve_pattern(int lines, int elim, int launches)
{
    if (lines < 1 || lines > 8192) throw Error(CEL_EINVAL, "lines must lie in [1, 8192]");
    if (elim < 0 || elim > 1) throw Error(CEL_EINVAL, "elim must be 0 or 1");
    if (launches < 1 || launches > 10000) throw Error(CEL_EINVAL, "launches must lie in [1, 10000]");
    SolvePatternPlanes q{};
    q.lines = lines;
    q.pitch = (2048 + 20 + 63) / 64 * 64;
    const size_t plane = (size_t)q.pitch * (lines + 4) * sizeof(float);
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
    try {
        check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
        check(hipEventCreate(&e0), "hipEventCreate");
        check(hipEventCreate(&e1), "hipEventCreate");
        // bed planes filled with 0.1 (b = 1.1, a = c = -0.1: diagonally dominant), the rest zero
        const std::vector<float> fill(plane / sizeof(float), 0.1f);
        for (int k = 0; k < 6; ++k) {
            void *m = nullptr;
            check(hipMalloc(&m, plane), "hipMalloc(solve probe)");
            bufs.push_back(m);
            if (k == 2 || k == 3) check(hipMemcpyAsync(m, fill.data(), plane, hipMemcpyHostToDevice, st), "upload");
            else check(hipMemsetAsync(m, 0, plane, st), "hipMemset(solve probe)");
        }
        q.D[0] = static_cast<float *>(bufs[0]); q.D[1] = static_cast<float *>(bufs[1]);
        q.B[0] = static_cast<float *>(bufs[2]); q.B[1] = static_cast<float *>(bufs[3]);
        q.out[0] = static_cast<float *>(bufs[4]); q.out[1] = static_cast<float *>(bufs[5]);
        check(hipStreamSynchronize(st), "sync");
        auto launch = [&]() {
            if (elim == 0) hipLaunchKernelGGL(probe_solve<0>, dim3(2 * lines), dim3(256), 0, st, q);
            else hipLaunchKernelGGL(probe_solve<1>, dim3(2 * lines), dim3(256), 0, st, q);
            check(hipGetLastError(), "solve probe launch");
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
