// engine.cpp -- host orchestration of the Celeris step on one MI355X.
//
// One call of timestep() issues four kernels on the engine's stream:
//   explicit_step<scheme>  (Pass1+Pass2+Pass3 fused, engine.cpp:955-1048)
//   solve_lines            (x-line and y-line tridiagonal solves + irregular-wave lines,
//                           engine.cpp:1060-1233 and 1811-2151)
//   boundary_apply x2      (four sides in the order E, N, W, S, engine.cpp:1235-1438)
// then advances the scheme state machine exactly as engine.cpp:1476-1521.
// Compiled once per working precision (real = float in namespace cel, double in cel64).
#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace CEL_NS {

namespace {

constexpr real kMinDepth = real(0.0001);   // MIN_DEPTH, compute.hlsl:1016

size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

real sponge_function(real PI, real L, real x)   // compute.hlsl:1011-1014
{
    return real(0.5) + real(0.5) * std::cos(PI * std::max(L - x, real(0)) / L);
}

real wavenumber(real omega, real d, real g)     // compute.hlsl:1080-1084
{
    return omega * omega / (g * std::sqrt(std::tanh(omega * omega * d / g)));
}

}  // namespace

int count_of_matrices(int n)
{
    int c = 1;
    while (n > 1) { n /= 2; ++c; }
    return c;
}

void ShallowWaterEngine::check(hipError_t e, const char *what) const
{
    if (e != hipSuccess) throw Error(e == hipErrorOutOfMemory ? CEL_ENOMEM : CEL_EDEVICE,
                                     std::string(what) + ": " + hipGetErrorString(e));
}

ShallowWaterEngine::ShallowWaterEngine(const cel_settings &s, const cel_boundary bnd[4], const float *bottom3,
                                       const float *state3)
    : s_(s)
{
    if (!bottom3 || !state3 || !bnd) throw Error(CEL_EINVAL, "null bottom, state or boundary array");
    if (s.nx < 4 || s.ny < 4) throw Error(CEL_EINVAL, "nx and ny must be at least 4");
    if (s.nx > CEL_MAX_LINE || s.ny > CEL_MAX_LINE) throw Error(CEL_EINVAL, "nx or ny above CEL_MAX_LINE");
    if (!(s.width > 0.f) || !(s.length > 0.f)) throw Error(CEL_EINVAL, "width and length must be positive");
    if (!(s.timestep > 0.f)) throw Error(CEL_EINVAL, "timestep must be positive");
    if (!(s.gravity > 0.f)) throw Error(CEL_EINVAL, "gravity must be positive");
    if (s.time_scheme != CEL_PREDICTOR && s.time_scheme != CEL_PREDICTOR_ADAPTIVE)
        throw Error(CEL_EINVAL, "time_scheme must be CEL_PREDICTOR or CEL_PREDICTOR_ADAPTIVE");
    if (s.correction_steps_num < 0) throw Error(CEL_EINVAL, "correction_steps_num must be >= 0");
#ifdef CEL_FP64
    if (s.state_precision != CEL_STATE_FP64) throw Error(CEL_EINVAL, "an FP64 engine needs state_precision CEL_STATE_FP64");
#else
    if (s.state_precision != CEL_STATE_FP32 && s.state_precision != CEL_STATE_COMPENSATED)
        throw Error(CEL_EINVAL, "state_precision must be CEL_STATE_FP32, CEL_STATE_COMPENSATED or CEL_STATE_FP64");
    comp_ = s.state_precision == CEL_STATE_COMPENSATED;
#endif
    if (comp_ && s.correction_steps_num > 0)
        throw Error(CEL_EINVAL, "the compensated state update does not take corrector steps");
    nx_ = s.nx; ny_ = s.ny; X_ = nx_ + 4; Y_ = ny_ + 4;
    P_ = (int)round_up((size_t)X_ + 16, 64);
    PT_ = (int)round_up((size_t)Y_ + 16, 64);
    Lx_ = count_of_matrices(nx_);
    Ly_ = count_of_matrices(ny_);
    shape_x_ = default_shape(nx_);
    shape_y_ = default_shape(ny_);
    dx_ = (real)s.width / (real)(nx_ - 1);
    dy_ = (real)s.length / (real)(ny_ - 1);
    PI_ = real(4) * std::atan(real(1));
    for (int i = 0; i < 4; ++i) {
        if (bnd[i].type < CEL_SOLID || bnd[i].type > CEL_UNIFORMTIMESERIES) throw Error(CEL_EINVAL, "bad boundary type");
        side_[i].cfg = bnd[i];
    }
    current_timestep_ = old_timestep_ = old_old_timestep_ = s.timestep;
    uniform_dt_ = s.timestep;
    check(hipGetDevice(&device_), "hipGetDevice");
    rows_ = default_band_rows(device_);
    check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    for (auto &e : ev_) check(hipEventCreate(&e), "hipEventCreate");
    allocate();
    uploadBottom(bottom3);
    for (int i = 0; i < 4; ++i) setBoundary(i, bnd[i]);
    setState(state3);
}

ShallowWaterEngine::~ShallowWaterEngine()
{
    if (stream_) (void)hipStreamSynchronize(stream_);
    for (int i = 0; i < 4; ++i) freeSideTables(i);
    if (arena_) (void)hipFree(arena_);
    if (d_stats_) (void)hipFree(d_stats_);
    if (h_stats_) (void)hipHostFree(h_stats_);
    if (aux_mem_) (void)hipFree(aux_mem_);
    clearGraphs();
    if (d_clock_) (void)hipFree(d_clock_);
    for (auto &h : h_clock_) if (h) (void)hipHostFree(h);
    for (auto &e : clock_ev_) if (e) (void)hipEventDestroy(e);
    for (auto &e : ev_) if (e) (void)hipEventDestroy(e);
    for (auto &e : pool_) (void)hipEventDestroy(e);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void ShallowWaterEngine::allocate()
{
    const size_t plane = (size_t)P_ * Y_ * sizeof(real);
    const size_t tplane = (size_t)PT_ * X_ * sizeof(real);
    // w,hu x2; BY,BX,BA; Dx; history ring (3 slots of a quad plane + a G* plane) and the
    // predictor's slot; Cw,Chu; compensated w_lo, hu_lo x2
    const int nrow = 2 + 2 + 3 + 1 + 3 * 5 + 5 + 2 + (comp_ ? 4 : 0);
    const int ncol = 2 + 1 + 1 + 1 + (comp_ ? 2 : 0);         // hvT x2; DyT; ChvT; BAT; hvT_lo x2
    arena_bytes_ = plane * nrow + tplane * ncol;
    void *mem = nullptr;
    check(hipMalloc(&mem, arena_bytes_), "hipMalloc(state)");
    arena_ = static_cast<char *>(mem);
    check(hipMemsetAsync(arena_, 0, arena_bytes_, stream_), "hipMemset(state)");
    char *q = arena_;
    auto rp = [&]() { real *r = reinterpret_cast<real *>(q); q += plane; return r; };
    auto cp = [&]() { real *r = reinterpret_cast<real *>(q); q += tplane; return r; };
    w_[0] = rp(); w_[1] = rp(); hu_[0] = rp(); hu_[1] = rp();
    BY_ = rp(); BX_ = rp(); BA_ = rp(); Dx_ = rp();
    auto rp4 = [&]() { real *r = reinterpret_cast<real *>(q); q += 4 * plane; return r; };
    for (int k = 0; k < 3; ++k) { H_[k] = rp4(); H1_[k] = rp(); }
    Hp_ = rp4(); Hp1_ = rp();
    Cw_ = rp(); Chu_ = rp();
    hvT_[0] = cp(); hvT_[1] = cp(); DyT_ = cp(); ChvT_ = cp(); BAT_ = cp();
    if (comp_) {
        w_lo_[0] = rp(); w_lo_[1] = rp(); hu_lo_[0] = rp(); hu_lo_[1] = rp();
        hvT_lo_[0] = cp(); hvT_lo_[1] = cp();
    }
    const int tiles = ((4 * (nx_ / 4) + kStatsTileX - 1) / kStatsTileX) * ((4 * (ny_ / 4) + kStatsTileY - 1) / kStatsTileY);
    stats_blocks_ = tiles;
    const size_t sbytes = ((size_t)tiles + 1) * kStatsValues * sizeof(double);
    check(hipMalloc(&mem, sbytes + 256), "hipMalloc(stats)");
    d_stats_ = static_cast<double *>(mem);
    d_flag_ = reinterpret_cast<int *>(static_cast<char *>(mem) + sbytes);   // after the partial sums
    d_status_ = d_flag_ + 1;
    check(hipMemsetAsync(d_status_, 0, sizeof(int), stream_), "clear status");
    table_bytes_ += sbytes + 256;
    check(hipHostMalloc(&mem, kStatsValues * sizeof(double), hipHostMallocDefault), "hipHostMalloc(stats)");
    h_stats_ = static_cast<double *>(mem);
}

size_t ShallowWaterEngine::estimateBytes(const cel_settings &s, const cel_boundary bnd[4])
{
    const int X = s.nx + 4, Y = s.ny + 4;
    const size_t P = round_up((size_t)X + 16, 64), PT = round_up((size_t)Y + 16, 64);
    const bool comp = s.state_precision == CEL_STATE_COMPENSATED;
    const size_t plane = P * Y * sizeof(real), tplane = PT * X * sizeof(real);
    size_t b = plane * (2 + 2 + 3 + 1 + 3 * 5 + 5 + 2 + (comp ? 4 : 0)) + tplane * (2 + 1 + 1 + 1 + (comp ? 2 : 0));
    const int tiles = ((4 * (s.nx / 4) + kStatsTileX - 1) / kStatsTileX) * ((4 * (s.ny / 4) + kStatsTileY - 1) / kStatsTileY);
    b += ((size_t)tiles + 1) * kStatsValues * sizeof(double) + 256;
    for (int side = 0; side < 4; ++side) {
        const cel_boundary &c = bnd[side];
        const size_t ncell = (size_t)c.width * ((side == CEL_WEST || side == CEL_EAST) ? Y : X);
        if (ncell == 0) continue;
        const size_t e = sizeof(real);
        if (c.type == CEL_SPONGE) b += ncell * e;
        else if (c.type == CEL_SINEWAVE) b += ncell * 5 * e;
        else if (c.type == CEL_IRREGULARWAVES) {
            const int L = count_of_matrices((side == CEL_WEST || side == CEL_EAST) ? s.nx : s.ny);
            const int nact = std::min(c.n_waves, 1 << (L - 1));
            const size_t nw1 = (size_t)std::max(nact, 1);
            b += ncell * nw1 * 4 * e + nw1 * 3 * e + ncell + ncell * 3 * e;
        }
    }
    return b;
}

void ShallowWaterEngine::uploadBottom(const float *bottom3)
{
    const size_t N = (size_t)X_ * Y_;
    hostBottom_.assign(bottom3, bottom3 + 3 * N);
    uploadHostBottom();
}

// Uploads hostBottom_ (BY, BX, BA planes), rebuilds the transposed BA and decides whether
// the mesh admits the fast variants.
void ShallowWaterEngine::uploadHostBottom()
{
    const size_t N = (size_t)X_ * Y_;
    const size_t e = sizeof(real);
    (void)e;
    const std::vector<real> &b = hostBottom_;
    real *dst[3] = {BY_, BX_, BA_};
    for (int c = 0; c < 3; ++c) stagePlane(dst[c], b.data() + c * N);
    hostBA_.assign(b.begin() + 2 * N, b.begin() + 3 * N);
    StaticParams sp;
    fillStatic(sp);
    check(hipMemsetAsync(d_flag_, 0, sizeof(int), stream_), "clear flag");
    check(launch_static_fields(sp, stream_), "static_fields");
    int unsafe = 1;
    check(hipMemcpyAsync(&unsafe, d_flag_, sizeof(int), hipMemcpyDeviceToHost, stream_), "read flag");
    check(hipStreamSynchronize(stream_), "upload sync");
    fast_mesh_ = unsafe == 0 && fastSettings();
}

void ShallowWaterEngine::stagePlane(real *dst, const real *src)
{
    const size_t e = sizeof(real);
    check(hipMemcpyAsync(Dx_, src, (size_t)X_ * Y_ * e, hipMemcpyHostToDevice, stream_), "upload plane");
    check(hipMemcpy2DAsync(dst, (size_t)P_ * e, Dx_, (size_t)X_ * e, (size_t)X_ * e, Y_, hipMemcpyDeviceToDevice,
                           stream_), "pitch plane");
}

void ShallowWaterEngine::stagePlaneT(real *dstT, const real *src)
{
    check(hipMemcpyAsync(Dx_, src, (size_t)X_ * Y_ * sizeof(real), hipMemcpyHostToDevice, stream_), "upload plane");
    check(launch_transpose(Dx_, X_, dstT, PT_, Y_, X_, stream_), "transpose plane");
}

// The explicit kernel's FAST variant needs: FP32 state, the Boussinesq coefficients (their
// sea level is the still-water level, so the kernel forms each depth once), epsilon >=
// 2^-96 (the CalcUV root argument), dx^2 and dy^2 in [2^-20, 2^20], and every numerator of
// its divisions by dx^2, dy^2, 72 and 12 in [2^-80, 2^80] or zero (checked per cell by
// static_fields).  Its results are then bit-identical to the plain kernel's.
bool ShallowWaterEngine::fastSettings() const
{
    if (sizeof(real) != 4 || !s_.is_boussinesq) return false;      // coefficient sea level == SWL
    const real m = std::max(dx_, dy_);
    const real eps = std::min((real)s_.epsilon, (m * m) * (m * m));   // as fillExplicit
    auto in = [](real v) { return v >= real(0x1p-20f) && v <= real(0x1p20f); };
    return eps >= real(0x1p-96f) && in(dx_ * dx_) && in(dy_ * dy_);
}

// createSimTextures ring/scheme init, engine.cpp:3254-3269, and resetTimestep's first
// call (engine.cpp:1714-1720).
void ShallowWaterEngine::resetScheme()
{
    cur_ = 0;
    ring_old_ = 0; ring_oo_ = 1; ring_scr_ = 2;
    time_scheme_now_ = CEL_EULER;
    timestep_count_ = 1;
    correction_steps_count_ = s_.correction_steps_num;
    total_time_ = 0.f;
    current_timestep_ = old_timestep_ = old_old_timestep_ = s_.timestep;
    for (auto &sd : side_) sd.uts_iter = 1;
}

void ShallowWaterEngine::setState(const float *state3)
{
    if (!state3) throw Error(CEL_EINVAL, "null state");
    check(hipStreamSynchronize(stream_), "sync");
    const size_t N = (size_t)X_ * Y_;
    const size_t e = sizeof(real);
    // the state in the working precision (the caller's planes themselves in FP32); both
    // ping-pong buffers receive it
    std::vector<real> conv;
    const real *st = reinterpret_cast<const real *>(state3);
    if (sizeof(real) != sizeof(float)) {
        conv.assign(state3, state3 + 3 * N);
        st = conv.data();
    }
    const size_t plane = (size_t)P_ * Y_ * e, tplane0 = (size_t)PT_ * X_ * e;
    stagePlane(w_[0], st);
    stagePlane(hu_[0], st + N);
    stagePlaneT(hvT_[0], st + 2 * N);
    check(hipMemcpyAsync(w_[1], w_[0], plane, hipMemcpyDeviceToDevice, stream_), "copy w");
    check(hipMemcpyAsync(hu_[1], hu_[0], plane, hipMemcpyDeviceToDevice, stream_), "copy hu");
    check(hipMemcpyAsync(hvT_[1], hvT_[0], tplane0, hipMemcpyDeviceToDevice, stream_), "copy hv");
    // history textures start at zero (engine.cpp:3024-3034 creates them without data)
    for (int k = 0; k < 3; ++k) {
        check(hipMemsetAsync(H_[k], 0, 4 * plane, stream_), "zero history");
        check(hipMemsetAsync(H1_[k], 0, plane, stream_), "zero history");
    }
    if (comp_) {
        const size_t tplane = (size_t)PT_ * X_ * e;
        for (int k = 0; k < 2; ++k) {
            check(hipMemsetAsync(w_lo_[k], 0, plane, stream_), "zero low words");
            check(hipMemsetAsync(hu_lo_[k], 0, plane, stream_), "zero low words");
            check(hipMemsetAsync(hvT_lo_[k], 0, tplane, stream_), "zero low words");
        }
    }
    if (aux_on_) zeroAux();
    check(hipMemsetAsync(d_status_, 0, sizeof(int), stream_), "clear status");
    check(hipStreamSynchronize(stream_), "upload sync");
    resetScheme();
}

// ---- checkpoints
//
// Layout: a header (magic, version, precision, grid, pitches, plane count, scalar count),
// the host scalars of the step state machine, then every device plane the next steps can
// read, copied whole (pitch included) in checkpointPlanes' order.  The reference's hot start
// (engine.cpp:2943-2970) keeps only (w, hu, hv) and restarts the scheme with Euler steps;
// this keeps the Adams-Bashforth ring too.

namespace {
constexpr unsigned kCheckpointMagic = 0x43454c4bu;   // "CELK"
constexpr int kCheckpointVersion = 2;   // 2: history slots as a quad plane + a G* plane
struct CheckpointHeader {
    unsigned magic;
    int version, real_bytes, comp, nx, ny, P, PT, nplanes, nscalars, correction_steps, aux;
};
}  // namespace

std::vector<std::pair<real *, size_t>> ShallowWaterEngine::checkpointPlanes() const
{
    const size_t plane = (size_t)P_ * Y_ * sizeof(real), tplane = (size_t)PT_ * X_ * sizeof(real);
    std::vector<std::pair<real *, size_t>> v;
    for (int k = 0; k < 2; ++k) { v.emplace_back(w_[k], plane); v.emplace_back(hu_[k], plane); v.emplace_back(hvT_[k], tplane); }
    if (comp_)
        for (int k = 0; k < 2; ++k) {
            v.emplace_back(w_lo_[k], plane); v.emplace_back(hu_lo_[k], plane); v.emplace_back(hvT_lo_[k], tplane);
        }
    for (int k = 0; k < 3; ++k) { v.emplace_back(H_[k], 4 * plane); v.emplace_back(H1_[k], plane); }
    if (s_.correction_steps_num > 0) {
        v.emplace_back(Hp_, 4 * plane); v.emplace_back(Hp1_, plane);
        v.emplace_back(Cw_, plane); v.emplace_back(Chu_, plane); v.emplace_back(ChvT_, tplane);
    }
    if (aux_on_) v.emplace_back(aux_mem_, 3 * plane);
    return v;
}

size_t ShallowWaterEngine::checkpoint(void *buf, size_t cap)
{
    const auto planes = checkpointPlanes();
    const double scal[] = {(double)cur_, (double)ring_old_, (double)ring_oo_, (double)ring_scr_,
                           (double)time_scheme_now_, (double)timestep_count_, (double)correction_steps_count_,
                           (double)total_time_, (double)current_timestep_, (double)old_timestep_,
                           (double)old_old_timestep_, (double)side_[0].uts_iter, (double)side_[1].uts_iter,
                           (double)side_[2].uts_iter, (double)side_[3].uts_iter, (double)uniform_dt_,
                           (double)dt_initialized_, (double)target_cfl_set_, (double)target_cfl_};
    constexpr int nscal = sizeof(scal) / sizeof(scal[0]);
    size_t bytes = sizeof(CheckpointHeader) + sizeof(scal);
    for (const auto &pl : planes) bytes += pl.second;
    if (!buf || cap < bytes) return bytes;
    sync();
    char *q = static_cast<char *>(buf);
    const CheckpointHeader h{kCheckpointMagic, kCheckpointVersion, (int)sizeof(real), comp_ ? 1 : 0, nx_, ny_, P_, PT_,
                             (int)planes.size(), nscal, s_.correction_steps_num, aux_on_ ? 1 : 0};
    std::memcpy(q, &h, sizeof(h));
    q += sizeof(h);
    std::memcpy(q, scal, sizeof(scal));
    q += sizeof(scal);
    for (const auto &pl : planes) {
        check(hipMemcpyAsync(q, pl.first, pl.second, hipMemcpyDeviceToHost, stream_), "checkpoint copy");
        q += pl.second;
    }
    sync();
    return bytes;
}

void ShallowWaterEngine::restore(const void *buf, size_t bytes)
{
    if (!buf || bytes < sizeof(CheckpointHeader)) throw Error(CEL_EINVAL, "checkpoint too short");
    CheckpointHeader h;
    std::memcpy(&h, buf, sizeof(h));
    if (h.magic != kCheckpointMagic || h.version != kCheckpointVersion) throw Error(CEL_EINVAL, "not a checkpoint");
    if (h.real_bytes != (int)sizeof(real) || h.comp != (comp_ ? 1 : 0) || h.nx != nx_ || h.ny != ny_ || h.P != P_ ||
        h.PT != PT_ || h.correction_steps != s_.correction_steps_num)
        throw Error(CEL_EINVAL, "checkpoint of an engine with another grid, precision or corrector setting");
    if ((h.aux != 0) != aux_on_)
        throw Error(CEL_EINVAL, "checkpoint visual outputs differ: set cel_engine_set_visual as when it was taken");
    const auto planes = checkpointPlanes();
    constexpr int nscal = 19;
    size_t need = sizeof(CheckpointHeader) + nscal * sizeof(double);
    for (const auto &pl : planes) need += pl.second;
    if (h.nplanes != (int)planes.size() || h.nscalars != nscal || bytes != need)
        throw Error(CEL_EINVAL, "checkpoint size does not match this engine");
    const char *q = static_cast<const char *>(buf) + sizeof(h);
    double scal[nscal];
    std::memcpy(scal, q, sizeof(scal));
    q += sizeof(scal);
    // the scalars index host arrays (buffer and ring slots, time-series rows): a damaged
    // checkpoint is rejected before anything is changed
    auto whole = [](double v, double lo, double hi) { return v == std::floor(v) && v >= lo && v <= hi; };
    const double imax = 2147483647.0;
    bool ok = whole(scal[0], 0, 1) && whole(scal[1], 0, 2) && whole(scal[2], 0, 2) && whole(scal[3], 0, 2) &&
              scal[1] != scal[2] && scal[1] != scal[3] && scal[2] != scal[3] && whole(scal[4], CEL_EULER, 3) &&
              whole(scal[5], 0, imax) && whole(scal[6], 0, imax) && std::isfinite(scal[7]) &&
              whole(scal[16], 0, 1) && whole(scal[17], 0, 1) && std::isfinite(scal[18]) && scal[15] > 0.0 &&
              std::isfinite(scal[15]);
    for (int k = 8; k <= 10; ++k) ok = ok && scal[k] > 0.0 && std::isfinite(scal[k]);
    for (int k = 0; k < 4; ++k) ok = ok && whole(scal[11 + k], 1, imax);
    if (!ok) throw Error(CEL_EINVAL, "checkpoint holds an invalid step state");
    sync();
    clearGraphs();
    for (const auto &pl : planes) {
        check(hipMemcpyAsync(pl.first, q, pl.second, hipMemcpyHostToDevice, stream_), "restore copy");
        q += pl.second;
    }
    check(hipMemsetAsync(d_status_, 0, sizeof(int), stream_), "clear status");
    sync();
    cur_ = (int)scal[0]; ring_old_ = (int)scal[1]; ring_oo_ = (int)scal[2]; ring_scr_ = (int)scal[3];
    time_scheme_now_ = (int)scal[4]; timestep_count_ = (int)scal[5]; correction_steps_count_ = (int)scal[6];
    total_time_ = (real)scal[7]; current_timestep_ = (real)scal[8]; old_timestep_ = (real)scal[9];
    old_old_timestep_ = (real)scal[10];
    for (int k = 0; k < 4; ++k) side_[k].uts_iter = (int)scal[11 + k];
    uniform_dt_ = (float)scal[15];
    dt_initialized_ = scal[16] != 0.0; target_cfl_set_ = scal[17] != 0.0; target_cfl_ = (float)scal[18];
}

int ShallowWaterEngine::numericStatus()
{
    int v = 0;
    check(hipMemcpyAsync(&v, d_status_, sizeof(int), hipMemcpyDeviceToHost, stream_), "read status");
    check(hipStreamSynchronize(stream_), "status sync");
    return v;
}

// Device planes (w, hu row-major, hv transposed) of one buffer set -> three host planes.
static void download_state(hipStream_t st, real *const w, real *const hu, real *const hvT, int X, int Y, int P, int PT,
                           std::vector<real> &out)
{
    const size_t N = (size_t)X * Y, e = sizeof(real);
    out.assign(3 * N, real(0));
    std::vector<real> t(N);
    hipError_t err = hipMemcpy2DAsync(out.data(), (size_t)X * e, w, (size_t)P * e, (size_t)X * e, Y,
                                      hipMemcpyDeviceToHost, st);
    if (err == hipSuccess)
        err = hipMemcpy2DAsync(out.data() + N, (size_t)X * e, hu, (size_t)P * e, (size_t)X * e, Y,
                               hipMemcpyDeviceToHost, st);
    if (err == hipSuccess)
        err = hipMemcpy2DAsync(t.data(), (size_t)Y * e, hvT, (size_t)PT * e, (size_t)Y * e, X, hipMemcpyDeviceToHost, st);
    if (err == hipSuccess) err = hipStreamSynchronize(st);
    if (err != hipSuccess) throw Error(CEL_EDEVICE, std::string("download state: ") + hipGetErrorString(err));
    for (int y = 0; y < Y; ++y)
        for (int x = 0; x < X; ++x) out[2 * N + (size_t)y * X + x] = t[(size_t)x * Y + y];
}

void ShallowWaterEngine::getState(float *state3)
{
    if (!state3) throw Error(CEL_EINVAL, "null state");
    std::vector<real> st;
    download_state(stream_, w_[cur_], hu_[cur_], hvT_[cur_], X_, Y_, P_, PT_, st);
    for (size_t i = 0; i < st.size(); ++i) state3[i] = (float)st[i];
}

void ShallowWaterEngine::getStateF64(double *state3)
{
    if (!state3) throw Error(CEL_EINVAL, "null state");
    std::vector<real> hi, lo;
    download_state(stream_, w_[cur_], hu_[cur_], hvT_[cur_], X_, Y_, P_, PT_, hi);
    if (comp_) download_state(stream_, w_lo_[cur_], hu_lo_[cur_], hvT_lo_[cur_], X_, Y_, P_, PT_, lo);
    for (size_t i = 0; i < hi.size(); ++i) state3[i] = (double)hi[i] + (comp_ ? (double)lo[i] : 0.0);
}

void ShallowWaterEngine::freeSideTables(int side)
{
    SideState &sd = side_[side];
    for (void *p : {(void *)sd.d_tab, (void *)sd.d_irr_out, (void *)sd.d_irr_wave, (void *)sd.d_wet})
        if (p) (void)hipFree(p);
    sd.d_tab = nullptr; sd.d_irr_out = nullptr; sd.d_irr_wave = nullptr; sd.d_wet = nullptr;
}

void ShallowWaterEngine::setBoundary(int side, const cel_boundary &b)
{
    if (side < 0 || side > 3) throw Error(CEL_EINVAL, "side out of range");
    if (b.type < CEL_SOLID || b.type > CEL_UNIFORMTIMESERIES) throw Error(CEL_EINVAL, "bad boundary type");
    if (b.width < 0) throw Error(CEL_EINVAL, "negative boundary width");
    if ((b.type == CEL_SINEWAVE) && !(b.period > 0.f)) throw Error(CEL_EINVAL, "sine-wave period must be positive");
    if (b.type == CEL_IRREGULARWAVES) {
        if (b.n_waves < 0 || b.n_waves > CEL_MAX_IRREGULAR_WAVES || (b.n_waves > 0 && !b.waves))
            throw Error(CEL_EINVAL, "irregular wave table must hold 0..1000 rows");
        for (int i = 0; i < b.n_waves; ++i)
            if (!(b.waves[4 * i + 1] > 0.f)) throw Error(CEL_EINVAL, "irregular wave period must be positive");
    }
    if (b.type == CEL_UNIFORMTIMESERIES && (b.n_uts < 2 || !b.uts))
        throw Error(CEL_EINVAL, "uniform time series needs at least two rows");
    check(hipStreamSynchronize(stream_), "sync");
    SideState &sd = side_[side];
    const cel_boundary old = sd.cfg;
    sd.cfg = b;
    try {
        validateBoundaries();
    } catch (...) {
        sd.cfg = old;
        throw;
    }
    sd.waves.assign(b.waves ? b.waves : nullptr, b.waves ? b.waves + 4 * (size_t)std::max(b.n_waves, 0) : nullptr);
    sd.uts.assign(b.uts ? b.uts : nullptr, b.uts ? b.uts + 4 * (size_t)std::max(b.n_uts, 0) : nullptr);
    sd.cfg.waves = sd.waves.empty() ? nullptr : sd.waves.data();
    sd.cfg.uts = sd.uts.empty() ? nullptr : sd.uts.data();
    sd.uts_iter = 1;
    clearGraphs();
    buildSideTables(side);
}

// The boundary kernel evaluates each boundary cell's chain of side operations from the
// pre-boundary state.  Cells whose chain passes a solid side run in the first launch; they
// must not be the mirror source of another such cell.
void ShallowWaterEngine::validateBoundaries() const
{
    int w[4];
    for (int i = 0; i < 4; ++i) w[i] = side_[i].cfg.width;
    const int bl = std::max(w[CEL_WEST], 2), br = std::max(w[CEL_EAST], 2);
    const int bb = std::max(w[CEL_SOUTH], 2), bt = std::max(w[CEL_NORTH], 2);
    if (bl + br > X_ || bb + bt > Y_) throw Error(CEL_EINVAL, "boundary strips overlap across the domain");
    auto in_strip = [&](int sd, int x, int y) {
        switch (sd) {
        case CEL_WEST: return x < w[sd];
        case CEL_EAST: return x >= X_ - w[sd];
        case CEL_SOUTH: return y < w[sd];
        default: return y >= Y_ - w[sd];
        }
    };
    const int order[4] = {CEL_EAST, CEL_NORTH, CEL_WEST, CEL_SOUTH};
    auto walk = [&](int x, int y, int &sx, int &sy) {
        bool mirrored = false;
        for (int k = 3; k >= 0; --k) {
            const int sd = order[k];
            if (!in_strip(sd, x, y)) continue;
            const int ty = side_[sd].cfg.type;
            if (ty == CEL_SOLID) {
                mirrored = true;
                if (sd == CEL_WEST) x = 4 - x;
                else if (sd == CEL_SOUTH) y = 4 - y;
                else if (sd == CEL_EAST) x = 2 * nx_ + 2 - x;
                else y = 2 * ny_ + 2 - y;
            } else if (ty != CEL_SPONGE) {
                break;
            }
        }
        sx = x; sy = y;
        return mirrored;
    };
    auto boundary_cell = [&](int x, int y) { return x < bl || x >= X_ - br || y < bb || y >= Y_ - bt; };
    for (int y = 0; y < Y_; ++y)
        for (int x = 0; x < X_; ++x) {
            if (!boundary_cell(x, y)) continue;
            int sx, sy;
            if (!walk(x, y, sx, sy)) continue;
            if (sx == x && sy == y) continue;
            if (sx < 0 || sy < 0 || sx >= X_ || sy >= Y_) continue;
            if (!boundary_cell(sx, sy)) continue;
            int tx, ty2;
            if (walk(sx, sy, tx, ty2))
                throw Error(CEL_EINVAL, "solid boundary strips this wide are not supported (mirror source is itself a mirrored cell)");
        }
}

// Per-strip-cell tables: sponge damping factors (they depend on dt), sine-wave phase
// terms, irregular-wave component terms.  Computed in FP32 with the reference formulas.
void ShallowWaterEngine::buildSideTables(int side)
{
    freeSideTables(side);
    SideState &sd = side_[side];
    const cel_boundary &b = sd.cfg;
    const int wdt = b.width;
    const bool xs = side == CEL_WEST || side == CEL_EAST;
    const int nlines = xs ? Y_ : X_;
    const int ncell = wdt * nlines;
    const real g = s_.gravity, sl = b.water_level;
    // strip cell (strip index s, line position ln) -> grid cell
    auto cell_xy = [&](int s, int ln, int &x, int &y) {
        switch (side) {
        case CEL_WEST: x = s; y = ln; break;
        case CEL_EAST: x = X_ - wdt + s; y = ln; break;
        case CEL_SOUTH: x = ln; y = s; break;
        default: x = ln; y = Y_ - wdt + s; break;
        }
    };
    auto BAat = [&](int x, int y) { return hostBA_[(size_t)y * X_ + x]; };
    if (ncell == 0) return;
    if (b.type == CEL_SPONGE) {
        std::vector<real> gam(ncell);
        const real del = xs ? dx_ : dy_;
        const real Lw = (real)wdt * del;
        for (int s = 0; s < wdt; ++s)
            for (int ln = 0; ln < nlines; ++ln) {
                int x, y;
                cell_xy(s, ln, x, y);
                const real d_here = std::max(kMinDepth, sl - BAat(x, y));
                real dist;
                switch (side) {
                case CEL_WEST: dist = (real)x * del; break;
                case CEL_SOUTH: dist = (real)y * del; break;
                case CEL_EAST: dist = (real)(X_ - x) * del; break;
                default: dist = (real)(Y_ - y) * del; break;
                }
                gam[(size_t)s * nlines + ln] = sponge_function(PI_, Lw, dist) /
                    sponge_function(PI_, Lw, dist + std::sqrt(g * d_here) * current_timestep_);
            }
        check(hipMalloc(&sd.d_tab, gam.size() * sizeof(real)), "hipMalloc(sponge)");
        check(hipMemcpy(sd.d_tab, gam.data(), gam.size() * sizeof(real), hipMemcpyHostToDevice), "upload sponge");
        table_bytes_ += gam.size() * sizeof(real);
    } else if (b.type == CEL_SINEWAVE) {
        std::vector<real> tab((size_t)ncell * 5, real(0));
        const real omega = 2 * PI_ / b.period;
        sd.sine_omega = omega;
        for (int s = 0; s < wdt; ++s)
            for (int ln = 0; ln < nlines; ++ln) {
                int x, y;
                cell_xy(s, ln, x, y);
                const real d_here = std::max(real(0), sl - BAat(x, y));
                real px, py;
                switch (side) {
                case CEL_EAST: px = (real)(x - X_) * dx_; py = (real)y * dy_; break;
                case CEL_NORTH: px = (real)x * dx_; py = (real)(y - Y_) * dy_; break;
                default: px = (real)x * dx_; py = (real)y * dy_; break;
                }
                real *t = &tab[((size_t)s * nlines + ln) * 5];
                if (d_here > kMinDepth) {
                    const real k = wavenumber(omega, d_here, g);
                    const real c = omega / k;
                    t[0] = std::cos((real)b.theta) * px * k;
                    t[1] = std::sin((real)b.theta) * py * k;
                    t[2] = c * std::cos((real)b.theta);
                    t[3] = c * std::sin((real)b.theta);
                    t[4] = 1.f;
                }
            }
        check(hipMalloc(&sd.d_tab, tab.size() * sizeof(real)), "hipMalloc(sine)");
        check(hipMemcpy(sd.d_tab, tab.data(), tab.size() * sizeof(real), hipMemcpyHostToDevice), "upload sine");
        table_bytes_ += tab.size() * sizeof(real);
    } else if (b.type == CEL_IRREGULARWAVES) {
        const int L = xs ? Lx_ : Ly_;
        const int pn = 1 << (L - 1);
        const int nact = std::min(b.n_waves, pn);
        sd.irr_nw = nact;
        std::vector<real> wave((size_t)std::max(nact, 1) * 3, real(0));
        for (int i = 0; i < nact; ++i) {
            wave[i] = b.waves[4 * i];
            wave[std::max(nact, 1) + i] = 2 * PI_ / b.waves[4 * i + 1];
            wave[2 * std::max(nact, 1) + i] = b.waves[4 * i + 3];
        }
        std::vector<real> tab((size_t)ncell * std::max(nact, 1) * 4, real(0));
        std::vector<unsigned char> wet(ncell, 0);
        for (int s = 0; s < wdt; ++s)
            for (int ln = 0; ln < nlines; ++ln) {
                const size_t cell = (size_t)s * nlines + ln;
                if (ln < 2 || ln >= nlines - 2) continue;
                int col_or_row = xs ? (side == CEL_WEST ? s : X_ - wdt + s) : (side == CEL_SOUTH ? s : Y_ - wdt + s);
                real px, py, Bh;
                if (xs) {
                    px = side == CEL_WEST ? (real)col_or_row * dx_ : (real)(col_or_row - X_) * dx_;
                    py = (real)ln * dy_;
                    Bh = BAat(col_or_row, ln);
                } else {
                    px = (real)ln * dx_;
                    py = side == CEL_SOUTH ? (real)col_or_row * dy_ : (real)(col_or_row - Y_) * dy_;
                    Bh = BAat(ln, col_or_row);
                }
                const real d_here = std::max(real(0), sl - Bh);
                wet[cell] = d_here > kMinDepth;
                if (!wet[cell]) continue;
                for (int i = 0; i < nact; ++i) {
                    const real theta = b.waves[4 * i + 2];
                    const real omega = 2 * PI_ / b.waves[4 * i + 1];
                    const real k = wavenumber(omega, d_here, g);
                    const real c = omega / k;
                    real *t = &tab[(cell * nact + i) * 4];
                    t[0] = std::cos(theta) * px * k;
                    t[1] = std::sin(theta) * py * k;
                    t[2] = c * std::cos(theta);
                    t[3] = c * std::sin(theta);
                }
            }
        check(hipMalloc(&sd.d_tab, tab.size() * sizeof(real)), "hipMalloc(irregular)");
        check(hipMemcpy(sd.d_tab, tab.data(), tab.size() * sizeof(real), hipMemcpyHostToDevice), "upload irregular");
        check(hipMalloc(&sd.d_irr_wave, wave.size() * sizeof(real)), "hipMalloc(irregular waves)");
        check(hipMemcpy(sd.d_irr_wave, wave.data(), wave.size() * sizeof(real), hipMemcpyHostToDevice), "upload waves");
        check(hipMalloc(&sd.d_wet, wet.size()), "hipMalloc(wet)");
        check(hipMemcpy(sd.d_wet, wet.data(), wet.size(), hipMemcpyHostToDevice), "upload wet");
        check(hipMalloc(&sd.d_irr_out, (size_t)ncell * 3 * sizeof(real)), "hipMalloc(irregular out)");
        table_bytes_ += tab.size() * sizeof(real) + wave.size() * sizeof(real) + wet.size() + (size_t)ncell * 3 * sizeof(real);
    }
}

void ShallowWaterEngine::resetTimestep(float dt)
{
    if (!(dt > 0.f)) throw Error(CEL_EINVAL, "dt must be positive");
    uniform_dt_ = dt;
    if (dt != current_timestep_) {
        check(hipStreamSynchronize(stream_), "sync");
        current_timestep_ = dt;
        clearGraphs();
        for (int i = 0; i < 4; ++i)
            if (side_[i].cfg.type == CEL_SPONGE) buildSideTables(i);
    }
}

// timeSeriesHelper, engine.cpp:1893-1912
void ShallowWaterEngine::timeSeries(int side, real t, real &eta, real &hu, real &hv)
{
    SideState &sd = side_[side];
    const int n = (int)sd.uts.size() / 4;
    int &it = sd.uts_iter;
    eta = 0; hu = 0; hv = 0;
    while (it < n && (real)sd.uts[4 * it] < t) ++it;
    if (it >= n) return;
    const float *a = &sd.uts[4 * (it - 1)], *b = &sd.uts[4 * it];
    const real m = (t - (real)a[0]) / ((real)b[0] - (real)a[0]);
    eta = (real)a[1] + ((real)b[1] - (real)a[1]) * m;
    hu = (real)a[2] + ((real)b[2] - (real)a[2]) * m;
    hv = (real)a[3] + ((real)b[3] - (real)a[3]) * m;
}

void ShallowWaterEngine::fillExplicit(ExplicitParams &p) const
{
    std::memset(&p, 0, sizeof(p));
    p.nx = nx_; p.ny = ny_; p.P = P_; p.PT = PT_;
    p.nstrips = (nx_ + kStripCols - 1) / kStripCols;
    p.nbands = (ny_ + rows_ - 1) / rows_;
    const int c = cur_, n = 1 - cur_;
    p.w = w_[c]; p.hu = hu_[c]; p.hvT = hvT_[c];
    p.BY = BY_; p.BX = BX_; p.BA = BA_;
    p.Ho = H_[ring_old_]; p.Ho1 = H1_[ring_old_];
    p.Hoo = H_[ring_oo_]; p.Hoo1 = H1_[ring_oo_];
    p.Hs = H_[ring_scr_]; p.Hs1 = H1_[ring_scr_];
    p.Hp = Hp_; p.Hp1 = Hp1_;
    p.Cw = Cw_; p.Chu = Chu_; p.ChvT = ChvT_;
    p.w_new = w_[n]; p.Dx = Dx_; p.DyT = DyT_;
    p.w_lo = w_lo_[c]; p.w_lo_new = w_lo_[n];
    p.status = d_status_;
    // fillConstantBuffers, engine.cpp:4099-4136
    const real g = s_.gravity;
    p.two_theta = real(2) * (real)s_.theta;
    p.g = g;
    p.half_g = real(0.5) * g;
    p.Bcoef_g = (real(1) / real(15)) * g;
    p.g_over_dx = g / dx_;
    p.g_over_dy = g / dy_;
    p.one_over_dx = real(1) / dx_;
    p.one_over_dy = real(1) / dy_;
    p.dt = current_timestep_;
    const real m = std::max(dx_, dy_);
    p.epsilon = std::min((real)s_.epsilon, (m * m) * (m * m));   // CalcEpsilon, engine.cpp:290-299
    p.friction = s_.friction;
    p.sea_level = s_.still_water_elevation;
    p.is_manning = s_.is_manning;
    // AdamsBashforthCoef and SecondOrderDerivatives weights (compute.hlsl:560-605)
    const real dt = current_timestep_, d1 = old_timestep_, d2 = old_old_timestep_;
    p.ab[0] = (dt / d1) * ((2 * dt + 6 * d1 + 3 * d2) / (d1 + d2)) + 6;
    p.ab[1] = -(dt / d1) * ((2 * dt + 3 * d1 + 3 * d2) / (d2));
    p.ab[2] = (dt / d2) * ((2 * dt + 3 * d1) / (d1 + d2));
    p.der[0] = (2 * d1 + d2) / (d1 * (d1 + d2));
    p.der[1] = -(d1 + d2) / (d1 * d2);
    p.der[2] = (d1) / (d2 * (d1 + d2));
    p.der[3] = (d2) / (d1 * (d1 + d2));
    p.der[4] = (d1 - d2) / (d1 * d2);
    p.der[5] = -(d1) / (d2 * (d1 + d2));
    p.der[6] = -(d2) / (d1 * (d1 + d2));
    p.der[7] = (d1 + d2) / (d1 * d2);
    p.der[8] = -(d1 + 2 * d2) / (d2 * (d1 + d2));
    p.kcoef = real(1) / real(15) + real(1) / real(3);     // as fillStatic
    p.dx2 = dx_ * dx_;
    p.dy2 = dy_ * dy_;
    p.coef_sl = s_.is_boussinesq ? (real)s_.still_water_elevation : real(-9999);
    p.rdx2 = real(1) / p.dx2;
    p.rdy2 = real(1) / p.dy2;
}

// createTridiagonalCoefTextures' level-0 inputs (engine.cpp:2619-2625) and the depth
// gradients of Pass3 (compute.hlsl:641-697)
void ShallowWaterEngine::fillStatic(StaticParams &p) const
{
    std::memset(&p, 0, sizeof(p));
    p.nx = nx_; p.ny = ny_; p.P = P_; p.PT = PT_;
    p.BA = BA_;
    p.BAT = BAT_;
    p.kcoef = real(1) / real(15) + real(1) / real(3);
    p.dx2 = dx_ * dx_;
    p.dy2 = dy_ * dy_;
    p.coef_sl = s_.is_boussinesq ? (real)s_.still_water_elevation : real(-9999);
    p.rdx2 = real(1) / p.dx2;
    p.rdy2 = real(1) / p.dy2;
    p.sl = s_.still_water_elevation;
    p.odx = real(1) / dx_;
    p.ody = real(1) / dy_;
    p.unsafe = d_flag_;
}

void ShallowWaterEngine::fillLines(LineParams &p) const
{
    std::memset(&p, 0, sizeof(p));
    p.nx = nx_; p.ny = ny_; p.P = P_; p.PT = PT_;
    p.nlines_x = ny_; p.nlines_y = nx_;
    p.Dx = Dx_;
    p.DyT = DyT_;
    p.BA = BA_; p.BAT = BAT_;
    p.kcoef = real(1) / real(15) + real(1) / real(3);     // as fillStatic
    p.dx2 = dx_ * dx_;
    p.dy2 = dy_ * dy_;
    p.rdx2 = real(1) / p.dx2;
    p.rdy2 = real(1) / p.dy2;
    p.coef_sl = s_.is_boussinesq ? (real)s_.still_water_elevation : real(-9999);
    p.fast = fastArith() ? 1 : 0;
    const int n = 1 - cur_;
    p.hu_out = hu_[n]; p.hvT_out = hvT_[n];
    const int c = cur_;
    p.hu_prev = hu_[c]; p.hu_prev_lo = hu_lo_[c]; p.hu_lo_out = hu_lo_[n];
    p.hvT_prev = hvT_[c]; p.hvT_prev_lo = hvT_lo_[c]; p.hvT_lo_out = hvT_lo_[n];
    p.Lx = Lx_; p.Ly = Ly_;
}

// Everything one step issues, from the host state at the start of the step
// (fillConstantBuffers + the pass setup of engine.cpp:849-1438); then the host state
// advances as after the step (engine.cpp:1476-1521).
void ShallowWaterEngine::planStep(StepLaunch &L, bool advance)
{
    std::memset(&L, 0, sizeof(L));
    const int scheme = time_scheme_now_;
    const int n = 1 - cur_;
    L.scheme = scheme;
    L.rows = rows_;
    L.shape_x = shape_x_;
    L.shape_y = shape_y_;
    // host state the plan moves on (restored when only peeking)
    const int keep_i[8] = {cur_, ring_old_, ring_oo_, ring_scr_, time_scheme_now_, timestep_count_,
                           correction_steps_count_, 0};
    const real keep_f[4] = {total_time_, current_timestep_, old_timestep_, old_old_timestep_};
    int keep_uts[4];
    for (int s = 0; s < 4; ++s) keep_uts[s] = side_[s].uts_iter;
    fillExplicit(L.ep);
    L.aux = aux_on_;
    if (aux_on_) fillAux(L.ap);
    // predictor copies for a following corrector (engine.cpp:1050-1058); without corrector
    // steps nothing ever reads them
    if ((scheme == CEL_PREDICTOR || scheme == CEL_PREDICTOR_ADAPTIVE) && s_.correction_steps_num > 0) {
        const size_t plane = (size_t)P_ * Y_ * sizeof(real), tplane = (size_t)PT_ * X_ * sizeof(real);
        auto add = [&](real *d, const real *src, size_t bytes) { L.copy[L.ncopy++] = StepLaunch::Copy{d, src, bytes}; };
        add(Hp_, H_[ring_scr_], 4 * plane);
        add(Hp1_, H1_[ring_scr_], plane);
        add(Cw_, w_[cur_], plane);
        add(Chu_, hu_[cur_], plane);
        add(ChvT_, hvT_[cur_], tplane);
    }
    // line solves + irregular-wave lines
    fillLines(L.lp);
    IrregularAll &irr = L.irr;
    irr.t = total_time_;
    for (int s = 0; s < 4; ++s) {
        const SideState &sd = side_[s];
        if (sd.cfg.type != CEL_IRREGULARWAVES || sd.cfg.width == 0) continue;
        const bool xs = s == CEL_WEST || s == CEL_EAST;
        IrregularSide &S = irr.side[s];
        const int Lv = xs ? Lx_ : Ly_;
        const int pn = 1 << (Lv - 1);
        S.active = 1;
        S.width = sd.cfg.width;
        S.nlines = xs ? Y_ : X_;
        S.nleaf_active = sd.irr_nw;
        int p2 = 1;
        while (p2 < sd.irr_nw) p2 <<= 1;
        S.nleaf_pow2 = p2;
        int above = 0;
        for (int v = p2; v < pn; v <<= 1) ++above;
        S.levels_above = above;
        S.zero_line = Lv == 1;
        S.sl = sd.cfg.water_level;
        const int nw1 = std::max(sd.irr_nw, 1);
        S.amp = sd.d_irr_wave; S.omega = sd.d_irr_wave + nw1; S.phase = sd.d_irr_wave + 2 * nw1;
        S.tab = sd.d_tab; S.wet = sd.d_wet; S.nw = nw1; S.out = sd.d_irr_out;
        irr.count[s] = S.width * S.nlines;
    }
    // boundaries (engine.cpp:1235-1438), boundary time = t^n (engine.cpp:1324)
    BoundaryParams &bp = L.bp;
    bp.nx = nx_; bp.ny = ny_; bp.P = P_; bp.PT = PT_; bp.X = X_; bp.Y = Y_;
    bp.band_l = std::max(side_[CEL_WEST].cfg.width, 2);
    bp.band_r = std::max(side_[CEL_EAST].cfg.width, 2);
    bp.band_b = std::max(side_[CEL_SOUTH].cfg.width, 2);
    bp.band_t = std::max(side_[CEL_NORTH].cfg.width, 2);
    bp.ncells = (bp.band_l + bp.band_r) * Y_ + (X_ - bp.band_l - bp.band_r) * (bp.band_b + bp.band_t);
    // one launch when no walk can pass both solid sides of an axis (boundary_cell, pass 2)
    bp.mirror_pass = nx_ > std::max(bp.band_l, bp.band_r) + 2 && ny_ > std::max(bp.band_b, bp.band_t) + 2 ? 2 : 1;
    bp.t = total_time_;
    bp.w = w_[n]; bp.hu = hu_[n]; bp.hvT = hvT_[n];
    bp.w_lo = w_lo_[n]; bp.hu_lo = hu_lo_[n]; bp.hvT_lo = hvT_lo_[n];
    L.clock[0] = total_time_;
    for (int s = 0; s < 4; ++s) {
        SideParams &S = bp.side[s];
        const SideState &sd = side_[s];
        S.type = sd.cfg.type; S.width = sd.cfg.width; S.sl = sd.cfg.water_level;
        S.tab = sd.d_tab; S.amp = sd.cfg.amplitude; S.omega = sd.sine_omega;
        S.irr = sd.d_irr_out;
        if (S.type == CEL_UNIFORMTIMESERIES) {
            timeSeries(s, total_time_, S.uts_eta, S.uts_hu, S.uts_hv);
            L.clock[1 + 3 * s] = S.uts_eta; L.clock[2 + 3 * s] = S.uts_hu; L.clock[3 + 3 * s] = S.uts_hv;
        }
    }
    // time-scheme state machine for the next step (engine.cpp:1476-1521)
    if (timestep_count_ >= 2) {
        if (correction_steps_count_ == s_.correction_steps_num) {
            time_scheme_now_ = s_.time_scheme;
            correction_steps_count_ = 0;
        } else {
            time_scheme_now_ = CEL_CORRECTOR;
            correction_steps_count_++;
        }
    }
    if (time_scheme_now_ == CEL_PREDICTOR || time_scheme_now_ == CEL_PREDICTOR_ADAPTIVE || time_scheme_now_ == CEL_EULER) {
        timestep_count_++;
        const int t = ring_old_;
        ring_old_ = ring_scr_;
        ring_scr_ = ring_oo_;
        ring_oo_ = t;
        old_old_timestep_ = old_timestep_;
        old_timestep_ = current_timestep_;
        total_time_ += current_timestep_;
    }
    cur_ = n;
    if (!advance) {
        cur_ = keep_i[0]; ring_old_ = keep_i[1]; ring_oo_ = keep_i[2]; ring_scr_ = keep_i[3];
        time_scheme_now_ = keep_i[4]; timestep_count_ = keep_i[5]; correction_steps_count_ = keep_i[6];
        total_time_ = keep_f[0]; current_timestep_ = keep_f[1]; old_timestep_ = keep_f[2]; old_old_timestep_ = keep_f[3];
        for (int s = 0; s < 4; ++s) side_[s].uts_iter = keep_uts[s];
    }
}

// Issues one planned step on the stream.  With dclock set, the kernels take t^n and the
// time-series values from that device entry instead of their parameter blocks.
void ShallowWaterEngine::issueStep(const StepLaunch &L, const real *dclock, bool profile)
{
    hipEvent_t pe[4] = {};
    if (profile) {
        for (auto &e : pe) e = nextEvent();
        check(hipEventRecord(pe[0], stream_), "event");
    }
    // the visual outputs read only the step's starting state, which no kernel of the step writes
    if (L.aux) check(launch_aux(L.ap, stream_), "aux_update");
    check(launch_explicit(L.scheme, comp_, L.rows, fastArith(), L.ep, stream_), "explicit_step");
    if (profile) check(hipEventRecord(pe[1], stream_), "event");
    for (int i = 0; i < L.ncopy; ++i)
        check(hipMemcpyAsync(L.copy[i].dst, L.copy[i].src, L.copy[i].bytes, hipMemcpyDeviceToDevice, stream_), "copy");
    IrregularAll irr = L.irr;
    BoundaryParams bp = L.bp;
    irr.clock = dclock;
    bp.clock = dclock;
    if (profile) check(hipEventRecord(pe[2], stream_), "event");
    check(launch_solve(comp_, L.shape_x, L.shape_y, L.lp, irr, stream_), "solve_lines");
    if (profile) check(hipEventRecord(pe[3], stream_), "event");
    if (bp.mirror_pass == 2) {
        check(launch_boundary(bp, stream_), "boundary_apply");
    } else {
        bp.mirror_pass = 1;
        check(launch_boundary(bp, stream_), "boundary_apply");
        bp.mirror_pass = 0;
        check(launch_boundary(bp, stream_), "boundary_apply");
    }
    if (profile) {
        hipEvent_t e4 = nextEvent();
        check(hipEventRecord(e4, stream_), "event");
    }
}

void ShallowWaterEngine::timestep()
{
    StepLaunch L;
    planStep(L);
    issueStep(L, nullptr, profiling_);
}

// ---- captured multi-step blocks
//
// Every kernel parameter of a step is a function of the discrete engine state (ping-pong
// index, history ring, scheme counters, dt history) except the step clock.  A block of K
// steps is captured once per starting discrete state; each replay first copies that
// block's clock table (computed by planning the K steps on the host) to the device.

void ShallowWaterEngine::setGraphSteps(int k)
{
    if (k < 0 || k > kMaxGraphSteps) throw Error(CEL_EINVAL, "steps per graph must lie in [0, 1024]");
    sync();
    clearGraphs();
    if (k > 0 && !d_clock_) {
        void *mem = nullptr;
        check(hipMalloc(&mem, (size_t)kMaxGraphSteps * kClockStride * sizeof(real)), "hipMalloc(clock)");
        d_clock_ = static_cast<real *>(mem);
        for (auto &h : h_clock_) {
            check(hipHostMalloc(&mem, (size_t)kMaxGraphSteps * kClockStride * sizeof(real), hipHostMallocDefault),
                  "hipHostMalloc(clock)");
            h = static_cast<real *>(mem);
        }
        for (auto &e : clock_ev_) check(hipEventCreate(&e), "hipEventCreate");
        table_bytes_ += (size_t)kMaxGraphSteps * kClockStride * sizeof(real);
    }
    graph_steps_ = k;
}

void ShallowWaterEngine::clearGraphs()
{
    for (auto &g : graphs_) (void)hipGraphExecDestroy(g.exec);
    graphs_.clear();
}

ShallowWaterEngine::GraphKey ShallowWaterEngine::discreteKey() const
{
    GraphKey k;
    k.v[0] = cur_; k.v[1] = ring_old_; k.v[2] = ring_oo_; k.v[3] = ring_scr_;
    k.v[4] = time_scheme_now_; k.v[5] = correction_steps_count_; k.v[6] = std::min(timestep_count_, 3);
    k.v[7] = graph_steps_;
    std::memcpy(&k.v[8], &current_timestep_, 4);
    std::memcpy(&k.v[9], &old_timestep_, 4);
    std::memcpy(&k.v[10], &old_old_timestep_, 4);
    return k;
}

void ShallowWaterEngine::graphBlock()
{
    const int K = graph_steps_;
    const GraphKey key = discreteKey();
    std::vector<StepLaunch> plans(K);
    for (int k = 0; k < K; ++k) planStep(plans[k]);
    hipGraphExec_t exec = nullptr;
    for (const auto &g : graphs_)
        if (g.key == key) exec = g.exec;
    if (!exec) {
        hipGraph_t graph = nullptr;
        check(hipStreamBeginCapture(stream_, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
        try {
            for (int k = 0; k < K; ++k) issueStep(plans[k], d_clock_ + (size_t)k * kClockStride, false);
        } catch (...) {
            (void)hipStreamEndCapture(stream_, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        check(hipStreamEndCapture(stream_, &graph), "hipStreamEndCapture");
        const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        check(e, "hipGraphInstantiate");
        graphs_.push_back(GraphEntry{key, exec});
    }
    // the host clock buffer of this slot was last read by the copy two blocks ago
    real *h = h_clock_[clock_slot_];
    if (clock_used_[clock_slot_]) check(hipEventSynchronize(clock_ev_[clock_slot_]), "clock event");
    for (int k = 0; k < K; ++k) std::memcpy(h + (size_t)k * kClockStride, plans[k].clock, sizeof(plans[k].clock));
    check(hipMemcpyAsync(d_clock_, h, (size_t)K * kClockStride * sizeof(real), hipMemcpyHostToDevice, stream_),
          "upload clock");
    check(hipEventRecord(clock_ev_[clock_slot_], stream_), "clock event");
    clock_used_[clock_slot_] = true;
    clock_slot_ ^= 1;
    check(hipGraphLaunch(exec, stream_), "hipGraphLaunch");
}

void ShallowWaterEngine::step(int nsteps)
{
    if (nsteps < 0) throw Error(CEL_EINVAL, "negative step count");
    int done = 0;
    if (graph_steps_ > 0 && !profiling_)
        for (; done + graph_steps_ <= nsteps; done += graph_steps_) graphBlock();
    for (; done < nsteps; ++done) timestep();
}

void ShallowWaterEngine::sync() { check(hipStreamSynchronize(stream_), "hipStreamSynchronize"); }

float ShallowWaterEngine::timedSteps(int nsteps)
{
    check(hipEventRecord(ev_[4], stream_), "event");
    step(nsteps);
    check(hipEventRecord(ev_[5], stream_), "event");
    check(hipEventSynchronize(ev_[5]), "event sync");
    float ms = 0.f;
    check(hipEventElapsedTime(&ms, ev_[4], ev_[5]), "event");
    if (profiling_) collectProfile();
    return ms;
}

hipEvent_t ShallowWaterEngine::nextEvent()
{
    if (pool_used_ == pool_.size()) {
        hipEvent_t e;
        check(hipEventCreate(&e), "hipEventCreate");
        pool_.push_back(e);
    }
    return pool_[pool_used_++];
}

void ShallowWaterEngine::setProfiling(bool on)
{
    sync();
    if (profiling_) collectProfile();
    profiling_ = on;
    for (auto &k : kt_) k = KTime{};
}

// Each profiled step recorded five events: before explicit_step, after it, before
// solve_lines, after it, after the boundary launches.
void ShallowWaterEngine::collectProfile()
{
    sync();
    for (size_t i = 0; i + 5 <= pool_used_; i += 5) {
        float ms;
        check(hipEventElapsedTime(&ms, pool_[i], pool_[i + 1]), "event"); kt_[0].ms += ms; kt_[0].n++;
        check(hipEventElapsedTime(&ms, pool_[i + 2], pool_[i + 3]), "event"); kt_[1].ms += ms; kt_[1].n++;
        check(hipEventElapsedTime(&ms, pool_[i + 3], pool_[i + 4]), "event"); kt_[2].ms += ms; kt_[2].n++;
    }
    pool_used_ = 0;
}

bool ShallowWaterEngine::kernelTime(const std::string &name, float *ms, int *launches)
{
    if (profiling_) collectProfile();
    int i = -1;
    if (name == "explicit_step") i = 0;
    else if (name == "solve_lines") i = 1;
    else if (name == "boundary_apply") i = 2;
    if (i < 0) return false;
    if (ms) *ms = kt_[i].ms;
    if (launches) *launches = kt_[i].n;
    return true;
}

void ShallowWaterEngine::runStats(double *out9)
{
    StatsParams p;
    std::memset(&p, 0, sizeof(p));
    p.ncx = 4 * (nx_ / 4); p.ncy = 4 * (ny_ / 4);
    p.P = P_; p.hv_sx = PT_; p.hv_sy = 1;
    p.tiles_x = (p.ncx + kStatsTileX - 1) / kStatsTileX;
    p.nblocks = stats_blocks_;
    p.w = w_[cur_]; p.hu = hu_[cur_]; p.hv = hvT_[cur_]; p.BA = BA_;
    p.g = s_.gravity;
    p.one_over_dx = real(1) / dx_;
    p.one_over_dy = real(1) / dy_;
    const real m = std::max(dx_, dy_);
    p.epsilon = std::min((real)s_.epsilon, (m * m) * (m * m));
    p.partial = d_stats_;
    p.out = d_stats_ + (size_t)stats_blocks_ * kStatsValues;
    check(launch_stats(p, stream_), "stats");
    check(hipMemcpyAsync(h_stats_, p.out, kStatsValues * sizeof(double), hipMemcpyDeviceToHost, stream_), "download stats");
    check(hipStreamSynchronize(stream_), "stats sync");
    std::memcpy(out9, h_stats_, kStatsValues * sizeof(double));
}

// resetTimestep's post-processing of the reductions, engine.cpp:1686-1696 (in FP64)
void ShallowWaterEngine::deriveStats(const double *r, cel_stats *o) const
{
    cel::derive_stats(s_, (double)current_timestep_, r, o);
}

void ShallowWaterEngine::stats(cel_stats *out)
{
    if (!out) throw Error(CEL_EINVAL, "null stats");
    double raw[kStatsValues];
    runStats(raw);
    deriveStats(raw, out);
}

// resetTimestep, engine.cpp:1698-1729 (dt arithmetic in float, as the reference)
void ShallowWaterEngine::resetTimestepAdaptive(const cel_adaptive &a, cel_stats *out)
{
    double raw[kStatsValues];
    runStats(raw);
    float cfl = (float)raw[7];
    if (cfl > 0 && !target_cfl_set_ && s_.time_scheme == CEL_PREDICTOR_ADAPTIVE) {
        target_cfl_ = s_.timestep * cfl;
        target_cfl_set_ = true;
    }
    const float safety = target_cfl_set_ ? target_cfl_ : a.target_cfl;
    float cur = (float)current_timestep_;
    if (!dt_initialized_) {
        cur = s_.timestep;
        cfl = safety / cur;
        old_old_timestep_ = old_timestep_ = cur;
        dt_initialized_ = true;
    }
    if (time_scheme_now_ == CEL_PREDICTOR_ADAPTIVE) {
        const float old = (float)old_timestep_;
        const float target = std::max(safety / cfl, s_.timestep * a.min_timestep_ratio);
        cur = target > old ? a.ema_alpha * target + (1 - a.ema_alpha) * old : target;
    } else {
        cur = uniform_dt_;
    }
    if (!(cur > 0.f) || !std::isfinite(cur)) throw Error(CEL_ENUMERIC, "adaptive dt is not a positive finite number");
    const float keep_uniform = uniform_dt_;
    resetTimestep(cur);
    uniform_dt_ = keep_uniform;
    if (out) deriveStats(raw, out);
}

// ---- work decomposition

// The taller of 12- and 8-row bands whose one-wave tiles still number at least the device's
// resident waves (4 per SIMD), else 2-row bands: taller bands re-read fewer halo rows, but a
// grid whose bands do not fill the device is latency-bound (each wave's rows run one after
// another) and runs faster in more, shorter bands (measured on MI355X: 2 rows up to 1024^2,
// 8 at 1536^2, 12 at 2048^2).
int ShallowWaterEngine::default_band_rows(int device) const
{
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0)
        return kDefaultBandRows;
    const long long slots = 16LL * cus;
    const long long strips = (nx_ + kStripCols - 1) / kStripCols;
    for (int r : {12, 8})
        if (strips * ((ny_ + r - 1) / r) >= slots) return r;
    return 2;
}

void ShallowWaterEngine::getTuning(cel_tuning *out) const
{
    if (!out) throw Error(CEL_EINVAL, "null tuning");
    out->band_rows = rows_;
    out->x_chunk = kShapeC[shape_x_]; out->x_waves = kShapeW[shape_x_];
    out->y_chunk = kShapeC[shape_y_]; out->y_waves = kShapeW[shape_y_];
    out->explicit_us = tuned_us_[0]; out->x_lines_us = tuned_us_[1]; out->y_lines_us = tuned_us_[2];
    out->fast_arith = fastArith() ? 1 : 0;
}

void ShallowWaterEngine::setTuning(const cel_tuning &t)
{
    auto shape_of = [](int c, int w) {
        for (int i = 0; i < kNumShapes; ++i)
            if (kShapeC[i] == c && kShapeW[i] == w) return i;
        return -1;
    };
    if (t.band_rows != 2 && t.band_rows != 4 && t.band_rows != 8 && t.band_rows != 12 && t.band_rows != 16)
        throw Error(CEL_EINVAL, "band_rows must be 2, 4, 8, 12 or 16");
    const int sx = shape_of(t.x_chunk, t.x_waves), sy = shape_of(t.y_chunk, t.y_waves);
    if (sx < 0 || sy < 0) throw Error(CEL_EINVAL, "unsupported line-solve block shape");
    if (shape_capacity(sx) < nx_ || shape_capacity(sy) < ny_)
        throw Error(CEL_EINVAL, "line-solve block shape too small for the line length");
    sync();
    clearGraphs();
    rows_ = t.band_rows;
    shape_x_ = sx;
    shape_y_ = sy;
    fast_allowed_ = t.fast_arith != 0;
    tuned_us_[0] = tuned_us_[1] = tuned_us_[2] = 0.f;
}

void ShallowWaterEngine::setLineShapes(int sx, int sy)
{
    if (sx < 0 || sx >= kNumShapes || sy < 0 || sy >= kNumShapes || shape_capacity(sx) < nx_ ||
        shape_capacity(sy) < ny_)
        throw Error(CEL_EINVAL, "line-solve block shape too small for the line length");
    clearGraphs();
    shape_x_ = sx;
    shape_y_ = sy;
}

// Times each candidate on the current state.  The explicit kernel writes the next state's
// w, the line right-hand sides and the scratch history slot; the solves write the next
// state's hu and hv: all of it is rewritten by the next step, so tuning changes no result.
void ShallowWaterEngine::autotune(int repeats, cel_tuning *out)
{
    if (repeats < 1 || repeats > 1000) throw Error(CEL_EINVAL, "repeats must lie in [1, 1000]");
    sync();
    clearGraphs();
    const int scheme = s_.time_scheme;     // the steady-state scheme
    auto time_it = [&](auto &&launch) {
        launch();                           // warm-up
        check(hipEventRecord(ev_[6], stream_), "event");
        for (int r = 0; r < repeats; ++r) launch();
        check(hipEventRecord(ev_[7], stream_), "event");
        check(hipEventSynchronize(ev_[7]), "event sync");
        float ms = 0.f;
        check(hipEventElapsedTime(&ms, ev_[6], ev_[7]), "event");
        return 1000.f * ms / repeats;
    };
    // candidates are timed in three interleaved rounds and judged by their fastest round,
    // so clock ramps and other transients do not pick the winner
    constexpr int kRounds = 3;
    constexpr int kRowsCand = 5;
    const int rows_cand[kRowsCand] = {2, 4, 8, 12, 16};
    float rows_us[kRowsCand] = {-1.f, -1.f, -1.f, -1.f, -1.f};
    for (int round = 0; round < kRounds; ++round)
        for (int c = 0; c < kRowsCand; ++c) {
            rows_ = rows_cand[c];
            ExplicitParams p;
            fillExplicit(p);
            const float us = time_it([&] { check(launch_explicit(scheme, comp_, rows_, fastArith(), p, stream_), "explicit_step"); });
            if (rows_us[c] < 0.f || us < rows_us[c]) rows_us[c] = us;
        }
    int bc = 0;
    for (int c = 1; c < kRowsCand; ++c)
        if (rows_us[c] < rows_us[bc]) bc = c;
    rows_ = rows_cand[bc];
    tuned_us_[0] = rows_us[bc];
    LineParams lp;
    fillLines(lp);
    IrregularAll none;
    std::memset(&none, 0, sizeof(none));
    float dir_us[2][kNumShapes];
    for (int dir = 0; dir < 2; ++dir) {
        const int n = dir == 0 ? nx_ : ny_;
        LineParams q = lp;
        if (dir == 0) q.nlines_y = 0; else q.nlines_x = 0;
        float *shape_us = dir_us[dir];
        for (int sh = 0; sh < kNumShapes; ++sh) shape_us[sh] = -1.f;
        for (int round = 0; round < kRounds; ++round)
            for (int sh = 0; sh < kNumShapes; ++sh) {
                const int cap = shape_capacity(sh);
                if (cap < n || cap > std::max(4 * n, 64)) continue;
                const int sx = dir == 0 ? sh : shape_x_, sy = dir == 0 ? shape_y_ : sh;
                const float us = time_it([&] { check(launch_solve(comp_, sx, sy, q, none, stream_), "solve_lines"); });
                if (shape_us[sh] < 0.f || us < shape_us[sh]) shape_us[sh] = us;
            }
        int best_shape = -1;
        for (int sh = 0; sh < kNumShapes; ++sh)
            if (shape_us[sh] >= 0.f && (best_shape < 0 || shape_us[sh] < shape_us[best_shape])) best_shape = sh;
        if (dir == 0) shape_x_ = best_shape; else shape_y_ = best_shape;
        tuned_us_[1 + dir] = shape_us[best_shape];
    }
    // the step launches both directions at once (one launch when the shapes agree): time
    // that launch for the pair chosen above and for every common shape, keep the fastest
    {
        const int nmax = std::max(nx_, ny_), nmin = std::min(nx_, ny_);
        int best_x = shape_x_, best_y = shape_y_;
        float best_us = -1.f;
        for (int round = 0; round < kRounds; ++round)
            for (int sh = -1; sh < kNumShapes; ++sh) {
                int sx = shape_x_, sy = shape_y_;
                if (sh >= 0) {
                    const int cap = shape_capacity(sh);
                    if (cap < nmax || cap > std::max(4 * nmin, 64)) continue;
                    sx = sy = sh;
                }
                const float us = time_it([&] { check(launch_solve(comp_, sx, sy, lp, none, stream_), "solve_lines"); });
                if (best_us < 0.f || us < best_us) { best_us = us; best_x = sx; best_y = sy; }
            }
        shape_x_ = best_x;
        shape_y_ = best_y;
        tuned_us_[1] = dir_us[0][shape_x_];
        tuned_us_[2] = dir_us[1][shape_y_];
    }
    if (out) getTuning(out);
}

void ShallowWaterEngine::sample(const int *xy, int n, float *out3)
{
    if (n < 0 || (n > 0 && (!xy || !out3))) throw Error(CEL_EINVAL, "bad sample arguments");
    if (n == 0) return;
    for (int k = 0; k < n; ++k)
        if (xy[2 * k] < 0 || xy[2 * k] >= X_ || xy[2 * k + 1] < 0 || xy[2 * k + 1] >= Y_)
            throw Error(CEL_EINVAL, "sample cell outside the grid");
    void *mem = nullptr;
    const size_t out_off = ((size_t)n * 2 * sizeof(int) + 15) / 16 * 16;   // 16-byte aligned output
    const size_t bytes = out_off + (size_t)n * 3 * sizeof(real);
    check(hipMalloc(&mem, bytes), "hipMalloc(sample)");
    int *dxy = static_cast<int *>(mem);
    real *dout = reinterpret_cast<real *>(static_cast<char *>(mem) + out_off);
    std::vector<real> h((size_t)3 * n);
    hipError_t e = hipMemcpyAsync(dxy, xy, (size_t)n * 2 * sizeof(int), hipMemcpyHostToDevice, stream_);
    if (e == hipSuccess) e = launch_gather(w_[cur_], hu_[cur_], hvT_[cur_], P_, PT_, dxy, n, dout, stream_);
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), dout, (size_t)n * 3 * sizeof(real), hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    (void)hipFree(mem);
    check(e, "sample");
    for (size_t i = 0; i < h.size(); ++i) out3[i] = (float)h[i];
}

// ---- terrain and water edits

// The w plane of the current state on the host (ghost-inclusive, row-major) and back.
void ShallowWaterEngine::downloadW(std::vector<real> &w)
{
    const size_t e = sizeof(real);
    w.resize((size_t)X_ * Y_);
    check(hipMemcpy2DAsync(w.data(), (size_t)X_ * e, w_[cur_], (size_t)P_ * e, (size_t)X_ * e, Y_,
                           hipMemcpyDeviceToHost, stream_), "download w");
    check(hipStreamSynchronize(stream_), "sync");
}

void ShallowWaterEngine::uploadW(const std::vector<real> &w)
{
    const size_t e = sizeof(real);
    check(hipMemcpy2DAsync(w_[cur_], (size_t)P_ * e, w.data(), (size_t)X_ * e, (size_t)X_ * e, Y_,
                           hipMemcpyHostToDevice, stream_), "upload w");
}

// After a bottom change: the transposed bed and the fast-variant check, the
// depth-dependent boundary tables, no captured blocks (their tables moved).
void ShallowWaterEngine::bottomChanged()
{
    clearGraphs();
    uploadHostBottom();
    for (int i = 0; i < 4; ++i) buildSideTables(i);
}

// fillTerrainTexture (keep_depth) / fillTerrainTextureLite, engine.cpp:3313-3412.
void ShallowWaterEngine::setBottom(const float *bottom3, bool keep_depth)
{
    if (!bottom3) throw Error(CEL_EINVAL, "null bottom");
    sync();
    const size_t N = (size_t)X_ * Y_;
    if (keep_depth) {
        std::vector<real> w;
        downloadW(w);
        for (size_t i = 0; i < N; ++i) w[i] = (w[i] - hostBottom_[2 * N + i]) + (real)bottom3[2 * N + i];
        uploadW(w);
    }
    hostBottom_.assign(bottom3, bottom3 + 3 * N);
    if (aux_on_) zeroAux();
    bottomChanged();
}

namespace {
// GaussianBrush, engine.cpp:93-99
real gaussian_brush(real x, real y, real r)
{
    const real sigma = real(1.0f) / real(3.0f);
    const real z = (x * x + y * y) / (r * r);
    if (z > real(1)) return real(0);
    return real(2) * std::exp(-z / (real(2) * sigma * sigma));
}
}  // namespace

// raiseLowerTerrain, engine.cpp:4685-4843 (CPU code in the reference too).
void ShallowWaterEngine::raiseTerrain(float world_x, float world_y, float radius, float displacement)
{
    sync();
    const size_t N = (size_t)X_ * Y_;
    const float W = s_.width, L = s_.length;
    const float brush_radius = radius * 2;
    const float sx = (nx_ - 1) / W, bx = (nx_ + 4) / 2.0f, sy = (ny_ - 1) / L, by = 2.5f;
    const int cix = (int)(sx * world_x + bx), ciy = (int)(sy * world_y + by);
    const int rx = (int)(sx * brush_radius) + 3, ry = (int)(sy * brush_radius) + 3;
    const int ix0 = std::max(0, std::min(nx_ + 4, cix - rx)), ix1 = std::max(0, std::min(nx_ + 4, cix + rx + 1));
    const int iy0 = std::max(0, std::min(ny_ + 4, ciy - ry)), iy1 = std::max(0, std::min(ny_ + 4, ciy + ry + 1));
    std::vector<real> w;
    downloadW(w);
    real *BY = hostBottom_.data(), *BX = BY + N, *BA = BY + 2 * N;
    const real br = (real)brush_radius, disp = (real)displacement;
    for (int iy = iy0; iy < iy1; ++iy) {
        const real wyp = (real)((iy + 0.5f - by) / sy - world_y), wym = (real)((iy - 0.5f - by) / sy - world_y);
        for (int ix = ix0; ix < ix1; ++ix) {
            const real wxp = (real)((ix + 0.5f - bx) / sx - world_x), wxm = (real)((ix - 0.5f - bx) / sx - world_x);
            const real pp = gaussian_brush(wxp, wyp, br) * disp, pm = gaussian_brush(wxp, wym, br) * disp;
            const real mp = gaussian_brush(wxm, wyp, br) * disp, mm = gaussian_brush(wxm, wym, br) * disp;
            const size_t i = (size_t)iy * X_ + ix;
            w[i] += (pp + pm + mp + mm) / real(4);
            BX[i] += (pp + pm) / real(2);
            BY[i] += (pp + mp) / real(2);
            BA[i] += (pp + pm + mp + mm) / real(4);
        }
    }
    uploadW(w);
    bottomChanged();
}

// shiftTerrainSlider (engine.cpp:4845-4924) with the auxiliary reset resetTimestep does after
// it (engine.cpp:1544-1553).
void ShallowWaterEngine::shiftTerrain(float shift_up)
{
    sync();
    const size_t N = (size_t)X_ * Y_;
    std::vector<real> w;
    downloadW(w);
    real *BY = hostBottom_.data(), *BX = BY + N, *BA = BY + 2 * N;
    const real swl = (real)s_.still_water_elevation, sh = (real)shift_up;
    for (size_t i = 0; i < N; ++i) {
        if ((double)(w[i] - BA[i]) < 0.01) {
            if (BA[i] < swl) w[i] = swl;
            else w[i] += sh;
        }
        BX[i] += sh;
        BY[i] += sh;
        BA[i] += sh;
    }
    uploadW(w);
    if (aux_on_) zeroAux();
    bottomChanged();
}

// The left-mouse water brush: 12-triangle fan rasterised with D3D11's rules on the cells'
// centres (applyMouseShader, engine.cpp:4561-4683; left_mouse.hlsl:59-103).
void ShallowWaterEngine::editWater(float world_x, float world_y, float radius, float disp_A, float disp_B)
{
    if (!(radius > 0.f)) throw Error(CEL_EINVAL, "brush radius must be positive");
    WaterBrushParams p;
    const float W = s_.width, L = s_.length;
    const float sxw = (nx_ - 1) / W, syw = (ny_ - 1) / L;
    const float scx = sxw * radius, bcx = sxw * world_x + (nx_ + 4) / 2.0f;
    const float scy = syw * radius, bcy = syw * world_y + 2.5f;
    const float PI_F = 4.0f * std::atan(1.0f), dth = 2 * PI_F / 12;
    double lo_x = 1e300, hi_x = -1e300, lo_y = 1e300, hi_y = -1e300;
    for (int k = 0; k <= 12; ++k) {
        float px = 0.f, py = 0.f;
        if (k < 12) { const float th = dth * k; px = std::cos(th); py = std::sin(th); }
        const float tx = scx * px + bcx, ty = scy * py + bcy;
        // vertex positions snapped to 1/256 of a cell (k = 12: the centre)
        p.vx[k] = std::nearbyint((double)tx * 256.0) / 256.0;
        p.vy[k] = std::nearbyint((double)ty * 256.0) / 256.0;
        lo_x = std::min(lo_x, p.vx[k]); hi_x = std::max(hi_x, p.vx[k]);
        lo_y = std::min(lo_y, p.vy[k]); hi_y = std::max(hi_y, p.vy[k]);
    }
    if (!(hi_x > lo_x) || !(hi_y > lo_y) || !std::isfinite(hi_x + hi_y + lo_x + lo_y)) return;
    // cells whose centres can lie inside the fan, clipped to the grid
    p.x0 = (int)std::max(0.0, std::floor(lo_x - 1.0));
    p.y0 = (int)std::max(0.0, std::floor(lo_y - 1.0));
    const int x1 = (int)std::min((double)X_, std::ceil(hi_x + 1.0));
    const int y1 = (int)std::min((double)Y_, std::ceil(hi_y + 1.0));
    if (x1 <= p.x0 || y1 <= p.y0) return;
    p.bw = x1 - p.x0;
    p.bh = y1 - p.y0;
    p.P = P_; p.PT = PT_;
    p.w = w_[cur_]; p.hu = hu_[cur_]; p.hvT = hvT_[cur_];
    p.w_lo = w_lo_[cur_]; p.hu_lo = hu_lo_[cur_]; p.hvT_lo = hvT_lo_[cur_];
    p.BA = BA_;
    p.disp_A = disp_A;
    p.disp_B = disp_B;
    check(launch_water_brush(p, stream_), "water_brush");
}

// ---- visual outputs (compute.hlsl:437-457, 535-540) and addSolitaryWave

void ShallowWaterEngine::fillAux(AuxParams &p) const
{
    const size_t plane = (size_t)P_ * Y_;
    p.nx = nx_; p.ny = ny_; p.P = P_;
    p.w = w_[cur_]; p.BX = BX_; p.BY = BY_;
    p.depth = aux_mem_; p.slope = aux_mem_ + plane; p.white = aux_mem_ + 2 * plane;
    p.two_theta = real(2) * (real)s_.theta;
    p.one_over_dx = real(1) / dx_;
    p.one_over_dy = real(1) / dy_;
    p.threshold = aux_threshold_;
    // pow(abs(whiteWaterDecayRate), dt) is uniform over the grid: evaluated once per step
    p.decay = std::pow(std::fabs(aux_rate_), current_timestep_);
}

void ShallowWaterEngine::zeroAux()
{
    check(hipMemsetAsync(aux_mem_, 0, (size_t)P_ * Y_ * 3 * sizeof(real), stream_), "zero aux");
}

void ShallowWaterEngine::setVisual(bool on, float threshold, float decay)
{
    sync();
    clearGraphs();
    if (on) {
        if (!aux_mem_) {
            void *mem = nullptr;
            check(hipMalloc(&mem, (size_t)P_ * Y_ * 3 * sizeof(real)), "hipMalloc(aux)");
            aux_mem_ = static_cast<real *>(mem);
            table_bytes_ += (size_t)P_ * Y_ * 3 * sizeof(real);
        }
        zeroAux();
        aux_threshold_ = (real)threshold;
        aux_rate_ = (real)(1.0f / (1 + 2 * decay));     // engine.cpp:4129
        check(hipStreamSynchronize(stream_), "sync");
    } else if (aux_mem_) {
        (void)hipFree(aux_mem_);
        aux_mem_ = nullptr;
        table_bytes_ -= (size_t)P_ * Y_ * 3 * sizeof(real);
    }
    aux_on_ = on;
}

void ShallowWaterEngine::getAux(float *aux3)
{
    if (!aux_on_) throw Error(CEL_EINVAL, "visual outputs are off");
    if (!aux3) throw Error(CEL_EINVAL, "null output");
    const size_t N = (size_t)X_ * Y_, e = sizeof(real);
    std::vector<real> h(3 * N);
    for (int c = 0; c < 3; ++c)
        check(hipMemcpy2DAsync(h.data() + c * N, (size_t)X_ * e, aux_mem_ + (size_t)c * P_ * Y_, (size_t)P_ * e,
                               (size_t)X_ * e, Y_, hipMemcpyDeviceToHost, stream_), "download aux");
    check(hipStreamSynchronize(stream_), "sync");
    for (size_t i = 0; i < 3 * N; ++i) aux3[i] = (float)h[i];
}

void ShallowWaterEngine::getNormals(float *normal3)
{
    if (!normal3) throw Error(CEL_EINVAL, "null output");
    const size_t N = (size_t)X_ * Y_;
    void *mem = nullptr;
    check(hipMalloc(&mem, 3 * N * sizeof(real)), "hipMalloc(normals)");
    NormalParams p;
    p.X = X_; p.Y = Y_; p.P = P_;
    p.w = w_[cur_];
    p.one_over_dx = real(1) / dx_;
    p.one_over_dy = real(1) / dy_;
    p.out = static_cast<real *>(mem);
    std::vector<real> h(3 * N);
    hipError_t e = launch_normals(p, stream_);
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), p.out, 3 * N * sizeof(real), hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    (void)hipFree(mem);
    check(e, "normals");
    for (size_t i = 0; i < 3 * N; ++i) normal3[i] = (float)h[i];
}

void ShallowWaterEngine::addSolitaryWave(float H, float theta_deg, float x0, float y0)
{
    SolitonParams p;
    p.X = X_; p.Y = Y_; p.P = P_; p.PT = PT_;
    p.w = w_[cur_]; p.hu = hu_[cur_]; p.hvT = hvT_[cur_];
    p.w_lo = w_lo_[cur_]; p.hu_lo = hu_lo_[cur_]; p.hvT_lo = hvT_lo_[cur_];
    p.BA = BA_;
    p.g = (real)s_.gravity;
    p.dx = dx_; p.dy = dy_;
    p.H = (real)H;
    // sw_theta = theta * PI / 180.0f with PI = 4 atan(1) in FP32 (engine.cpp:90, 2167)
    const float PI_F = 4.0f * std::atan(1.0f);
    const real theta = (real)(theta_deg * PI_F / 180.0f);
    p.cos_t = std::cos(theta);
    p.sin_t = std::sin(theta);
    p.x0 = (real)x0; p.y0 = (real)y0;
    check(launch_soliton(p, stream_), "add_soliton");
}

void ShallowWaterEngine::coefficients(float *ax, float *ay)
{
    StaticParams sp;
    fillStatic(sp);
    const size_t N = (size_t)X_ * Y_;
    real *d = nullptr;
    check(hipMalloc(&d, N * 6 * sizeof(real)), "hipMalloc");
    std::vector<real> h(N * 6);
    check(launch_export_coefficients(sp, d, d + 3 * N, stream_), "export_coefficients");
    check(hipMemcpyAsync(h.data(), d, N * 6 * sizeof(real), hipMemcpyDeviceToHost, stream_), "download");
    check(hipStreamSynchronize(stream_), "sync");
    (void)hipFree(d);
    for (size_t i = 0; i < 3 * N; ++i) { ax[i] = (float)h[i]; ay[i] = (float)h[3 * N + i]; }
}

}  // namespace CEL_NS
