// kernels.hip -- gfx950 kernels of the Celeris Boussinesq step.
//
//   explicit_step   : MUSCL/minmod reconstruction + KP07 central-upwind fluxes + flux
//                     divergence, bed slope, friction, Madsen-Sorensen dispersion and the
//                     Euler / AB3 / AM4 / variable-step AB3 update, fused.  One wave
//                     marches down a band of rows; the 64 lanes own 64 adjacent columns
//                     (60 produce output, 2 halo lanes each side) and exchange
//                     horizontal neighbours through cross-lane moves, vertical
//                     neighbours stay in registers.  (compute.hlsl Pass1/Pass2/Pass3)
//   solve_lines_block : one workgroup per x-line or y-line; the reference's cyclic
//                     reduction with each thread holding a chunk of the line in registers
//                     and the upper levels in LDS; coefficients recomputed from BA.
//                     (engine.cpp:2612-2887 + compute.hlsl:835-947)
//   irregular_lines : irregular-wave superposition with the reference's pairwise-tree
//                     sum order (compute.hlsl:1196-1309), merged into the solve launch.
//   boundary_apply  : the four boundary strips in the order E, N, W, S evaluated as one
//                     chain per cell (compute.hlsl:985-1194, engine.cpp:1345-1438).
//
// Arithmetic keeps the shader's operation order; the build disables FMA contraction so
// that every product and sum rounds as written.  The file is compiled twice: with
// real = float (namespace cel, the reference's precision) and, with CEL_FP64 defined,
// real = double (namespace cel64, the FP64 variant).
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "params.hpp"


namespace CEL_NS {

// math in the working precision
__device__ __forceinline__ float SQRT(float x) { return sqrtf(x); }
__device__ __forceinline__ double SQRT(double x) { return sqrt(x); }
__device__ __forceinline__ float SIN(float x) { return sinf(x); }
__device__ __forceinline__ double SIN(double x) { return sin(x); }
__device__ __forceinline__ float FMAX(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ double FMAX(double a, double b) { return fmax(a, b); }
__device__ __forceinline__ float FMIN(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ double FMIN(double a, double b) { return fmin(a, b); }
__device__ __forceinline__ float FABS(float x) { return fabsf(x); }
__device__ __forceinline__ double FABS(double x) { return fabs(x); }
__device__ __forceinline__ float POW(float x, float y) { return powf(x, y); }
__device__ __forceinline__ double POW(double x, double y) { return pow(x, y); }

// Correctly rounded quotient and square root in fewer instructions, for the FAST variants
// of the explicit kernel (FP32; the FP64 build keeps the plain operations).  On the
// arguments for which the engine selects them they return the bits of a / b and sqrtf
// (up to the sign of a zero result, which never changes a nonzero value downstream):
//   div_k(a, b, rb): b fixed for the launch, rb = RN(1/b).  q = a rb lies within one ulp of
//     a / b, the remainder a - b q is exact in one FMA, and q + (a - b q) rb rounds to the
//     correctly rounded quotient (Markstein's theorem) for a == 0 or 2^-80 <= |a| <= 2^80
//     and 2^-20 <= b <= 2^20 (no intermediate leaves the normal range).  The library
//     division spends about ten instructions on a general b.
//   sqrt_k(x): x == 0 or 2^-96 <= x < inf.  The hardware root (within one ulp) is corrected
//     by the sign of the residual on either neighbour, as the compiler's own expansion of
//     sqrtf does for arguments that need no rescaling; the rescaling and the special-value
//     tests it adds for the others are dropped.
//   sqrt_ks(x): any x that is not NaN.  The compiler's expansion with its rescaling of
//     x < 2^-96 (by 2^32, the root by 2^-16) but without its final test for +-0 and +inf:
//     the corrected root already returns x there (the neighbour residuals come out NaN or
//     zero and select nothing).
template <bool FAST>
__device__ __forceinline__ real div_k(real a, real b, real rb)
{
#ifndef CEL_FP64
    if (FAST) {
        const float q = a * rb;
        return __builtin_fmaf(__builtin_fmaf(-b, q, a), rb, q);
    }
#endif
    (void)rb;
    return a / b;
}
template <bool FAST>
__device__ __forceinline__ real sqrt_ks(real x)
{
#ifndef CEL_FP64
    if (FAST) {
        const bool small = x < 0x1p-96f;
        const float xs = small ? x * 0x1p32f : x;
        const float s = __builtin_amdgcn_sqrtf(xs);
        const float sm = __int_as_float(__float_as_int(s) - 1);
        const float sp = __int_as_float(__float_as_int(s) + 1);
        float r = __builtin_fmaf(-sm, s, xs) <= 0.f ? sm : s;
        r = __builtin_fmaf(-sp, s, xs) > 0.f ? sp : r;
        return small ? r * 0x1p-16f : r;
    }
#endif
    return SQRT(x);
}
template <bool FAST>
__device__ __forceinline__ real sqrt_k(real x)
{
#ifndef CEL_FP64
    if (FAST) {
        const float s = __builtin_amdgcn_sqrtf(x);
        const float sm = __int_as_float(__float_as_int(s) - 1);
        const float sp = __int_as_float(__float_as_int(s) + 1);
        const float r = __builtin_fmaf(-sm, s, x) <= 0.f ? sm : s;
        return __builtin_fmaf(-sp, s, x) > 0.f ? sp : r;
    }
#endif
    return SQRT(x);
}

// 16 bytes of consecutive elements (4 floats or 2 doubles) and a pair of elements, for
// vector loads and stores
constexpr int kVec = 16 / (int)sizeof(real);
struct alignas(16) Vec { real v[kVec]; };
struct alignas(2 * sizeof(real)) Pair { real x, y; };
// the four history values of one cell in a quad plane (one 16-byte access in FP32)
struct alignas(4 * sizeof(real)) Quad { real v[4]; };
// the pieces a band of R rows stages and stores its transposed columns in: 16 bytes, or
// two elements for 2-row bands (whose first row is only even)
template <int R>
struct alignas(sizeof(real) * (R % kVec == 0 ? kVec : 2)) BandPiece {
    static constexpr int n = R % kVec == 0 ? kVec : 2;
    real v[n];
};

// ------------------------------------------------------------------ helpers

// IEEE min/max (v_min_f32 / v_max_f32, fused into v_min3/v_max3).  They agree with the
// shader's max/min on every non-NaN input up to the sign of a zero result, which never
// changes a nonzero value downstream.
__device__ __forceinline__ real fmaxr(real a, real b) { return FMAX(a, b); }
__device__ __forceinline__ real fminr(real a, real b) { return FMIN(a, b); }

// lane i receives lane i+1 / i-1 through a DPP wave shift (edge lanes keep their own
// value; they are halo lanes); a double moves as its two 32-bit halves
template <int CTRL>
__device__ __forceinline__ float dpp_move(float v)
{
    const int b = __builtin_bit_cast(int, v);
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(b, b, CTRL, 0xf, 0xf, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const int lo = (int)(unsigned)(b & 0xffffffffull), hi = (int)(unsigned)(b >> 32);
    const unsigned nlo = (unsigned)__builtin_amdgcn_update_dpp(lo, lo, CTRL, 0xf, 0xf, false);
    const unsigned nhi = (unsigned)__builtin_amdgcn_update_dpp(hi, hi, CTRL, 0xf, 0xf, false);
    return __builtin_bit_cast(double, ((unsigned long long)nhi << 32) | nlo);
}
__device__ __forceinline__ real from_right(real v) { return dpp_move<0x130>(v); }   // wave_shl:1
__device__ __forceinline__ real from_left(real v) { return dpp_move<0x138>(v); }    // wave_shr:1

// MinMod, compute.hlsl:292-296: the minimum of three positive slopes, the maximum of three
// negative ones, else zero.  That is the median of (0, min, max): min when all three are
// positive, max when all are negative, 0 otherwise -- one v_med3 after min3/max3.
__device__ __forceinline__ float MED3(float a, float b, float c) { return __builtin_amdgcn_fmed3f(a, b, c); }
__device__ __forceinline__ double MED3(double a, double b, double c) { return fmax(fmin(a, b), fmin(fmax(a, b), c)); }
__device__ __forceinline__ real minmod3(real a, real b, real c)
{
    const real mn = FMIN(FMIN(a, b), c);
    const real mx = FMAX(FMAX(a, b), c);
    return MED3(real(0), mn, mx);
}

// Reconstruct, compute.hlsl:298-315: half-slope of the generalized minmod limiter
__device__ __forceinline__ real half_slope(real two_theta, real west, real here, real east)
{
    const real z1 = two_theta * (here - west);
    const real z2 = (east - west);
    const real z3 = two_theta * (east - here);
    return 0.25f * minmod3(z1, z2, z3);
}

// CorrectW, compute.hlsl:317-328, as selects
__device__ __forceinline__ void correct_w(real B_west, real B_east, real w_bar, real &w_west, real &w_east)
{
    const bool c1 = w_east < B_east;
    const bool c2 = !c1 && w_west < B_west;
    const real alt_w = fmaxr(B_west, 2.f * w_bar - B_east);
    const real alt_e = fmaxr(B_east, 2.f * w_bar - B_west);
    const real ne = c1 ? B_east : (c2 ? alt_e : w_east);
    const real nw = c1 ? alt_w : (c2 ? B_west : w_west);
    w_east = ne;
    w_west = nw;
}

// desingularised 1/h of CalcUV, compute.hlsl:330-340 (the root's argument is >= eps)
template <bool FAST = false>
__device__ __forceinline__ real divide_by_h(real h, real eps)
{
    const real h2 = h * h;
    const real h4 = h2 * h2;
    return SQRT(real(2)) * h / sqrt_k<FAST>(h4 + fmaxr(h4, eps));
}

// NumericalFlux, compute.hlsl:363-370
__device__ __forceinline__ real num_flux(real ap, real am, real Fp, real Fm, real dU)
{
    const real den = ap - am;
    const real q = (ap * Fm - am * Fp + ap * am * dU) / den;   // discarded when den == 0
    return den != 0.f ? q : 0.f;
}

// Knuth's TwoSum: s + e == a + b exactly (round-to-nearest, no contraction).
__device__ __forceinline__ void two_sum(real a, real b, real &s, real &e)
{
    s = a + b;
    const real bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// Element access through a 32-bit byte offset from a uniform base: maps to the
// scalar-base + vector-offset global addressing mode (planes are < 4 GiB).
__device__ __forceinline__ real ld32(const real *base, unsigned idx)
{
    return *reinterpret_cast<const real *>(reinterpret_cast<const char *>(base) + idx * (unsigned)sizeof(real));
}
__device__ __forceinline__ void st32(real *base, unsigned idx, real v)
{
    *reinterpret_cast<real *>(reinterpret_cast<char *>(base) + idx * (unsigned)sizeof(real)) = v;
}
// quad-plane element idx (grids up to 8192 x 8192 keep the byte offset below 4 GiB in FP64)
__device__ __forceinline__ Quad ldq(const real *base, unsigned idx)
{
    return *reinterpret_cast<const Quad *>(reinterpret_cast<const char *>(base) + idx * (unsigned)sizeof(Quad));
}
__device__ __forceinline__ void stq(real *base, unsigned idx, const Quad &v)
{
    *reinterpret_cast<Quad *>(reinterpret_cast<char *>(base) + idx * (unsigned)sizeof(Quad)) = v;
}

struct Face { real h, u, v; };

// Face values of one cell along one axis (Pass1, compute.hlsl:408-433): returns the
// "low" face (W or S) and the "high" face (E or N).
template <bool FAST>
__device__ __forceinline__ void faces_1d(real two_theta, real eps,
                                         real w_lo, real w_c, real w_hi,
                                         real hu_lo, real hu_c, real hu_hi,
                                         real hv_lo, real hv_c, real hv_hi,
                                         real B_lo, real B_hi, Face &lo, Face &hi)
{
    const real sw = half_slope(two_theta, w_lo, w_c, w_hi);
    real w_hi_f = w_c + sw, w_lo_f = w_c - sw;
    const real su = half_slope(two_theta, hu_lo, hu_c, hu_hi);
    const real hu_hi_f = hu_c + su, hu_lo_f = hu_c - su;
    const real sv = half_slope(two_theta, hv_lo, hv_c, hv_hi);
    const real hv_hi_f = hv_c + sv, hv_lo_f = hv_c - sv;
    correct_w(B_lo, B_hi, w_c, w_lo_f, w_hi_f);
    hi.h = w_hi_f - B_hi;
    lo.h = w_lo_f - B_lo;
    const real qh = divide_by_h<FAST>(hi.h, eps);
    hi.u = qh * hu_hi_f;
    hi.v = qh * hv_hi_f;
    const real ql = divide_by_h<FAST>(lo.h, eps);
    lo.u = ql * hu_lo_f;
    lo.v = ql * hv_lo_f;
}

// Pass2 x-direction flux between cell E face (here) and the W face of the east cell.
// normal velocity u, tangential v; for y-direction call with (v, u) swapped and the
// momentum components mapped as in compute.hlsl:517-533.
template <bool FAST>
__device__ __forceinline__ void flux_x(real g, real half_g, const Face &E, const Face &Wn, real &f0, real &f1, real &f2)
{
    const real cE = sqrt_ks<FAST>(g * E.h);
    const real cW = sqrt_ks<FAST>(g * Wn.h);
    const real ap = fmaxr(fmaxr(E.u + cE, Wn.u + cW), 0.f);
    const real am = fminr(fminr(E.u - cE, Wn.u - cW), 0.f);
    f0 = num_flux(ap, am, Wn.h * Wn.u, E.h * E.u, Wn.h - E.h);
    f1 = num_flux(ap, am, Wn.h * (Wn.u * Wn.u + half_g * Wn.h), E.h * (E.u * E.u + half_g * E.h),
                  Wn.h * Wn.u - E.h * E.u);
    f2 = num_flux(ap, am, Wn.h * Wn.u * Wn.v, E.h * E.u * E.v, Wn.h * Wn.v - E.h * E.v);
}

template <bool FAST>
__device__ __forceinline__ void flux_y(real g, real half_g, const Face &N, const Face &Sn, real &f0, real &f1, real &f2)
{
    const real cN = sqrt_ks<FAST>(g * N.h);
    const real cS = sqrt_ks<FAST>(g * Sn.h);
    const real bp = fmaxr(fmaxr(N.v + cN, Sn.v + cS), 0.f);
    const real bm = fminr(fminr(N.v - cN, Sn.v - cS), 0.f);
    f0 = num_flux(bp, bm, Sn.h * Sn.v, N.h * N.v, Sn.h - N.h);
    f1 = num_flux(bp, bm, Sn.h * Sn.u * Sn.v, N.h * N.u * N.v, Sn.h * Sn.u - N.h * N.u);
    f2 = num_flux(bp, bm, Sn.h * (Sn.v * Sn.v + half_g * Sn.h), N.h * (N.v * N.v + half_g * N.h),
                  Sn.h * Sn.v - N.h * N.v);
}

// Level-0 tridiagonal coefficients (engine.cpp:2664-2680) of a line cell: K1 and K2/72
// from the depths at offsets -2..+2 along the line, then A, B, C.  rd2 = RN(1/d2).
template <bool FAST = false>
__device__ __forceinline__ void line_kq(real kcoef, real d2, real rd2, real dm2, real dm1, real d0, real dp1,
                                        real dp2, real &K1, real &Q)
{
    K1 = kcoef * div_k<FAST>(d0 * d0, d2, rd2);
    const real K2 = div_k<FAST>(d0, d2, rd2) * (-dp2 + 8.f * dp1 - 8.f * dm1 + dm2);
    Q = div_k<FAST>(K2, real(72), real(1) / real(72));
}
__device__ __forceinline__ void line_abc(real K1, real Q, bool first, bool last, real &A, real &B, real &C)
{
    A = first ? 0.f : -K1 + Q;
    B = 1.f + 2.f * K1;
    C = last ? 0.f : -(K1 + Q);
}

// ------------------------------------------------------------------ explicit step

// Per-cell inputs of Pass3 that are read once: the Adams-Bashforth history (and the low
// word of w in the compensated state).
struct CellIn {
    real o0, o1, o2, oo0, oo1, oo2, Fo, Go, Foo, Goo;
    real wlo;      // low word of w (compensated state)
};

template <int SCHEME, bool COMP>
__device__ __forceinline__ CellIn load_cell(const ExplicitParams &p, unsigned i)
{
    CellIn c;
    if (SCHEME != 0) {
        const Quad a = ldq(p.Ho, i), b = ldq(p.Hoo, i);
        c.o0 = a.v[0]; c.o1 = a.v[1]; c.o2 = a.v[2]; c.Fo = a.v[3];
        c.oo0 = b.v[0]; c.oo1 = b.v[1]; c.oo2 = b.v[2]; c.Foo = b.v[3];
        c.Go = ld32(p.Ho1, i); c.Goo = ld32(p.Hoo1, i);
    } else {
        c.o0 = c.o1 = c.o2 = c.oo0 = c.oo1 = c.oo2 = c.Fo = c.Go = c.Foo = c.Goo = 0.f;
    }
    c.wlo = COMP ? ld32(p.w_lo, i) : 0.f;
    return c;
}

// LDS of one explicit tile: the band's transposed hv (and the corrector's start hv),
// and the y-line right-hand sides, stored transposed after the band.
template <bool CORRECTOR, int R>
struct ExplicitSmem {
    real hv[R + 4][64];
    real chv[CORRECTOR ? R + 4 : 1][64];
    real dy[R][64];
};

// One wave computes interior columns [x0, x0 + 60) of rows [ys, ys + R) of one scenario.
// SCHEME: 0 Euler, 1 AB3 predictor, 2 AM4 corrector, 3 variable-step AB3
template <int SCHEME, bool COMP, int R, bool CS, bool FAST = false>
__device__ __forceinline__ void explicit_tile(const ExplicitParams &p, int tile, ExplicitSmem<CS, R> &sm)
{
    real (*hvblk)[64] = sm.hv;
    real (*chvblk)[64] = sm.chv;
    real (*dystg)[64] = sm.dy;
    const int lane = threadIdx.x;
    const int strip = tile % p.nstrips;
    const int band = tile / p.nstrips;
    const int X = p.nx + 4;
    const int x0 = 2 + strip * kStripCols;
    const int x = x0 - 2 + lane;
    const int xc = x < X ? x : X - 1;               // clamped column for loads
    const bool out_lane = lane >= 2 && lane < 62 && x < p.nx + 2;
    const int ys = 2 + band * R;
    const int ye = min(ys + R, p.ny + 2);
    const int P = p.P, PT = p.PT;

    // transposed hv: each lane stages rows [ys-2, ye+2) of its own column in pieces of E
    // elements (16 bytes; 8 bytes for 2-row bands, whose ys-2 is only even); ys-2 is a
    // multiple of E and the column pitch leaves >= 16 elements of padding
    using Piece = BandPiece<R>;
    constexpr int E = Piece::n;
    {
        const int nrows = ye - ys + 4;
        const Piece *src = reinterpret_cast<const Piece *>(p.hvT + (size_t)xc * PT + (ys - 2));
        for (int r = 0; r < nrows; r += E) {
            const Piece v = src[r / E];
#pragma unroll
            for (int j = 0; j < E; ++j)
                if (r + j < R + 4) hvblk[r + j][lane] = v.v[j];
        }
        if (SCHEME == 2) {
            const Piece *csrc = reinterpret_cast<const Piece *>(p.ChvT + (size_t)xc * PT + (ys - 2));
            for (int r = 0; r < nrows; r += E) {
                const Piece v = csrc[r / E];
#pragma unroll
                for (int j = 0; j < E; ++j)
                    if (r + j < R + 4) chvblk[r + j][lane] = v.v[j];
            }
        }
    }
    __syncthreads();

    const real tt = p.two_theta, eps = p.epsilon, g = p.g, half_g = p.half_g;
    const real sl = p.sea_level;

    // register windows over rows; index 0 = y-2 ... 4 = y+2
    real w_[5], BA_[5], hu_[4], hv_[4], BY_[4];
    // fast variant (coefficient sea level == still-water level): the depths max(0, sl - BA)
    // of rows y-2 .. y+2, formed once per row and shared by K1, K2/72, the gradients and Psi
    real D_[5];
    // window slot j holds row (y - 2 + j) for w/BA; (y - 1 + j) for hu/hv/BY
    // 32-bit element offsets: planes are far below 4 G elements, and a uniform base with a
    // 32-bit offset maps to the scalar-base addressing mode (no 64-bit address arithmetic)
    auto ldrow = [&](const real *base, int row) { return ld32(base, (unsigned)row * (unsigned)P + (unsigned)xc); };

    // prologue: rows ys-2 .. ys+1
    for (int j = 0; j < 4; ++j) {
        const int row = ys - 2 + j;
        w_[j] = ldrow(p.w, row);
        BA_[j] = ldrow(p.BA, row);
        if (FAST) D_[j] = fmaxr(0.f, sl - BA_[j]);
    }
    for (int j = 0; j < 3; ++j) {
        const int row = ys - 1 + j;
        hu_[j] = ldrow(p.hu, row);
        hv_[j] = hvblk[row - (ys - 2)][lane];
        BY_[j] = ldrow(p.BY, row);
    }
    const real hu_m2 = ldrow(p.hu, ys - 2);
    const real hv_m2 = hvblk[0][lane];
    const real BY_m2 = ldrow(p.BY, ys - 2);

    // faces of row ys-1 (need rows ys-2..ys) -> keep N; faces of row ys -> S, N
    Face Nprev, Scur, Ncur;
    {
        Face lo, hi;
        faces_1d<FAST>(tt, eps, w_[0], w_[1], w_[2], hu_m2, hu_[0], hu_[1], hv_m2, hv_[0], hv_[1], BY_m2, BY_[0], lo, hi);
        Nprev = hi;
        faces_1d<FAST>(tt, eps, w_[1], w_[2], w_[3], hu_[0], hu_[1], hu_[2], hv_[0], hv_[1], hv_[2], BY_[0], BY_[1], lo, hi);
        Scur = lo;
        Ncur = hi;
    }
    real yf_s0, yf_s1, yf_s2;   // yflux at row y-1
    flux_y<FAST>(g, half_g, Nprev, Scur, yf_s0, yf_s1, yf_s2);

    // a lane's new values were not all finite (numeric sentinel)
    bool bad = false;

    // software pipelining: the loads a row needs are issued one row ahead
    const int ylast = p.ny + 3;
    real nx_w = ldrow(p.w, ys + 2), nx_hu = ldrow(p.hu, ys + 2), nx_BY = ldrow(p.BY, ys + 2);
    real nx_BX = ldrow(p.BX, ys), nx_BA = ldrow(p.BA, ys + 2);
    CellIn nx_c = load_cell<SCHEME, COMP>(p, (unsigned)ys * (unsigned)P + (unsigned)xc);

    auto row = [&](const int y) __attribute__((always_inline)) {
        // ---- rows y+2 (w, hu, hv, BY) and y (BX, BA, per-cell inputs), prefetched
        w_[4] = nx_w;
        hu_[3] = nx_hu;
        hv_[3] = hvblk[y + 2 - (ys - 2)][lane];
        BY_[3] = nx_BY;
        const real BXc = nx_BX;
        BA_[4] = nx_BA;
        if (FAST) D_[4] = fmaxr(0.f, sl - BA_[4]);
        const real BAc = BA_[2];
        const CellIn cin = nx_c;
        {
            const int yn = min(y + 3, ylast), yc = min(y + 1, ye - 1);
            nx_w = ldrow(p.w, yn);
            nx_hu = ldrow(p.hu, yn);
            nx_BY = ldrow(p.BY, yn);
            nx_BX = ldrow(p.BX, yc);
            nx_BA = ldrow(p.BA, yn);
            nx_c = load_cell<SCHEME, COMP>(p, (unsigned)yc * (unsigned)P + (unsigned)xc);
        }

        // ---- faces of row y+1 along y: S face used now, N face kept
        Face S1, N1;
        faces_1d<FAST>(tt, eps, w_[2], w_[3], w_[4], hu_[1], hu_[2], hu_[3], hv_[1], hv_[2], hv_[3], BY_[1], BY_[2], S1, N1);
        real yf0, yf1, yf2;   // yflux at row y
        flux_y<FAST>(g, half_g, Ncur, S1, yf0, yf1, yf2);

        // ---- faces of row y along x
        const real wc = w_[2], huc = hu_[1], hvc = hv_[1];
        const real wl = from_left(wc), wr = from_right(wc);
        const real hul = from_left(huc), hur = from_right(huc);
        const real hvl = from_left(hvc), hvr = from_right(hvc);
        const real BXw = from_left(BXc);
        Face Wf, Ef;
        faces_1d<FAST>(tt, eps, wl, wc, wr, hul, huc, hur, hvl, hvc, hvr, BXw, BXc, Wf, Ef);
        // Pass1 stores u = (1/h)*hu and v = (1/h)*hv; along x the normal velocity is u.
        Face Wn;
        Wn.h = from_right(Wf.h);
        Wn.u = from_right(Wf.u);
        Wn.v = from_right(Wf.v);
        real xf0, xf1, xf2;   // xflux at (x, y)
        flux_x<FAST>(g, half_g, Ef, Wn, xf0, xf1, xf2);
        const real xw0 = from_left(xf0), xw1 = from_left(xf1), xw2 = from_left(xf2);

        // ---- Pass3 at (x, y), compute.hlsl:616-812
        const real wll = from_left(wl), wrr = from_right(wr);
        const real wul = from_left(w_[3]), wur = from_right(w_[3]);
        const real wdl = from_left(w_[1]), wdr = from_right(w_[1]);
        const real BYs = BY_[0];
        const real BYc = BY_[1];
        // the bed at x-2 .. x+2 for the line coefficients and the depth gradients
        // (fast variant: the depths at x-2 .. x+2 instead)
        const real bal = from_left(FAST ? D_[2] : BAc), bar = from_right(FAST ? D_[2] : BAc);
        const real ball = from_left(bal), barr = from_right(bar);

        if (out_lane) {
            const unsigned i = (unsigned)y * (unsigned)P + (unsigned)x;
            const real h = fmaxr(0.f, wc - BAc);
            const real d_here = FAST ? D_[2] : fmaxr(0.f, sl - BAc);
            const real d2_here = d_here * d_here;
            const real d3_here = d2_here * d_here;

            const real eta_here = wc - sl;
            const real eta_left = wl - sl, eta_right = wr - sl;
            const real eta_down = w_[1] - sl, eta_up = w_[3] - sl;
            const real eta_left_left = wll - sl, eta_right_right = wrr - sl;
            const real eta_down_down = w_[0] - sl, eta_up_up = w_[4] - sl;
            const real eta_up_left = wul - sl, eta_up_right = wur - sl;
            const real eta_down_left = wdl - sl, eta_down_right = wdr - sl;

            const real odx = p.one_over_dx, ody = p.one_over_dy;
            const real one_over_d2x = odx * odx;
            const real one_over_d3x = one_over_d2x * odx;
            const real one_over_d2y = ody * ody;
            const real one_over_d3y = one_over_d2y * ody;

            // K1, K2/72 of the x- and y-line through this cell (the same operations as the
            // line solves) and the fourth-order depth gradients, formed from the bed
            // rather than read from stored planes
            const real cs = p.coef_sl;
            real Kx, Qx, Ky, Qy;
            if (FAST) {
                line_kq<FAST>(p.kcoef, p.dx2, p.rdx2, ball, bal, D_[2], bar, barr, Kx, Qx);
                line_kq<FAST>(p.kcoef, p.dy2, p.rdy2, D_[0], D_[1], D_[2], D_[3], D_[4], Ky, Qy);
            } else {
                line_kq<FAST>(p.kcoef, p.dx2, p.rdx2, fmaxr(0.f, cs - ball), fmaxr(0.f, cs - bal),
                              fmaxr(0.f, cs - BAc), fmaxr(0.f, cs - bar), fmaxr(0.f, cs - barr), Kx, Qx);
                line_kq<FAST>(p.kcoef, p.dy2, p.rdy2, fmaxr(0.f, cs - BA_[0]), fmaxr(0.f, cs - BA_[1]),
                              fmaxr(0.f, cs - BA_[2]), fmaxr(0.f, cs - BA_[3]), fmaxr(0.f, cs - BA_[4]), Ky, Qy);
            }
            const real d_l = FAST ? bal : fmaxr(0.f, sl - bal), d_r = FAST ? bar : fmaxr(0.f, sl - bar);
            const real d_ll = FAST ? ball : fmaxr(0.f, sl - ball), d_rr = FAST ? barr : fmaxr(0.f, sl - barr);
            const real d_d = FAST ? D_[1] : fmaxr(0.f, sl - BA_[1]), d_u = FAST ? D_[3] : fmaxr(0.f, sl - BA_[3]);
            const real d_dd = FAST ? D_[0] : fmaxr(0.f, sl - BA_[0]), d_uu = FAST ? D_[4] : fmaxr(0.f, sl - BA_[4]);
            const real r12 = real(1) / real(12);
            const real dd_by_dx = div_k<FAST>((-d_rr + 8.f * d_r - 8.f * d_l + d_ll) * odx, real(12), r12);
            const real dd_by_dy = div_k<FAST>((-d_uu + 8.f * d_u - 8.f * d_d + d_dd) * ody, real(12), r12);
            const real eta_by_dx_dy = (0.25f * odx * ody * (eta_up_right - eta_down_right - eta_up_left + eta_down_left));
            const real eta_by_dx_dx = one_over_d2x * (eta_right - 2.f * eta_here + eta_left);
            const real eta_by_dy_dy = one_over_d2y * (eta_up - 2.f * eta_here + eta_down);

            const real Bg = p.Bcoef_g;
            const real Psi1x = Bg * (d3_here) *
                ((eta_right_right - 2.f * eta_right + 2.f * eta_left - eta_left_left) * (0.5f * one_over_d3x)
                 + (eta_up_right - eta_up_left - 2.f * eta_right + 2.f * eta_left + eta_down_right - eta_down_left)
                   * (0.5f * odx * one_over_d2y));
            const real Psi2x = Bg * (d2_here) * (dd_by_dx * (2.f * eta_by_dx_dx + eta_by_dy_dy) + dd_by_dy * eta_by_dx_dy);
            const real Psi1y = Bg * (d3_here) *
                ((eta_up_up - 2.f * eta_up + 2.f * eta_down - eta_down_down) * (0.5f * one_over_d3y)
                 + (eta_up_right + eta_up_left - 2.f * eta_up + 2.f * eta_down - eta_down_right - eta_down_left)
                   * (0.5f * ody * one_over_d2x));
            const real Psi2y = Bg * (d2_here) * (dd_by_dy * (eta_by_dx_dx + 2.f * eta_by_dy_dy) + dd_by_dx * eta_by_dx_dy);

            // FrictionCalc, compute.hlsl:544-553 (pow(h,4) as (h*h)*(h*h))
            real friction_ = 0.f;
            if (p.friction != 0.f) {
                real f = p.friction;
                const real dbh = divide_by_h(h, eps);
                if (p.is_manning == 1) f = g * p.friction * p.friction * POW(FABS(dbh), (real)0.3333f);
                friction_ = f * SQRT(huc * huc + hvc * hvc) * dbh * dbh;
            }
            const real src_g = -p.g_over_dx * h * (BXc - BXw) - huc * friction_ + (Psi1x + Psi2x);
            const real src_b = -p.g_over_dy * h * (BYc - BYs) - hvc * friction_ + (Psi1y + Psi2y);

            const real dr = (xw0 - xf0) * odx + (yf_s0 - yf0) * ody + 0.f;
            const real dg = (xw1 - xf1) * odx + (yf_s1 - yf1) * ody + src_g;
            const real db = (xw2 - xf2) * odx + (yf_s2 - yf2) * ody + src_b;

            // F*, G* (compute.hlsl:726-734); the "0 *" cross term of the reference adds zero
            const real hv_up = hv_[2], hv_down = hv_[0];
            const real hu_up = hu_[2], hu_down = hu_[0];
            const real F_star = (real(1) / real(6) * d_here *
                (dd_by_dx * ((ody / 2.f) * (hv_up - hv_down)) + (dd_by_dy * ((odx / 2.f) * (hvr - hvl))))) + 0.f;
            const real G_star = (real(1) / real(6) * d_here *
                (dd_by_dx * ((ody / 2.f) * (hu_up - hu_down)) + (dd_by_dy * ((odx / 2.f) * (hur - hul))))) + 0.f;

            // level-0 coefficients of the x-line (row y) and y-line (column x) through this cell
            real Ax, Bx, Cx, Ay, By, Cy;
            line_abc(Kx, Qx, x == 2, x == p.nx + 1, Ax, Bx, Cx);
            line_abc(Ky, Qy, y == 2, y == p.ny + 1, Ay, By, Cy);

            const real dt = p.dt;
            real wOut, huOut, hvOut;
            // increments of w and the line right-hand sides without the level-0 A, B, C
            // terms: the compensated update solves for the increment of hu and hv
            real dw, dhu, dhv;
            if (SCHEME == 0) {
                dw = dt * dr;
                dhu = dt * dg;
                dhv = dt * db;
                wOut = wc + dw;
                huOut = Ax * hul + Bx * huc + Cx * hur + dhu;
                hvOut = Ay * hv_down + By * hvc + Cy * hv_up + dhv;
            } else {
                const real o0 = cin.o0, o1 = cin.o1, o2 = cin.o2;
                const real oo0 = cin.oo0, oo1 = cin.oo1, oo2 = cin.oo2;
                const real Fo = cin.Fo, Go_ = cin.Go;
                const real Foo = cin.Foo, Goo_ = cin.Goo;
                if (SCHEME == 1) {
                    dw = dt / 12.f * (23.f * dr - 16.f * o0 + 5.f * oo0);
                    const real u1 = dt / 12.f * (23.f * dg - 16.f * o1 + 5.f * oo1), u2 = 2.f * F_star - 3.f * Fo + Foo;
                    const real v1 = dt / 12.f * (23.f * db - 16.f * o2 + 5.f * oo2), v2 = 2.f * G_star - 3.f * Go_ + Goo_;
                    wOut = wc + dw;
                    huOut = Ax * hul + Bx * huc + Cx * hur + u1 + u2;
                    hvOut = Ay * hv_down + By * hvc + Cy * hv_up + v1 + v2;
                    dhu = u1 + u2;
                    dhv = v1 + v2;
                } else if (SCHEME == 3) {
                    const real *c = p.der;
                    const real dF0 = (c[0] * F_star) + (c[1] * Fo) + (c[2] * Foo);
                    const real dF1 = (c[3] * F_star) + (c[4] * Fo) + (c[5] * Foo);
                    const real dF2 = (c[6] * F_star) + (c[7] * Fo) + (c[8] * Foo);
                    const real dG0 = (c[0] * G_star) + (c[1] * Go_) + (c[2] * Goo_);
                    const real dG1 = (c[3] * G_star) + (c[4] * Go_) + (c[5] * Goo_);
                    const real dG2 = (c[6] * G_star) + (c[7] * Go_) + (c[8] * Goo_);
                    const real a0 = p.ab[0], a1 = p.ab[1], a2 = p.ab[2];
                    dw = dt / 6.f * (a0 * dr + a1 * o0 + a2 * oo0);
                    const real u1 = dt / 6.f * (a0 * dg + a1 * o1 + a2 * oo1), u2 = dt / 6.f * (a0 * dF0 + a1 * dF1 + a2 * dF2);
                    const real v1 = dt / 6.f * (a0 * db + a1 * o2 + a2 * oo2), v2 = dt / 6.f * (a0 * dG0 + a1 * dG1 + a2 * dG2);
                    wOut = wc + dw;
                    huOut = Ax * hul + Bx * huc + Cx * hur + u1 + u2;
                    hvOut = Ay * hv_down + By * hvc + Cy * hv_up + v1 + v2;
                    dhu = u1 + u2;
                    dhv = v1 + v2;
                } else {
                    const real cw = ld32(p.Cw, i);
                    const real chu = ld32(p.Chu, i), chul = ld32(p.Chu, i - 1u), chur = ld32(p.Chu, i + 1u);
                    const real chv = chvblk[y - (ys - 2)][lane];
                    const real chvd = chvblk[y - 1 - (ys - 2)][lane];
                    const real chvu = chvblk[y + 1 - (ys - 2)][lane];
                    const Quad pq = ldq(p.Hp, i);
                    const real pr0 = pq.v[0], pr1 = pq.v[1], pr2 = pq.v[2], Fp = pq.v[3];
                    const real Gp_ = ld32(p.Hp1, i);
                    wOut = cw + dt / 24.f * (9.f * dr + 19.f * pr0 - 5.f * o0 + 1.f * oo0);
                    huOut = Ax * chul + Bx * chu + Cx * chur + dt / 24.f * (9.f * dg + 19.f * pr1 - 5.f * o1 + 1.f * oo1) +
                            F_star - Fp;
                    hvOut = Ay * chvd + By * chv + Cy * chvu + dt / 24.f * (9.f * db + 19.f * pr2 - 5.f * o2 + 1.f * oo2) +
                            G_star - Gp_;
                    dw = dhu = dhv = 0.f;   // the compensated update does not take corrector steps
                }
            }
            if (COMP) {
                // (w_hi, w_lo) += dw as a two-real sum; the lines solve for the increments
                real lo;
                two_sum(wc, dw + cin.wlo, wOut, lo);
                st32(p.w_lo_new, i, lo);
                huOut = dhu;
                hvOut = dhv;
            }
            bad = bad || !(__builtin_isfinite(wOut) && __builtin_isfinite(huOut) && __builtin_isfinite(hvOut));
            st32(p.w_new, i, wOut);
            st32(p.Dx, i - 2u, huOut);
            dystg[y - ys][lane] = hvOut;
            stq(p.Hs, i, Quad{{dr, dg, db, F_star}});
            st32(p.Hs1, i, G_star);
        }
        // ---- shift windows
        w_[0] = w_[1]; w_[1] = w_[2]; w_[2] = w_[3]; w_[3] = w_[4];
        BA_[0] = BA_[1]; BA_[1] = BA_[2]; BA_[2] = BA_[3]; BA_[3] = BA_[4];
        if (FAST) { D_[0] = D_[1]; D_[1] = D_[2]; D_[2] = D_[3]; D_[3] = D_[4]; }
        hu_[0] = hu_[1]; hu_[1] = hu_[2]; hu_[2] = hu_[3];
        hv_[0] = hv_[1]; hv_[1] = hv_[2]; hv_[2] = hv_[3];
        BY_[0] = BY_[1]; BY_[1] = BY_[2]; BY_[2] = BY_[3];
        Ncur = N1;
        yf_s0 = yf0; yf_s1 = yf1; yf_s2 = yf2;
    };
    // fast variant: two rows per trip, so that the window shifts between them become register
    // renames (the plain variant's longer arithmetic would need more than 128 VGPRs for it).
    // An odd row runs first, so no value of the prologue stays live past the paired loop.
    int y = ys;
    if (FAST) {
        if ((ye - ys) & 1) row(y++);
        for (; y < ye; y += 2) {
            row(y);
            row(y + 1);
        }
    } else {
        for (; y < ye; ++y) row(y);
    }
    // one atomic per wave that saw a non-finite value
    if (__ballot(bad) != 0ull && lane == 0 && p.status) atomicOr(p.status, 1);
    // column x, rows [ys, ye) at offset y-2: each lane writes its column in pieces of E
    if (out_lane) {
        real *dst = p.DyT + (size_t)x * PT + (ys - 2);
        const int nr = ye - ys;
        if (nr == R) {
            Piece *dv = reinterpret_cast<Piece *>(dst);
            for (int r = 0; r < R; r += E) {
                Piece v;
#pragma unroll
                for (int j = 0; j < E; ++j) v.v[j] = dystg[r + j][lane];
                dv[r / E] = v;
            }
        } else {
            for (int r = 0; r < nr; ++r) dst[r] = dystg[r][lane];
        }
    }
}

// FP32: 4 waves per SIMD (at most 128 VGPRs), the stencil's occupancy; the FP64 variants
// need more registers than that and take what they need instead of spilling
#ifdef CEL_FP64
#define CEL_EXPLICIT_BOUNDS __launch_bounds__(64)
#else
#define CEL_EXPLICIT_BOUNDS __launch_bounds__(64, 4)
#endif
template <int SCHEME, bool COMP, int R, bool FAST>
__global__ CEL_EXPLICIT_BOUNDS void explicit_step(const ExplicitParams p)
{
    __shared__ ExplicitSmem<SCHEME == 2, R> sm;
    // workgroups are dealt round-robin to the 8 XCDs; give each XCD a contiguous run of
    // tiles (strip-fastest) so neighbouring strips and bands share that XCD's L2
    const int ntiles = p.nstrips * p.nbands;
    const int per_xcd = (ntiles + 7) / 8;
    const int tile = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    if (tile >= ntiles) return;
    explicit_tile<SCHEME, COMP, R, SCHEME == 2, FAST>(p, tile, sm);
}

// ------------------------------------------------------------------ line solves


// Irregular waves: one wave per (side, strip index, line cell) computes that cell's
// pairwise-tree sum of the component texels (ColumnSumReduce/RowSumReduce order).
__device__ __forceinline__ void irregular_line(const IrregularSide &S, int strip, int line, real t)
{
    const int lane = threadIdx.x & 63;
    const int cell = strip * S.nlines + line;
    real r0 = 0.f, r1 = 0.f, r2 = 0.f;
    const bool ghost = line < 2 || line >= S.nlines - 2 || S.zero_line;
    if (!ghost) {
        const bool wet = S.wet[cell] != 0;
        const int A = S.nleaf_pow2;                 // power of two, <= 1024
        const int lanes_used = A > 64 ? 64 : A;
        const int nblk = A > 64 ? A / 64 : 1;       // blocks of 64 leaves
        real b0 = 0.f, b1 = 0.f, b2 = 0.f;         // lane j keeps the sum of block j
        for (int blk = 0; blk < nblk; ++blk) {
            const int wv = blk * 64 + lane;
            real v0 = 0.f, v1 = 0.f, v2 = 0.f;
            if (lane < lanes_used) {
                real e = 0.f, hu = 0.f, hv = 0.f;
                if (wet && wv < S.nleaf_active) {
                    const real *tb = S.tab + ((size_t)cell * S.nw + wv) * 4;
                    const real arg = S.omega[wv] * t - tb[0] - tb[1] + S.phase[wv];
                    e = S.amp[wv] * SIN(arg);
                    hu = tb[2] * e;
                    hv = tb[3] * e;
                }
                v0 = e + S.sl; v1 = hu; v2 = hv;
            }
            // pairwise tree over the block's leaves (ColumnSumReduce order)
            for (int off = 1; off < lanes_used; off <<= 1) {
                const real o0 = __shfl_down(v0, off, 64), o1 = __shfl_down(v1, off, 64), o2 = __shfl_down(v2, off, 64);
                if ((lane & (2 * off - 1)) == 0) { v0 = v0 + o0; v1 = v1 + o1; v2 = v2 + o2; }
            }
            v0 = __shfl(v0, 0, 64); v1 = __shfl(v1, 0, 64); v2 = __shfl(v2, 0, 64);
            if (lane == blk) { b0 = v0; b1 = v1; b2 = v2; }
        }
        // upper levels of the tree over the block sums
        for (int off = 1; off < nblk; off <<= 1) {
            const real o0 = __shfl_down(b0, off, 64), o1 = __shfl_down(b1, off, 64), o2 = __shfl_down(b2, off, 64);
            if ((lane & (2 * off - 1)) == 0) { b0 = b0 + o0; b1 = b1 + o1; b2 = b2 + o2; }
        }
        // subtrees beyond the active block hold only inactive texels (sl, 0, 0)
        real c0 = S.sl, c1 = 0.f, c2 = 0.f;
        for (int w = 1; w < A; w <<= 1) { c0 = c0 + c0; c1 = c1 + c1; c2 = c2 + c2; }
        for (int l = 0; l < S.levels_above; ++l) {
            b0 = b0 + c0; b1 = b1 + c1; b2 = b2 + c2;
            c0 = c0 + c0; c1 = c1 + c1; c2 = c2 + c2;
        }
        r0 = b0; r1 = b1; r2 = b2;
    }
    if (lane == 0) {
        S.out[(size_t)cell * 3 + 0] = r0;
        S.out[(size_t)cell * 3 + 1] = r1;
        S.out[(size_t)cell * 3 + 2] = r2;
    }
}


// One wave per line, each lane holding a contiguous chunk of C unknowns in registers.
// The eliminations are those of the reference's cyclic reduction (reduced equation m of
// level k built from unknowns 2m, 2m+1, 2m+2 of level k-1; back-substitution level by
// level), so the arithmetic, and therefore every rounding, is unchanged; only the
// placement differs.  Level k keeps the unknowns at positions 2^k (m+1) - 1: for
// 2^k <= C they lie inside one lane's chunk, above that at the chunk's last slot of
// lanes 2^(k-log2 C) (m+1) - 1.  Neighbours outside the chunk come from other lanes
// through cross-lane reads; positions >= n are the reference's (0, 1, 0 | 0) margin.
struct alignas(16) Eq { real a, b, c, d; };

__device__ __forceinline__ Eq shfl_eq(const Eq &e, int src)
{
    return Eq{__shfl(e.a, src, 64), __shfl(e.b, src, 64), __shfl(e.c, src, 64), __shfl(e.d, src, 64)};
}

__device__ __forceinline__ Eq reduce_eq(const Eq &L, const Eq &H, const Eq &R)
{
    const real k1 = H.a / L.b;
    const real k2 = H.c / R.b;
    Eq o;
    o.a = -L.a * k1;
    o.b = H.b - L.c * k1 - R.a * k2;
    o.c = -R.c * k2;
    o.d = H.d - L.d * k1 - R.d * k2;
    return o;
}

template <int C, int W>
struct LineSmem {
    // one unknown per thread above the chunk levels, as four planes.  Level j works on slots
    // 2^j apart, dealt to consecutive lanes; slot i is kept at i + i/32, so those lanes fall
    // on distinct banks for every j (a plain index would put up to 32 of them on one bank)
    static constexpr int kPad = 64 * W + 2 * W;
    real ta[kPad], tb[kPad], tc[kPad], td[kPad];
    Eq xch[2][64 * W];                         // chunk-level neighbour exchange (double-buffered)
    __device__ static int at(int i) { return i + (i >> 5); }
    __device__ Eq top(int i) const { const int k = at(i); return Eq{ta[k], tb[k], tc[k], td[k]}; }
    __device__ void set_top(int i, const Eq &e)
    {
        const int k = at(i);
        ta[k] = e.a; tb[k] = e.b; tc[k] = e.c; td[k] = e.d;
    }
    __device__ real x(int i) const { return td[at(i)]; }
    __device__ void set_x(int i, real v) { td[at(i)] = v; }
};

// Loads C consecutive elements starting at a 16-byte aligned element (C a multiple of
// kVec) or singly.
template <int C>
__device__ __forceinline__ void load_chunk(const real *src, int base, int n, real *v)
{
    if (C % kVec == 0 && base + C <= n) {
#pragma unroll
        for (int i = 0; i < C; i += kVec) {
            const Vec f = *reinterpret_cast<const Vec *>(src + base + i);
#pragma unroll
            for (int j = 0; j < kVec; ++j) v[i + j] = f.v[j];
        }
    } else {
#pragma unroll
        for (int i = 0; i < C; ++i) v[i] = base + i < n ? src[base + i] : 0.f;
    }
}

// The bed along one line and the constants of its level-0 coefficients.
struct LineBed {
    const real *B;        // bed at ghost-inclusive position 0 of the line (16-byte aligned)
    real kcoef, d2, rd2, cs;
    bool fast;            // divisions through div_k
};

// The C + 4 bed values [base, base + C + 4) of a chunk (16-byte loads when the chunk is
// whole; positions past the line's n + 4 read as 0 and are never used).
template <int C>
__device__ __forceinline__ void load_bed(const real *src, int base, int n, real *v)
{
    if ((C + 4) % kVec == 0 && C % kVec == 0 && base + C <= n) {
#pragma unroll
        for (int i = 0; i < C + 4; i += kVec) {
            const Vec f = *reinterpret_cast<const Vec *>(src + base + i);
#pragma unroll
            for (int j = 0; j < kVec; ++j) v[i + j] = f.v[j];
        }
    } else {
#pragma unroll
        for (int i = 0; i < C + 4; ++i) v[i] = base + i < n + 4 ? src[base + i] : 0.f;
    }
}

// Level-0 equations of a chunk: A, B, C from the depths at offsets -2..+2 (line_kq,
// line_abc: the same operations as the explicit kernel), right-hand sides from dv.
template <int C, bool FAST, bool FULL>
__device__ __forceinline__ void chunk_equations(const LineBed &lb, const real *dd, const real *dv, int base, int n,
                                                Eq *q)
{
#pragma unroll
    for (int i = 0; i < C; ++i) {
        const int el = base + i;
        if (FULL || el < n) {
            real K1, Q;
            line_kq<FAST>(lb.kcoef, lb.d2, lb.rd2, dd[i], dd[i + 1], dd[i + 2], dd[i + 3], dd[i + 4], K1, Q);
            line_abc(K1, Q, el == 0, el == n - 1, q[i].a, q[i].b, q[i].c);
            q[i].d = dv[i];
        } else {
            q[i] = Eq{0.f, 1.f, 0.f, 0.f};
        }
    }
}

// Line of n unknowns solved by a block of W waves; thread t owns positions [tC, tC+C).
// D points at element 0 of the line (16-byte aligned), lb.B at the line's ghost-inclusive
// position 0 (element e's bed stencil is lb.B[e .. e+4]); out at element 0 of the
// ghost-inclusive output line + 2 (8-byte aligned).
// comp: the solution is an increment; out/out_lo receive the two-real sum of the previous
// state (prev, prev_lo, same layout as out) and the increment.
// FULL: n == 64 W C, so every position of the block lies on the line and the tests against
// the line's end drop out (the same eliminations).
template <int C, int W, bool FULL>
__device__ __forceinline__ void cr_line_block(int n, const real *D, const LineBed &lb, real *out,
                                              LineSmem<C, W> &sm, bool comp, const real *prev, const real *prev_lo,
                                              real *out_lo)
{
    constexpr int NT = 64 * W;
    constexpr int K0 = C == 1 ? 0 : C == 2 ? 1 : C == 4 ? 2 : C == 8 ? 3 : C == 16 ? 4 : 5;
    constexpr int KT = W == 1 ? 6 : W == 2 ? 7 : W == 4 ? 8 : W == 8 ? 9 : 10;     // log2(NT)
    static_assert(W == 1 || W == 2 || W == 4 || W == 8 || W == 16, "waves per line");
    const int t = threadIdx.x;
    const int base = t * C;
    Eq q[C];
    {
        real dv[C], dd[C + 4];
        load_chunk<C>(D, base, FULL ? base + C : n, dv);
        load_bed<C>(lb.B, base, FULL ? base + C : n, dd);
#pragma unroll
        for (int i = 0; i < C + 4; ++i) dd[i] = fmaxr(0.f, lb.cs - dd[i]);
        if (lb.fast) chunk_equations<C, true, FULL>(lb, dd, dv, base, n, q);
        else chunk_equations<C, false, FULL>(lb, dd, dv, base, n, q);
    }
    const Eq margin{0.f, 1.f, 0.f, 0.f};
    // ---- reduction inside the chunk: levels 1..K0
#pragma unroll
    for (int k = 1; k <= K0; ++k) {
        const int s = 1 << (k - 1);
        // the chunk's last level-k unknown needs the next thread's level-(k-1) unknown at s-1
        sm.xch[k & 1][t] = q[s - 1];
        __syncthreads();
        Eq nr = sm.xch[k & 1][t + 1 < NT ? t + 1 : t];   // one 16-byte read, then the margin
        if (t + 1 >= NT) nr = margin;
        if (!FULL && base + C - 1 + s >= n) nr = margin;
#pragma unroll
        for (int j = 0; j < (C >> k); ++j) {
            const int pp = ((j + 1) << k) - 1;
            if (!FULL && base + pp >= n) continue;
            const Eq &L = q[pp - s];
            const Eq R = (pp + s < C) ? ((FULL || base + pp + s < n) ? q[pp + s] : margin) : nr;
            q[pp] = reduce_eq(L, q[pp], R);
        }
    }
    // ---- reduction across threads: levels K0+1 .. K0+KT, in LDS.  Slot i holds the chunk
    // top of thread i (line position iC + C - 1).  Level j updates slots (m+1) 2^j - 1; they
    // are dealt to threads m = 0, 1, ... so that the waves with no slot skip the level.
    sm.set_top(t, q[C - 1]);
    __syncthreads();
#pragma unroll
    for (int j = 1; j <= KT; ++j) {
        const int s = 1 << (j - 1);
        const int i = ((t + 1) << j) - 1;
        if (i < NT && (FULL || i * C + C - 1 < n)) {
            const Eq L = sm.top(i - s);
            const Eq R = (i + s < NT && (FULL || (i + s) * C + C - 1 < n)) ? sm.top(i + s) : margin;
            sm.set_top(i, reduce_eq(L, sm.top(i), R));
        }
        __syncthreads();
    }
    // ---- back-substitution across threads: levels K0+KT .. K0 (solution in .d); level j
    // solves slots (2m+1) 2^j - 1, dealt to threads m in the same way
#pragma unroll
    for (int j = KT; j >= 0; --j) {
        const int s = 1 << j;
        const int i = ((2 * t + 1) << j) - 1;
        if (i < NT && (FULL || i * C + C - 1 < n)) {
            const real xl = i - s >= 0 ? sm.x(i - s) : 0.f;
            const real xr = (i + s < NT && (FULL || (i + s) * C + C - 1 < n)) ? sm.x(i + s) : 0.f;
            const Eq e = sm.top(i);
            sm.set_x(i, (e.d - e.a * xl - e.c * xr) / e.b);
        }
        __syncthreads();
    }
    q[C - 1].d = sm.x(t);
    const real prev_last = t > 0 ? sm.x(t - 1) : 0.f;
    // ---- back-substitution inside the chunk: levels K0-1 .. 0
#pragma unroll
    for (int k = K0 - 1; k >= 0; --k) {
        const int s = 1 << k;
#pragma unroll
        for (int qq = 0; qq < (C >> (k + 1)); ++qq) {
            const int pp = s * (2 * qq + 1) - 1;
            if (!FULL && base + pp >= n) continue;
            const real xl = (pp - s >= 0) ? q[pp - s].d : prev_last;
            const real xr = (FULL || base + pp + s < n) ? q[pp + s].d : 0.f;
            q[pp].d = (q[pp].d - q[pp].a * xl - q[pp].c * xr) / q[pp].b;
        }
    }
    if (comp && C % 2 == 0 && (FULL || base + C <= n)) {
#pragma unroll
        for (int i = 0; i < C; i += 2) {
            const Pair pv = *reinterpret_cast<const Pair *>(prev + base + i);
            const Pair pl = *reinterpret_cast<const Pair *>(prev_lo + base + i);
            real h0, l0, h1, l1;
            two_sum(pv.x, q[i].d + pl.x, h0, l0);
            two_sum(pv.y, q[i + 1].d + pl.y, h1, l1);
            *reinterpret_cast<Pair *>(out + base + i) = Pair{h0, h1};
            *reinterpret_cast<Pair *>(out_lo + base + i) = Pair{l0, l1};
        }
    } else if (comp) {
#pragma unroll
        for (int i = 0; i < C; ++i)
            if (base + i < n) {
                real hi, lo;
                two_sum(prev[base + i], q[i].d + prev_lo[base + i], hi, lo);
                out[base + i] = hi;
                out_lo[base + i] = lo;
            }
    } else if (C % 2 == 0 && (FULL || base + C <= n)) {
#pragma unroll
        for (int i = 0; i < C; i += 2) *reinterpret_cast<Pair *>(out + base + i) = Pair{q[i].d, q[i + 1].d};
    } else {
#pragma unroll
        for (int i = 0; i < C; ++i)
            if (base + i < n) out[base + i] = q[i].d;
    }
}

__device__ __forceinline__ LineBed line_bed(const LineParams &p, bool xdir, int line)
{
    LineBed lb;
    lb.B = xdir ? p.BA + (size_t)line * p.P : p.BAT + (size_t)line * p.PT;
    lb.kcoef = p.kcoef;
    lb.d2 = xdir ? p.dx2 : p.dy2;
    lb.rd2 = xdir ? p.rdx2 : p.rdy2;
    lb.cs = p.coef_sl;
    lb.fast = p.fast != 0;
    return lb;
}

// One block: an x-line, a y-line, or W irregular-wave lines.  (The body stays in the
// kernel: routing the by-value parameter blocks through a device function makes the
// compiler copy them to scratch.)
template <int C, int W, bool COMP>
__global__ __launch_bounds__(64 * W) void solve_lines_block(const LineParams p, const IrregularAll irr)
{
    __shared__ LineSmem<C, W> sm;
    const int bid = blockIdx.x;
    const int nlines = p.nlines_x + p.nlines_y;
    if (bid >= nlines) {
        // irregular-wave lines: W per block (one per wave); time t^n from the step clock
        const real t = irr.clock ? irr.clock[0] : irr.t;
        int idx = (bid - nlines) * W + (threadIdx.x >> 6);
        for (int s = 0; s < 4; ++s) {
            if (idx < irr.count[s]) {
                const IrregularSide &S = irr.side[s];
                irregular_line(S, idx / S.nlines, idx % S.nlines, t);
                return;
            }
            idx -= irr.count[s];
        }
        return;
    }
    const bool xdir = bid < p.nlines_x;
    const int line = xdir ? bid + 2 : (bid - p.nlines_x) + 2;
    const size_t o = xdir ? (size_t)line * p.P : (size_t)line * p.PT;
    const int n = xdir ? p.nx : p.ny;
    const real *D = xdir ? p.Dx : p.DyT;
    real *out = (xdir ? p.hu_out : p.hvT_out) + o + 2;
    const LineBed lb = line_bed(p, xdir, line);
    const real *prev = COMP ? (xdir ? p.hu_prev : p.hvT_prev) + o + 2 : nullptr;
    const real *prev_lo = COMP ? (xdir ? p.hu_prev_lo : p.hvT_prev_lo) + o + 2 : nullptr;
    real *out_lo = COMP ? (xdir ? p.hu_lo_out : p.hvT_lo_out) + o + 2 : nullptr;
    if (n == 64 * W * C) cr_line_block<C, W, true>(n, D + o, lb, out, sm, COMP, prev, prev_lo, out_lo);
    else cr_line_block<C, W, false>(n, D + o, lb, out, sm, COMP, prev, prev_lo, out_lo);
}

// ------------------------------------------------------------------ boundaries

struct Val { real w, hu, hv; };

__device__ __forceinline__ bool in_strip(const BoundaryParams &p, int side, int x, int y)
{
    const int wdt = p.side[side].width;
    switch (side) {
    case kWest: return x < wdt;
    case kEast: return x >= p.X - wdt;
    case kSouth: return y < wdt;
    default: return y >= p.Y - wdt;
    }
}

// position of a cell inside a side's strip table: (strip index) * line length + line position
__device__ __forceinline__ int strip_index(const BoundaryParams &p, int side, int x, int y)
{
    switch (side) {
    case kWest: return x * p.Y + y;
    case kEast: return (x - (p.X - p.side[kEast].width)) * p.Y + y;
    case kSouth: return y * p.X + x;
    default: return (y - (p.Y - p.side[kNorth].width)) * p.X + x;
    }
}

__device__ __forceinline__ Val pre_value(const BoundaryParams &p, int x, int y)
{
    Val v{0.f, 0.f, 0.f};
    if (x >= 2 && x < p.nx + 2 && y >= 2 && y < p.ny + 2) {
        v.w = p.w[(size_t)y * p.P + x];
        v.hu = p.hu[(size_t)y * p.P + x];
        v.hv = p.hvT[(size_t)x * p.PT + y];
    }
    return v;
}

// The boundary cell set: left band, right band, then bottom/top bands between them (x
// fastest inside each band: the row-major w and hu accesses of a wave share rows).
__device__ __forceinline__ void boundary_cell_xy(const BoundaryParams &p, int tid, int &x, int &y)
{
    int r = tid;
    const int nl = p.band_l * p.Y, nr = p.band_r * p.Y;
    const int mid = p.X - p.band_l - p.band_r;
    if (r < nl) { y = r / p.band_l; x = r % p.band_l; }
    else if ((r -= nl) < nr) { y = r / p.band_r; x = p.X - p.band_r + r % p.band_r; }
    else {
        r -= nr;
        const int nb = p.band_b * mid;
        if (r < nb) { y = r / mid; x = p.band_l + r % mid; }
        else { r -= nb; y = p.Y - p.band_t + r / mid; x = p.band_l + r % mid; }
    }
}

__device__ __forceinline__ bool in_boundary_set(const BoundaryParams &p, int x, int y)
{
    return x >= 0 && x < p.X && y >= 0 && y < p.Y &&
           (x < p.band_l || x >= p.X - p.band_r || y < p.band_b || y >= p.Y - p.band_t);
}

// What the four strips (applied in the order E, N, W, S, each reading the state the
// previous ones left) make of one cell: walking the order backwards from the cell, slot k
// records where side order[k] acts; a solid side moves the walk to the mirrored cell, a
// set-valued side (sine, irregular, time series) ends it.  The cell's final value is the
// recorded sides applied in order to the pre-boundary value of the walk's end (cx, cy),
// or to nothing when the walk ended at a set-valued side.
struct Chain {
    int kx[4], ky[4];
    bool applied[4];
    int cx, cy;
    bool have_source, mirrored;
};

__device__ __forceinline__ Chain boundary_chain(const BoundaryParams &p, int x, int y)
{
    Chain c;
    c.cx = x; c.cy = y;
    c.have_source = true; c.mirrored = false;
    bool stop = false;
#pragma unroll
    for (int k = 3; k >= 0; --k) {
        const int sd = k == 0 ? kEast : (k == 1 ? kNorth : (k == 2 ? kWest : kSouth));
        c.applied[k] = !stop && in_strip(p, sd, c.cx, c.cy);
        c.kx[k] = c.cx; c.ky[k] = c.cy;
        if (c.applied[k]) {
            const int ty = p.side[sd].type;
            if (ty == kSolid) {
                c.mirrored = true;
                if (sd == kWest) c.cx = 4 - c.cx;
                else if (sd == kSouth) c.cy = 4 - c.cy;
                else if (sd == kEast) c.cx = (2 * p.nx + 2) - c.cx;
                else c.cy = (2 * p.ny + 2) - c.cy;
            } else if (ty != kSponge) {
                c.have_source = false;
                stop = true;
            }
        }
    }
    return c;
}

__device__ __forceinline__ Val apply_chain(const BoundaryParams &p, const Chain &c, Val v)
{
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (!c.applied[k]) continue;
        const int sd = k == 0 ? kEast : (k == 1 ? kNorth : (k == 2 ? kWest : kSouth));
        const SideParams &S = p.side[sd];
        const int ox = c.kx[k], oy = c.ky[k];
        switch (S.type) {
        case kSolid:
            if (sd == kWest || sd == kEast) v.hu = -v.hu; else v.hv = -v.hv;
            break;
        case kSponge: {
            const real gamma = S.tab[strip_index(p, sd, ox, oy)];
            v.w = gamma * (v.w - S.sl) + S.sl;
            v.hu = gamma * v.hu;
            v.hv = gamma * v.hv;
            break;
        }
        case kSine: {
            const real *tb = S.tab + (size_t)strip_index(p, sd, ox, oy) * 5;
            real e = 0.f, hu = 0.f, hv = 0.f;
            if (tb[4] != 0.f) {
                const real t = p.clock ? p.clock[0] : p.t;
                const real arg = S.omega * t - tb[0] - tb[1] + 0.f;
                e = S.amp * SIN(arg);
                hu = tb[2] * e;
                hv = tb[3] * e;
            }
            v.w = e + S.sl; v.hu = hu; v.hv = hv;
            break;
        }
        case kUTS:
            if (p.clock) {
                const real *c = p.clock + 1 + 3 * sd;
                v.w = c[0] + S.sl; v.hu = c[1]; v.hv = c[2];
            } else {
                v.w = S.uts_eta + S.sl; v.hu = S.uts_hu; v.hv = S.uts_hv;
            }
            break;
        default: {   // irregular
            const real *q = S.irr + (size_t)strip_index(p, sd, ox, oy) * 3;
            v.w = q[0]; v.hu = q[1]; v.hv = q[2];
            break;
        }
        }
    }
    return v;
}

// apply_chain for walks that end at a source cell: only solid and sponge sides lie on them
__device__ __forceinline__ Val apply_chain_sourced(const BoundaryParams &p, const Chain &c, Val v)
{
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (!c.applied[k]) continue;
        const int sd = k == 0 ? kEast : (k == 1 ? kNorth : (k == 2 ? kWest : kSouth));
        const SideParams &S = p.side[sd];
        if (S.type == kSolid) {
            if (sd == kWest || sd == kEast) v.hu = -v.hu; else v.hv = -v.hv;
        } else {
            const real gamma = S.tab[strip_index(p, sd, c.kx[k], c.ky[k])];
            v.w = gamma * (v.w - S.sl) + S.sl;
            v.hu = gamma * v.hu;
            v.hv = gamma * v.hv;
        }
    }
    return v;
}

__device__ __forceinline__ void write_boundary(const BoundaryParams &p, int x, int y, const Val &v)
{
    p.w[(size_t)y * p.P + x] = v.w;
    p.hu[(size_t)y * p.P + x] = v.hu;
    p.hvT[(size_t)x * p.PT + y] = v.hv;
    if (p.w_lo) {   // compensated state: boundary values are set, not accumulated
        p.w_lo[(size_t)y * p.P + x] = 0.f;
        p.hu_lo[(size_t)y * p.P + x] = 0.f;
        p.hvT_lo[(size_t)x * p.PT + y] = 0.f;
    }
}

// Boundary cell tid of one scenario (tid < p.ncells).
//   pass 1: cells whose chain passes a solid side; pass 0: all others (run after pass 1,
//   so that the mirrored cells read their sources before those are rewritten in place).
//   pass 2: every cell in one launch.  A cell whose walk ends at an interior cell of the
//   boundary set (one this launch rewrites) is written by that cell's thread instead,
//   which reads its own pre-boundary value before writing anything.  Such cells are
//   mirror images of their source in x, y or both (at most one solid side per axis on the
//   walk, which the host guarantees by the grid size).
__device__ __forceinline__ void boundary_cell(const BoundaryParams &p, int tid, int pass)
{
    int x, y;
    boundary_cell_xy(p, tid, x, y);
    const Chain c = boundary_chain(p, x, y);
    const bool interior = x >= 2 && x < p.nx + 2 && y >= 2 && y < p.ny + 2;
    const Val own = pre_value(p, x, y);    // read before this thread writes anything
    bool write_self = true;
    if (pass != 2) {
        if ((int)c.mirrored != pass) return;
    } else if (c.have_source && (c.cx != x || c.cy != y) && c.cx >= 2 && c.cx < p.nx + 2 && c.cy >= 2 &&
               c.cy < p.ny + 2 && in_boundary_set(p, c.cx, c.cy)) {
        write_self = false;                // the thread of (c.cx, c.cy) writes this cell
    }
    if (write_self) {
        Val v{0.f, 0.f, 0.f};
        if (c.have_source && c.cx >= 0 && c.cx < p.X && c.cy >= 0 && c.cy < p.Y)
            v = (c.cx == x && c.cy == y) ? own : pre_value(p, c.cx, c.cy);
        write_boundary(p, x, y, apply_chain(p, c, v));
    }
    if (pass == 2 && interior) {
        // the mirror images of this cell across solid sides whose walks end here
        const bool mw = p.side[kWest].type == kSolid && 4 - x >= 0 && 4 - x < p.side[kWest].width;
        const bool me = p.side[kEast].type == kSolid && (2 * p.nx + 2) - x >= p.X - p.side[kEast].width &&
                        (2 * p.nx + 2) - x < p.X;
        const bool ms = p.side[kSouth].type == kSolid && 4 - y >= 0 && 4 - y < p.side[kSouth].width;
        const bool mn = p.side[kNorth].type == kSolid && (2 * p.ny + 2) - y >= p.Y - p.side[kNorth].width &&
                        (2 * p.ny + 2) - y < p.Y;
        if (mw || me || ms || mn) {
#pragma unroll 1
            for (int m = 1; m < 9; ++m) {
                const int rx = m % 3, ry = m / 3;
                if ((rx == 1 && !mw) || (rx == 2 && !me) || (ry == 1 && !ms) || (ry == 2 && !mn)) continue;
                const int dx = rx == 0 ? x : (rx == 1 ? 4 - x : (2 * p.nx + 2) - x);
                const int dy = ry == 0 ? y : (ry == 1 ? 4 - y : (2 * p.ny + 2) - y);
                if (dx == x && dy == y) continue;
                const Chain d = boundary_chain(p, dx, dy);
                if (d.have_source && d.cx == x && d.cy == y) write_boundary(p, dx, dy, apply_chain_sourced(p, d, own));
            }
        }
    }
}

__global__ __launch_bounds__(256) void boundary_apply(const BoundaryParams p)
{
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid < p.ncells) boundary_cell(p, tid, p.mirror_pass);
}

#ifndef CEL_FP64   // batches run in the reference's precision only
// ------------------------------------------------------------------ batched scenarios
//
// Many independent scenarios advance in one launch per kernel: the launch's blocks (or
// threads) are a work-list concatenated over the scenarios, and each finds its scenario
// by binary search over the list's start offsets.  Parameter blocks live in device
// memory, one per scenario and step.

// largest j with starts[j] <= v (starts[0] == 0, nondecreasing)
__device__ __forceinline__ int find_job(const int *starts, int njobs, int v)
{
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (starts[mid] <= v) lo = mid; else hi = mid - 1;
    }
    return lo;
}

template <int R>
__global__ __launch_bounds__(64) void explicit_batched(const ExplicitJob *__restrict__ jobs,
                                                      const int *__restrict__ starts, int njobs)
{
    __shared__ ExplicitSmem<false, R> sm;
    const int total = starts[njobs];
    const int per_xcd = (total + 7) / 8;
    const int t = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    if (t >= total) return;
    const int j = find_job(starts, njobs, t);
    const ExplicitJob &J = jobs[j];
    const int tile = t - starts[j];
    switch (J.scheme * 4 + J.comp * 2 + J.fast) {
    case 0: explicit_tile<0, false, R, false, false>(J.p, tile, sm); break;
    case 1: explicit_tile<0, false, R, false, true>(J.p, tile, sm); break;
    case 2: explicit_tile<0, true, R, false, false>(J.p, tile, sm); break;
    case 3: explicit_tile<0, true, R, false, true>(J.p, tile, sm); break;
    case 4: explicit_tile<1, false, R, false, false>(J.p, tile, sm); break;
    case 5: explicit_tile<1, false, R, false, true>(J.p, tile, sm); break;
    case 6: explicit_tile<1, true, R, false, false>(J.p, tile, sm); break;
    case 7: explicit_tile<1, true, R, false, true>(J.p, tile, sm); break;
    case 12: explicit_tile<3, false, R, false, false>(J.p, tile, sm); break;
    case 13: explicit_tile<3, false, R, false, true>(J.p, tile, sm); break;
    case 14: explicit_tile<3, true, R, false, false>(J.p, tile, sm); break;
    case 15: explicit_tile<3, true, R, false, true>(J.p, tile, sm); break;
    default: break;   // corrector steps are not batched (rejected on the host)
    }
}

template <int C, int W>
__global__ __launch_bounds__(64 * W) void solve_batched(const SolveJob *__restrict__ jobs, const int *__restrict__ starts,
                                                       int njobs)
{
    __shared__ LineSmem<C, W> sm;
    const int j = find_job(starts, njobs, blockIdx.x);
    const SolveJob &J = jobs[j];
    const LineParams &p = J.p;
    const IrregularAll &irr = J.irr;
    const int bid = blockIdx.x - starts[j];
    const int nlines = p.nlines_x + p.nlines_y;
    if (bid >= nlines) {
        int idx = (bid - nlines) * W + (threadIdx.x >> 6);
        for (int s = 0; s < 4; ++s) {
            if (idx < irr.count[s]) {
                const IrregularSide &S = irr.side[s];
                irregular_line(S, idx / S.nlines, idx % S.nlines, irr.t);
                return;
            }
            idx -= irr.count[s];
        }
        return;
    }
    const bool xdir = bid < p.nlines_x;
    const int line = xdir ? bid + 2 : (bid - p.nlines_x) + 2;
    const size_t o = xdir ? (size_t)line * p.P : (size_t)line * p.PT;
    const int n = xdir ? p.nx : p.ny;
    const real *D = xdir ? p.Dx : p.DyT;
    real *out = (xdir ? p.hu_out : p.hvT_out) + o + 2;
    const bool comp = J.comp != 0;
    const LineBed lb = line_bed(p, xdir, line);
    cr_line_block<C, W, false>(n, D + o, lb, out, sm, comp, comp ? (xdir ? p.hu_prev : p.hvT_prev) + o + 2 : nullptr,
                               comp ? (xdir ? p.hu_prev_lo : p.hvT_prev_lo) + o + 2 : nullptr,
                               comp ? (xdir ? p.hu_lo_out : p.hvT_lo_out) + o + 2 : nullptr);
}

__global__ __launch_bounds__(256) void boundary_batched(const BoundaryJob *__restrict__ jobs,
                                                        const int *__restrict__ starts, int njobs,
                                                        int mirror_pass)
{
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= starts[njobs]) return;
    const int j = find_job(starts, njobs, tid);
    boundary_cell(jobs[j].p, tid - starts[j], mirror_pass);
}

#endif

// ------------------------------------------------------------------ diagnostics

// GetStats, compute.hlsl:1319-1378: per-cell values in FP32 exactly as the shader forms
// them; the five sums are accumulated in FP64, the four maxima in FP32 (order-free).
// Block partials go to p.partial; stats_final reduces them in a fixed order, so the
// result does not depend on scheduling.
__device__ __forceinline__ double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ real wave_max(real v)
{
    for (int o = 32; o > 0; o >>= 1) v = FMAX(v, __shfl_xor(v, o, 64));
    return v;
}

__global__ __launch_bounds__(256) void stats_partial(const StatsParams p)
{
    __shared__ real hvs[kStatsTileY][kStatsTileX + 1];
    __shared__ double red[4][kStatsValues];
    const int tid = threadIdx.x;
    const int tx = blockIdx.x % p.tiles_x, ty = blockIdx.x / p.tiles_x;
    const int x0 = 2 + tx * kStatsTileX, y0 = 2 + ty * kStatsTileY;
    // hv tile through LDS (the engine's hv is column-major: lanes walk y)
    for (int k = tid; k < kStatsTileX * kStatsTileY; k += 256) {
        const int xl = k / kStatsTileY, yl = k % kStatsTileY;
        const int x = x0 + xl, y = y0 + yl;
        hvs[yl][xl] = (x < 2 + p.ncx && y < 2 + p.ncy) ? p.hv[(size_t)x * p.hv_sx + (size_t)y * p.hv_sy] : 0.f;
    }
    __syncthreads();
    double sh = 0, sb = 0, shu = 0, shv = 0, sk = 0;
    real mu = 0.f, mh = 0.f, mc = 0.f, mf = 0.f;
    const int xl = tid % kStatsTileX;
    for (int yl = tid / kStatsTileX; yl < kStatsTileY; yl += 256 / kStatsTileX) {
        const int x = x0 + xl, y = y0 + yl;
        if (x >= 2 + p.ncx || y >= 2 + p.ncy) continue;
        const size_t i = (size_t)y * p.P + x;
        const real w = p.w[i], hu = p.hu[i], hv = hvs[yl][xl];
        const real B = p.BA[i];
        const real h = FMAX(real(0), w - B);
        const real c = SQRT(p.g * h);
        const real dbh = divide_by_h(h, p.epsilon);
        const real u = dbh * hu, v = dbh * hv;
        sh += (double)h;
        sb += (double)(h * (B + 0.5f * h));
        shu += (double)hu;
        shv += (double)hv;
        sk += (double)(hu * u + hv * v);
        const real u2v2 = u * u + v * v;
        mu = FMAX(mu, u2v2);
        mh = FMAX(mh, h);
        mc = FMAX(mc, (FABS(u) + c) * p.one_over_dx);
        mc = FMAX(mc, (FABS(v) + c) * p.one_over_dy);
        mf = FMAX(mf, u2v2 * dbh);
    }
    const double vals[kStatsValues] = {wave_sum(sh), wave_sum(sb), wave_sum(shu), wave_sum(shv), wave_sum(sk),
                                       (double)wave_max(mu), (double)wave_max(mh), (double)wave_max(mc),
                                       (double)wave_max(mf)};
    const int wv = tid >> 6, lane = tid & 63;
    if (lane == 0)
        for (int v = 0; v < kStatsValues; ++v) red[wv][v] = vals[v];
    __syncthreads();
    if (tid < kStatsValues) {
        double r = red[0][tid];
        for (int k = 1; k < 4; ++k) r = tid < 5 ? r + red[k][tid] : fmax(r, red[k][tid]);
        p.partial[(size_t)blockIdx.x * kStatsValues + tid] = r;
    }
}

__global__ __launch_bounds__(64) void stats_final(const StatsParams p)
{
    const int lane = threadIdx.x;
    double acc[kStatsValues];
    for (int v = 0; v < kStatsValues; ++v) acc[v] = 0.0;
    for (int b = lane; b < p.nblocks; b += 64)
        for (int v = 0; v < kStatsValues; ++v) {
            const double x = p.partial[(size_t)b * kStatsValues + v];
            acc[v] = v < 5 ? acc[v] + x : fmax(acc[v], x);
        }
    for (int v = 0; v < kStatsValues; ++v) {
        double r = acc[v];
        for (int o = 32; o > 0; o >>= 1) {
            const double q = __shfl_xor(r, o, 64);
            r = v < 5 ? r + q : fmax(r, q);
        }
        if (lane == 0) p.out[v] = r;
    }
}

// ------------------------------------------------------------------ utilities

// (w, hu, hv) of n cells, xy = (column, row) pairs
__global__ void gather_cells(const real *w, const real *hu, const real *hvT, int P, int PT, const int *xy, int n,
                             real *out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int x = xy[2 * k], y = xy[2 * k + 1];
    out[3 * k + 0] = w[(size_t)y * P + x];
    out[3 * k + 1] = hu[(size_t)y * P + x];
    out[3 * k + 2] = hvT[(size_t)x * PT + y];
}

// ------------------------------------------------------------------ visual outputs

// Standard deviation output of Reconstruct, compute.hlsl:306-312: of the three slope
// estimates about half their mean.
__device__ __forceinline__ real slope_deviation(real two_theta, real west, real here, real east)
{
    const real z1 = two_theta * (here - west);
    const real z2 = (east - west);
    const real z3 = two_theta * (east - here);
    const real mu = real(0.5) * (z1 + z2 + z3) / real(3);
    return SQRT(((z1 - mu) * (z1 - mu) + (z2 - mu) * (z2 - mu) + (z3 - mu) * (z3 - mu)) / real(3));
}

__device__ __forceinline__ float ATAN(float x) { return atanf(x); }
__device__ __forceinline__ double ATAN(double x) { return atan(x); }
__device__ __forceinline__ float COSH(float x) { return coshf(x); }
__device__ __forceinline__ double COSH(double x) { return cosh(x); }

// One step of the auxiliary texture, on Pass2's cells [1, nx+2) x [1, ny+2): Pass1 of the
// step's starting state gives aux2 = (max(mean corrected face depth, aux1.r), 0,
// atan(max slope deviation / dx), whitewater set to 1 above the threshold)
// (compute.hlsl:404-457); Pass2 gives aux1 = (aux2.r, 0, aux2.b, aux2.a |decay|^dt)
// (compute.hlsl:535-540).
__global__ __launch_bounds__(256) void aux_update(const AuxParams p)
{
    const int x = 1 + blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = 1 + blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= p.nx + 2 || y >= p.ny + 2) return;
    const size_t i = (size_t)y * p.P + x;
    const real here = p.w[i], west = p.w[i - 1], east = p.w[i + 1];
    const real south = p.w[i - p.P], north = p.w[i + p.P];
    const real BN = p.BY[i], BE = p.BX[i], BS = p.BY[i - p.P], BW = p.BX[i - 1];
    const real sx = half_slope(p.two_theta, west, here, east);
    const real sy = half_slope(p.two_theta, south, here, north);
    real wE = here + sx, wW = here - sx, wN = here + sy, wS = here - sy;
    correct_w(BW, BE, here, wW, wE);
    correct_w(BS, BN, here, wS, wN);
    const real hN = wN - BN, hE = wE - BE, hS = wS - BS, hW = wW - BW;
    real sd = slope_deviation(p.two_theta, west, here, east) * p.one_over_dx;
    sd = ATAN(fmaxr(sd, slope_deviation(p.two_theta, south, here, north) * p.one_over_dy));
    p.depth[i] = fmaxr((hN + hE + hS + hW) / real(4), p.depth[i]);
    p.slope[i] = sd;
    const real white = sd > p.threshold ? real(1) : p.white[i];
    p.white[i] = white * p.decay;
}

// The normal texture of Pass1 (compute.hlsl:437-445) on [1, X-1) x [1, Y-1): normalize of
// ((w_W - w_E)/dx, (w_S - w_N)/dy, 2), taken as v / |v|.
__global__ __launch_bounds__(256) void state_normals(const NormalParams p)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= p.X * p.Y) return;
    const int x = k % p.X, y = k / p.X;
    const size_t N = (size_t)p.X * p.Y;
    real n0 = 0, n1 = 0, n2 = 0;
    if (x >= 1 && x < p.X - 1 && y >= 1 && y < p.Y - 1) {
        const size_t i = (size_t)y * p.P + x;
        const real a = (p.w[i - 1] - p.w[i + 1]) * p.one_over_dx;
        const real b = (p.w[i - p.P] - p.w[i + p.P]) * p.one_over_dy;
        const real c = real(2);
        const real len = SQRT(a * a + b * b + c * c);
        n0 = a / len; n1 = b / len; n2 = c / len;
    }
    p.out[k] = n0; p.out[N + k] = n1; p.out[2 * N + k] = n2;
}

// addSolitaryWave + solitaryWave (engine.cpp:2154-2220, compute.hlsl:947-976): every cell of
// the ghost-inclusive grid at x = i dx, y = j dy, local depth d = max(w - BA, 0):
// eta = H / cosh(K ((x - x0) cos + (y - y0) sin))^2, K = sqrt(0.75 |H| / d^3),
// C = sqrt(g (H + d)); (w, hu, hv) += (eta, C cos eta, C sin eta).  A compensated state
// takes the increment into its low words with an exact two-sum.
__global__ __launch_bounds__(256) void add_soliton(const SolitonParams p)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= p.X * p.Y) return;
    const int x = k % p.X, y = k / p.X;
    const size_t i = (size_t)y * p.P + x, t = (size_t)x * p.PT + y;
    const real w = p.w[i];
    const real d = fmaxr(w - p.BA[i], real(0));
    const real K = SQRT(real(0.75) * FABS(p.H) / (d * d * d));
    const real C = SQRT(p.g * (p.H + d));
    const real ch = COSH(K * (((real)x * p.dx - p.x0) * p.cos_t + ((real)y * p.dy - p.y0) * p.sin_t));
    const real eta = p.H * real(1) / (ch * ch);
    const real dhu = C * p.cos_t * eta, dhv = C * p.sin_t * eta;
    if (p.w_lo) {
        real s, e;
        two_sum(w, eta, s, e);
        p.w[i] = s; p.w_lo[i] += e;
        two_sum(p.hu[i], dhu, s, e);
        p.hu[i] = s; p.hu_lo[i] += e;
        two_sum(p.hvT[t], dhv, s, e);
        p.hvT[t] = s; p.hvT_lo[t] += e;
    } else {
        p.w[i] = w + eta;
        p.hu[i] = p.hu[i] + dhu;
        p.hvT[t] = p.hvT[t] + dhv;
    }
}

// D3D11 coverage of a cell centre by one fan triangle (a = centre, b, c on the rim):
// clockwise on screen (y down), drawn when every edge function is positive or zero on a top
// or left edge; lc receives the barycentric weight of the centre vertex.  Coordinates are
// multiples of 1/256 below 2^15, so the edge functions are exact in double.
__device__ __forceinline__ bool fan_covers(double ax, double ay, double bx, double by, double cx, double cy, double px,
                                           double py, double &lc)
{
    const double xs[3] = {ax, bx, cx}, ys[3] = {ay, by, cy};
    const double area = (bx - ax) * (cy - ay) - (by - ay) * (cx - ax);
    if (area == 0.0) return false;
    const double sg = area > 0 ? 1.0 : -1.0;
    double e1 = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int k1 = (k + 1) % 3;
        double dx = xs[k1] - xs[k], dy = ys[k1] - ys[k];
        const double e = sg * (dx * (py - ys[k]) - dy * (px - xs[k]));
        dx *= sg; dy *= sg;
        const bool top_left = dy < 0 || (dy == 0 && dx > 0);
        if (e < 0 || (e == 0 && !top_left)) return false;
        if (k == 1) e1 = e;
    }
    lc = e1 / (sg * area);
    return true;
}

// LeftMousePixelShader (left_mouse.hlsl:80-103) on the cells the brush fan covers:
// dh = disp_A + disp_B radius, w' = max(B, w + dh), (hu, hv)' = (w' - B) (hu, hv) / max(0.5, w - B).
// A compensated state's low words of edited cells are cleared (the values are set).
__global__ __launch_bounds__(256) void water_brush(const WaterBrushParams p)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= p.bw * p.bh) return;
    const int i = p.x0 + k % p.bw, j = p.y0 + k / p.bw;
    const double px = i + 0.5, py = j + 0.5;
    double lc = 0.0;
    bool hit = false;
    for (int t = 0; t < 12 && !hit; ++t)
        hit = fan_covers(p.vx[12], p.vy[12], p.vx[t], p.vy[t], p.vx[(t + 1) % 12], p.vy[(t + 1) % 12], px, py, lc);
    if (!hit) return;
    const float rad = (float)(1.0 - lc);
    const size_t c = (size_t)j * p.P + i, ct = (size_t)i * p.PT + j;
    const real B = p.BA[c];
    const real w = p.w[c];
    const real dh = (real)p.disp_A + (real)p.disp_B * (real)rad;
    const real old_h = w - B;
    const real new_w = fmaxr(B, w + dh);
    const real new_h = new_w - B;
    const real lim = fmaxr(real(0.5f), old_h);
    const real u = p.hu[c] / lim, v = p.hvT[ct] / lim;
    p.w[c] = new_w;
    p.hu[c] = new_h * u;
    p.hvT[ct] = new_h * v;
    if (p.w_lo) { p.w_lo[c] = 0; p.hu_lo[c] = 0; p.hvT_lo[ct] = 0; }
}

// out[j * pout + i] = in[i * pin + j] for i < rows, j < cols (host-side layout changes)
__global__ void transpose_plane(const real *in, int pin, real *out, int pout, int rows, int cols)
{
    __shared__ real tile[32][33];
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
    for (int j = threadIdx.y; j < 32; j += 8) {
        const int r = by + j, c = bx + threadIdx.x;
        if (r < rows && c < cols) tile[j][threadIdx.x] = in[(size_t)r * pin + c];
    }
    __syncthreads();
    for (int j = threadIdx.y; j < 32; j += 8) {
        const int c = bx + j, r = by + threadIdx.x;
        if (r < rows && c < cols) out[(size_t)c * pout + r] = tile[threadIdx.x][j];
    }
}

// 0, or a magnitude in [2^-80, 2^80]: the numerators div_k rounds correctly
__device__ __forceinline__ bool div_k_range(real a)
{
    const real m = FABS(a);
    return a == real(0) || (m >= real(0x1p-80f) && m <= real(0x1p80f));
}

// Range check of every interior cell for the FAST variants (explicit kernel and line
// solves): the numerators of the divisions by dx^2, dy^2, 72 and 12 at the cell (K1's d^2,
// d, K2 of both lines, the depth-gradient sums, engine.cpp:2656-2680 and compute.hlsl:
// 641-697) must lie in div_k's range; *p.unsafe is raised otherwise.
__global__ void static_fields(const StaticParams p)
{
    const int x = 2 + blockIdx.x * blockDim.x + threadIdx.x;
    const int y = 2 + blockIdx.y;
    if (x >= p.nx + 2 || y >= p.ny + 2) return;
    const real *BA = p.BA;
    auto ba = [&](int xx, int yy) { return BA[(size_t)yy * p.P + xx]; };
    const real cs = p.coef_sl;
    auto dc = [&](int xx, int yy) { return fmaxr(0.f, cs - ba(xx, yy)); };
    const real ex[5] = {dc(x - 2, y), dc(x - 1, y), dc(x, y), dc(x + 1, y), dc(x + 2, y)};
    const real ey[5] = {dc(x, y - 2), dc(x, y - 1), ex[2], dc(x, y + 1), dc(x, y + 2)};
    {
        const real d0 = ex[2];
        const real K2x = (d0 / p.dx2) * (-ex[4] + 8.f * ex[3] - 8.f * ex[1] + ex[0]);
        const real K2y = (d0 / p.dy2) * (-ey[4] + 8.f * ey[3] - 8.f * ey[1] + ey[0]);
        auto ds = [&](int xx, int yy) { return fmaxr(0.f, p.sl - ba(xx, yy)); };
        const real Nx = (-ds(x + 2, y) + 8.f * ds(x + 1, y) - 8.f * ds(x - 1, y) + ds(x - 2, y)) * p.odx;
        const real Ny = (-ds(x, y + 2) + 8.f * ds(x, y + 1) - 8.f * ds(x, y - 1) + ds(x, y - 2)) * p.ody;
        if (!(div_k_range(d0 * d0) && div_k_range(d0) && div_k_range(K2x) && div_k_range(K2y) && div_k_range(Nx) &&
              div_k_range(Ny)))
            atomicOr(p.unsafe, 1);
    }
}

// level-0 coefficients of every cell for the x- and y-lines (diagnostic export): ghost rows
// of the x-lines repeat the nearest interior row, ghost columns of the y-lines the nearest
// interior column (engine.cpp:2641, 2797); line margins are (0, 1, 0).
__global__ void export_coefficients(const StaticParams p, real *ax, real *ay)
{
    const int X = p.nx + 4, Y = p.ny + 4;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= X * Y) return;
    const int x = i % X, y = i / X;
    const size_t N = (size_t)X * Y;
    auto d = [&](int xx, int yy) { return fmaxr(0.f, p.coef_sl - p.BA[(size_t)yy * p.P + xx]); };
    real A = 0.f, B = 1.f, C = 0.f, K1, Q;
    if (x >= 2 && x <= X - 3) {
        const int yy = y < 2 ? 2 : (y > p.ny + 1 ? p.ny + 1 : y);
        line_kq(p.kcoef, p.dx2, p.rdx2, d(x - 2, yy), d(x - 1, yy), d(x, yy), d(x + 1, yy), d(x + 2, yy), K1, Q);
        line_abc(K1, Q, x == 2, x == X - 3, A, B, C);
    }
    ax[i] = A; ax[N + i] = B; ax[2 * N + i] = C;
    A = 0.f; B = 1.f; C = 0.f;
    if (y >= 2 && y <= Y - 3) {
        const int xx = x < 2 ? 2 : (x > p.nx + 1 ? p.nx + 1 : x);
        line_kq(p.kcoef, p.dy2, p.rdy2, d(xx, y - 2), d(xx, y - 1), d(xx, y), d(xx, y + 1), d(xx, y + 2), K1, Q);
        line_abc(K1, Q, y == 2, y == Y - 3, A, B, C);
    }
    ay[i] = A; ay[N + i] = B; ay[2 * N + i] = C;
}


}  // namespace CEL_NS

// ------------------------------------------------------------------ launchers (host)

namespace CEL_NS {

hipError_t launch_explicit(int scheme, bool comp, int rows, bool fast, const ExplicitParams &p, hipStream_t st)
{
    const dim3 grid(8 * ((p.nstrips * p.nbands + 7) / 8)), block(64);
#define CEL_EXPLICIT(S, CP, R)                                                                  \
    do {                                                                                        \
        if (fast) hipLaunchKernelGGL((explicit_step<S, CP, R, true>), grid, block, 0, st, p);   \
        else hipLaunchKernelGGL((explicit_step<S, CP, R, false>), grid, block, 0, st, p);       \
    } while (0)
#define CEL_EXPLICIT_R(S, CP)                                  \
    do {                                                       \
        if (rows == 4) CEL_EXPLICIT(S, CP, 4);                 \
        else if (rows == 2) CEL_EXPLICIT(S, CP, 2);            \
        else if (rows == 12) CEL_EXPLICIT(S, CP, 12);          \
        else if (rows == 16) CEL_EXPLICIT(S, CP, 16);          \
        else CEL_EXPLICIT(S, CP, 8);                           \
    } while (0)
    if (rows != 2 && rows != 4 && rows != 8 && rows != 12 && rows != 16) return hipErrorInvalidValue;
    if (comp) {
        switch (scheme) {
        case 0: CEL_EXPLICIT_R(0, true); break;
        case 1: CEL_EXPLICIT_R(1, true); break;
        case 3: CEL_EXPLICIT_R(3, true); break;
        default: return hipErrorInvalidValue;   // no compensated corrector
        }
    } else {
        switch (scheme) {
        case 0: CEL_EXPLICIT_R(0, false); break;
        case 1: CEL_EXPLICIT_R(1, false); break;
        case 2: CEL_EXPLICIT_R(2, false); break;
        default: CEL_EXPLICIT_R(3, false); break;
        }
    }
#undef CEL_EXPLICIT_R
#undef CEL_EXPLICIT
    return hipGetLastError();
}

// Default line-solve shape for a line of n unknowns: (1,1) (2,1) (4,1) (8,1) (8,2) (8,4)
// (16,4) (16,8) by length class.
int default_shape(int n)
{
    return n <= 64 ? 0 : n <= 128 ? 1 : n <= 256 ? 3 : n <= 512 ? 5 : n <= 1024 ? 9 : n <= 2048 ? 12 : n <= 4096 ? 14
         : 16;
}

int shape_capacity(int shape) { return 64 * kShapeC[shape] * kShapeW[shape]; }

// Dispatch on a line-solve shape index (kShapeC/kShapeW).
#define CEL_SHAPE_SWITCH(shape, MACRO)                                                              \
    switch (shape) {                                                                                \
    case 0: MACRO(1, 1); break;   case 1: MACRO(2, 1); break;   case 2: MACRO(1, 2); break;          \
    case 3: MACRO(4, 1); break;   case 4: MACRO(2, 2); break;   case 5: MACRO(8, 1); break;          \
    case 6: MACRO(4, 2); break;   case 7: MACRO(2, 4); break;   case 8: MACRO(16, 1); break;         \
    case 9: MACRO(8, 2); break;   case 10: MACRO(4, 4); break;  case 11: MACRO(16, 2); break;        \
    case 12: MACRO(8, 4); break;  case 13: MACRO(4, 8); break;  case 14: MACRO(16, 4); break;        \
    case 15: MACRO(8, 8); break;  default: MACRO(16, 8); break;                                      \
    }

static void launch_solve_shape(int shape, bool comp, const LineParams &lp, const IrregularAll &ir, int nirr,
                               hipStream_t st)
{
    const int W = kShapeW[shape];
    const dim3 grid(lp.nlines_x + lp.nlines_y + (nirr + W - 1) / W);
#define CEL_SOLVE(C, W_)                                                                                    \
    do {                                                                                                    \
        if (comp) hipLaunchKernelGGL((solve_lines_block<C, W_, true>), grid, dim3(64 * W_), 0, st, lp, ir); \
        else hipLaunchKernelGGL((solve_lines_block<C, W_, false>), grid, dim3(64 * W_), 0, st, lp, ir);     \
    } while (0)
    CEL_SHAPE_SWITCH(shape, CEL_SOLVE)
#undef CEL_SOLVE
}

// One launch per line direction (x-lines with the irregular-wave lines, then y-lines), each
// with its own block shape; one launch when both directions share a shape.
hipError_t launch_solve(bool comp, int shape_x, int shape_y, const LineParams &p, const IrregularAll &irr,
                        hipStream_t st)
{
    if (shape_x < 0 || shape_x >= kNumShapes || shape_y < 0 || shape_y >= kNumShapes) return hipErrorInvalidValue;
    if (shape_capacity(shape_x) < p.nx || shape_capacity(shape_y) < p.ny) return hipErrorInvalidValue;
    int irr_lines = 0;
    for (int s = 0; s < 4; ++s) irr_lines += irr.count[s];
    if (shape_x == shape_y && p.nlines_x > 0 && p.nlines_y > 0) {
        launch_solve_shape(shape_x, comp, p, irr, irr_lines, st);
    } else {
        LineParams px = p, py = p;
        px.nlines_y = 0;
        py.nlines_x = 0;
        IrregularAll none = irr;
        for (int s = 0; s < 4; ++s) none.count[s] = 0;
        if (px.nlines_x > 0 || irr_lines > 0) launch_solve_shape(shape_x, comp, px, irr, irr_lines, st);
        if (py.nlines_y > 0) launch_solve_shape(shape_y, comp, py, none, 0, st);
    }
    return hipGetLastError();
}

#ifndef CEL_FP64
hipError_t launch_explicit_batched(int rows, const ExplicitJob *jobs, const int *starts, int njobs, int total_tiles,
                                   hipStream_t st)
{
    if (total_tiles <= 0) return hipSuccess;
    const dim3 grid(8 * ((total_tiles + 7) / 8));
    if (rows == 4) hipLaunchKernelGGL(explicit_batched<4>, grid, dim3(64), 0, st, jobs, starts, njobs);
    else if (rows == 16) hipLaunchKernelGGL(explicit_batched<16>, grid, dim3(64), 0, st, jobs, starts, njobs);
    else if (rows == 8) hipLaunchKernelGGL(explicit_batched<8>, grid, dim3(64), 0, st, jobs, starts, njobs);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_solve_batched(int shape, const SolveJob *jobs, const int *starts, int njobs, int total_blocks,
                                hipStream_t st)
{
    if (total_blocks <= 0) return hipSuccess;
    if (shape < 0 || shape >= kNumShapes) return hipErrorInvalidValue;
    const dim3 grid(total_blocks);
#define CEL_SB(C, W) hipLaunchKernelGGL((solve_batched<C, W>), grid, dim3(64 * (W)), 0, st, jobs, starts, njobs)
    CEL_SHAPE_SWITCH(shape, CEL_SB)
#undef CEL_SB
    return hipGetLastError();
}
#undef CEL_SHAPE_SWITCH

hipError_t launch_boundary_batched(const BoundaryJob *jobs, const int *starts, int njobs, int total_cells,
                                   int mirror_pass, hipStream_t st)
{
    if (total_cells > 0)
        hipLaunchKernelGGL(boundary_batched, dim3((total_cells + 255) / 256), dim3(256), 0, st, jobs, starts, njobs,
                           mirror_pass);
    return hipGetLastError();
}

#endif

hipError_t launch_boundary(const BoundaryParams &p, hipStream_t st)
{
    if (p.ncells == 0) return hipSuccess;
    const dim3 grid((p.ncells + 255) / 256), block(256);
    hipLaunchKernelGGL(boundary_apply, grid, block, 0, st, p);
    return hipGetLastError();
}

hipError_t launch_transpose(const real *in, int pin, real *out, int pout, int rows, int cols, hipStream_t st)
{
    hipLaunchKernelGGL(transpose_plane, dim3((cols + 31) / 32, (rows + 31) / 32), dim3(32, 8), 0, st, in, pin, out, pout,
                       rows, cols);
    return hipGetLastError();
}

hipError_t launch_static_fields(const StaticParams &p, hipStream_t st)
{
    const int X = p.nx + 4, Y = p.ny + 4;
    hipLaunchKernelGGL(transpose_plane, dim3((X + 31) / 32, (Y + 31) / 32), dim3(32, 8), 0, st, p.BA, p.P, p.BAT, p.PT,
                       Y, X);
    if (p.unsafe) hipLaunchKernelGGL(static_fields, dim3((p.nx + 127) / 128, p.ny), dim3(128), 0, st, p);
    return hipGetLastError();
}

hipError_t launch_gather(const real *w, const real *hu, const real *hvT, int P, int PT, const int *xy, int n,
                         real *out, hipStream_t st)
{
    if (n > 0) hipLaunchKernelGGL(gather_cells, dim3((n + 255) / 256), dim3(256), 0, st, w, hu, hvT, P, PT, xy, n, out);
    return hipGetLastError();
}

hipError_t launch_aux(const AuxParams &p, hipStream_t st)
{
    hipLaunchKernelGGL(aux_update, dim3((p.nx + 1 + 63) / 64, (p.ny + 1 + 3) / 4), dim3(256), 0, st, p);
    return hipGetLastError();
}

hipError_t launch_normals(const NormalParams &p, hipStream_t st)
{
    const int n = p.X * p.Y;
    hipLaunchKernelGGL(state_normals, dim3((n + 255) / 256), dim3(256), 0, st, p);
    return hipGetLastError();
}

hipError_t launch_soliton(const SolitonParams &p, hipStream_t st)
{
    const int n = p.X * p.Y;
    hipLaunchKernelGGL(add_soliton, dim3((n + 255) / 256), dim3(256), 0, st, p);
    return hipGetLastError();
}

hipError_t launch_water_brush(const WaterBrushParams &p, hipStream_t st)
{
    const int n = p.bw * p.bh;
    if (n > 0) hipLaunchKernelGGL(water_brush, dim3((n + 255) / 256), dim3(256), 0, st, p);
    return hipGetLastError();
}

hipError_t launch_stats(const StatsParams &p, hipStream_t st)
{
    if (p.nblocks > 0) hipLaunchKernelGGL(stats_partial, dim3(p.nblocks), dim3(256), 0, st, p);
    hipLaunchKernelGGL(stats_final, dim3(1), dim3(64), 0, st, p);
    return hipGetLastError();
}

#ifndef CEL_FP64
// ------------------------------------------------------------------ statistics without allocation
//
// GetStats over caller-owned device planes (row-major, ghost-inclusive, one pitch) with the
// per-block partial sums in a file-scope device array rather than engine memory, so the call
// allocates nothing: the same stats_partial / stats_final launches as an engine's statistics
// (same tiles, same orders), hence the same results.  Each GPU holds its own copy of the array
// until the process ends; host code uses the address hipGetSymbolAddress reports for the
// current device, looked up once per device.  Calls that overlap on one GPU share the
// array and race; the partials need no clearing (every block writes its own).
constexpr int kNoAllocStatsTiles = 16384;    // 64 x 16-cell tiles: grids up to 4096 x 4096
static __device__ __align__(256) double scratch[(size_t)kNoAllocStatsTiles * kStatsValues];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double *g_scratch_address[kMaxDevices] = {};  // per device; null until looked up
}  // namespace

hipError_t stats_noalloc(int nx, int ny, int pitch, const real *w, const real *hu, const real *hv, const real *BA,
                         real g, real one_over_dx, real one_over_dy, real epsilon, double *out9, hipStream_t st)
{
    StatsParams p;
    std::memset(&p, 0, sizeof(p));
    p.ncx = 4 * (nx / 4); p.ncy = 4 * (ny / 4);
    p.tiles_x = (p.ncx + kStatsTileX - 1) / kStatsTileX;
    p.nblocks = p.tiles_x * ((p.ncy + kStatsTileY - 1) / kStatsTileY);
    if (p.nblocks > kNoAllocStatsTiles) return hipErrorInvalidValue;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    double *partial = nullptr;
    {
        if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
        std::lock_guard<std::mutex> lock(g_scratch_mutex);
        if (!g_scratch_address[dev]) {
            void *addr = nullptr;
            e = hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch));
            if (e != hipSuccess) return e;
            g_scratch_address[dev] = static_cast<double *>(addr);
        }
        partial = g_scratch_address[dev];
    }
    p.P = pitch; p.hv_sx = 1; p.hv_sy = pitch;
    p.w = w; p.hu = hu; p.hv = hv; p.BA = BA;
    p.g = g; p.one_over_dx = one_over_dx; p.one_over_dy = one_over_dy; p.epsilon = epsilon;
    p.partial = partial;
    p.out = out9;
    return launch_stats(p, st);
}
#endif

hipError_t launch_export_coefficients(const StaticParams &p, real *ax, real *ay, hipStream_t st)
{
    const int n = (p.nx + 4) * (p.ny + 4);
    hipLaunchKernelGGL(export_coefficients, dim3((n + 255) / 256), dim3(256), 0, st, p, ax, ay);
    return hipGetLastError();
}

}  // namespace CEL_NS
