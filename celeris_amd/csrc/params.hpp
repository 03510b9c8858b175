// params.hpp -- launch parameter blocks shared by the host engine and the HIP kernels.
//
// Included once per working precision: namespace cel with real = float (the reference's
// precision) and, when CEL_FP64 is defined, namespace cel64 with real = double.
//
// Device memory layout (structure of arrays of `real`):
//   row-major planes    : element (x, y) at [y * P + x],  P  = padded row pitch  (>= nx+4+16)
//   transposed planes   : element (x, y) at [x * PT + y], PT = padded column pitch (>= ny+4+16)
// x and y are ghost-inclusive indices, interior cells are [2, nx+2) x [2, ny+2).
// hv and everything feeding the y-line solve live transposed so that both line
// solves read contiguous memory.
#undef CEL_NS
#ifdef CEL_FP64
#define CEL_NS cel64
#else
#define CEL_NS cel
#endif

#if (defined(CEL_FP64) && !defined(CEL_PARAMS_HPP_FP64)) || (!defined(CEL_FP64) && !defined(CEL_PARAMS_HPP_FP32))
#ifdef CEL_FP64
#define CEL_PARAMS_HPP_FP64
#else
#define CEL_PARAMS_HPP_FP32
#endif

namespace CEL_NS {

#ifdef CEL_FP64
using real = double;
#else
using real = float;
#endif

// Rows one wave marches through in the explicit kernel: 2, 4, 8, 12 or 16 for an engine (by
// default the tallest of 12 and 8 whose bands fill the device, else 2), 4, 8 or 16 for a
// batch (8 until its autotuner picks).
constexpr int kDefaultBandRows = 8;
constexpr int kStripCols = 60;  // interior columns one wave produces (64 lanes - 2 halo each side)

// Line-solve block shapes: C unknowns per thread, W waves per line (64 W C >= line length).
constexpr int kNumShapes = 17;
constexpr int kShapeC[kNumShapes] = {1, 2, 1, 4, 2, 8, 4, 2, 16, 8, 4, 16, 8, 4, 16, 8, 16};
constexpr int kShapeW[kNumShapes] = {1, 1, 2, 1, 2, 1, 2, 4, 1, 2, 4, 2, 4, 8, 4, 8, 8};

struct ExplicitParams {
    int nx, ny, P, PT;
    int nstrips, nbands;
    // state at step n
    const real *w, *hu, *hvT;
    // compensated state: low word of w at n (read) and n+1 (written)
    const real *w_lo;
    real *w_lo_new;
    // bottom: BY (north face), BX (east face), BA (centre)
    const real *BY, *BX, *BA;
    // Adams-Bashforth history, one ring slot per step: the dU/dt of (w, hu, hv) and F* of a
    // cell as four consecutive reals (a quad plane, [4 (y P + x) + k]) and G* in a plane of
    // its own.  Slots of steps n-1 and n-2:
    const real *Ho, *Ho1, *Hoo, *Hoo1;
    // corrector inputs: the predicted slot (same layout), and the state the predictor started from
    const real *Hp, *Hp1, *Cw, *Chu, *ChvT;
    // outputs
    real *w_new;    // w^{n+1} (interior)
    real *Dx;       // right-hand side of the x-line systems, [y * P + (x - 2)]
    real *DyT;      // right-hand side of the y-line systems, [x * PT + (y - 2)]
    real *Hs, *Hs1;   // history slot of step n (quad plane, G* plane)
    // numeric sentinel: bit 0 is raised when a new w or a line right-hand side is not finite
    // (null: not checked)
    int *status;
    // SimConstBuffer (compute.hlsl:65-94)
    real two_theta, g, half_g, Bcoef_g, g_over_dx, g_over_dy, one_over_dx, one_over_dy;
    real dt, epsilon, friction, sea_level;
    int is_manning;
    // variable-step AB3 coefficients (AdamsBashforthCoef) and derivative weights
    real ab[3];
    real der[9];
    // line-coefficient constants (BCOEF + 1/3, dx^2, dy^2, the coefficient sea level): the
    // kernel forms K1, K2/72 and the depth gradients from the bed as the line solves do
    real kcoef, dx2, dy2, coef_sl;
    real rdx2, rdy2;               // RN(1/dx^2), RN(1/dy^2) for the FAST variants
};

// Static per-cell fields derived from the bottom once per mesh (bit-identical to computing
// them inside the step: same operations on the same inputs).
struct StaticParams {
    int nx, ny, P, PT;
    const real *BA;
    real *BAT;                             // BA transposed, [x * PT + y] (ghost-inclusive)
    real kcoef, dx2, dy2, coef_sl;         // (BCOEF + 1/3), dx^2, dy^2, coefficient sea level
    real rdx2, rdy2;                       // RN(1/dx^2), RN(1/dy^2)
    real sl, odx, ody;                     // sea level, 1/dx, 1/dy (the depth gradients)
    int *unsafe;                           // raised when a cell rules out the FAST explicit kernel (or null)
};

struct LineParams {
    int nx, ny, P, PT;
    int nlines_x, nlines_y;         // blocks [0, nlines_x) solve x-lines, then y-lines
    const real *Dx;                // x-lines: right-hand sides, [y * P + (x - 2)]
    const real *DyT;               // y-lines: the same, transposed, [x * PT + (y - 2)]
    // the bed, from which each line forms its level-0 coefficients: row-major [y * P + x]
    // for the x-lines and transposed [x * PT + y] for the y-lines (ghost-inclusive)
    const real *BA, *BAT;
    real kcoef, dx2, dy2, rdx2, rdy2, coef_sl;   // as ExplicitParams
    int fast;                      // coefficient divisions through div_k (the mesh admits it)
    real *hu_out, *hvT_out;
    // compensated state: previous hi/lo words and the new low words (same layouts as out)
    const real *hu_prev, *hu_prev_lo, *hvT_prev, *hvT_prev_lo;
    real *hu_lo_out, *hvT_lo_out;
    int Lx, Ly;                     // cyclic-reduction level counts
    // irregular-wave strips, handled by blocks after the line solves
    int n_irr_lines;                // total (side, strip-index, line) triples
};

enum : int { kSolid = 0, kSponge = 1, kSine = 2, kIrregular = 3, kUTS = 4 };
enum : int { kWest = 0, kEast = 1, kSouth = 2, kNorth = 3 };

struct IrregularSide {
    int active;          // 1 if the side is IrregularWaves
    int width;           // columns (W/E) or rows (S/N) of the strip
    int nlines;          // cells along the strip (ny+4 for W/E, nx+4 for S/N)
    int nleaf_active;    // leaves that can hold a component: min(n_waves, 2^(L-1))
    int nleaf_pow2;      // smallest power of two >= nleaf_active (<= 2^(L-1))
    int levels_above;    // log2(2^(L-1) / nleaf_pow2)
    int zero_line;       // 1 if the reference copies an all-zero texture (L == 1)
    real sl;
    const real *amp, *omega, *phase;    // per component
    const real *tab;                    // [(strip * nlines + line) * nw + wave] * 4: kx, ky, c*cos, c*sin
    const unsigned char *wet;            // [strip * nlines + line]: d > MIN_DEPTH
    int nw;                              // components stored per line
    real *out;                          // [(strip * nlines + line) * 3]
};

// Step clock: per-step values that change every step while the launch pattern does not.
// A captured multi-step block reads them from a device table, one entry per step:
// [0] = t^n, [1 + 3 side + {0, 1, 2}] = the uniform-time-series (eta, hu, hv) of a side.
constexpr int kClockStride = 16;

struct IrregularAll {
    IrregularSide side[4];
    int count[4];      // width * nlines per side (0 if inactive)
    real t;
    const real *clock;   // if set, t = clock[0]
};

struct SideParams {
    int type, width;
    real sl;
    const real *tab;          // sponge: gamma; sine: kx, ky, c*cos, c*sin, wet (5 floats) per strip cell
    real amp, omega;          // sine
    real uts_eta, uts_hu, uts_hv;
    const real *irr;          // irregular: strip values from IrregularSide::out
};

struct BoundaryParams {
    int nx, ny, P, PT;
    int X, Y;
    int band_l, band_r, band_b, band_t;   // widths of the boundary cell set (>= 2 each)
    int ncells;
    int mirror_pass;                      // 1: cells whose chain passes a solid side; 0: all others; 2: all in one launch
    real t;
    const real *clock;                   // if set, t and the time-series values come from it
    real *w, *hu, *hvT;
    real *w_lo, *hu_lo, *hvT_lo;         // compensated state: low words (zeroed where set)
    SideParams side[4];
};

// One scenario's part of a batched launch (parameters of one step).
struct ExplicitJob {
    ExplicitParams p;
    int scheme, comp, fast;
};
struct SolveJob {
    LineParams p;
    IrregularAll irr;
    int comp;
};
struct BoundaryJob {
    BoundaryParams p;
};

// GetStats (compute.hlsl:1319-1378): cells [2, 2 + ncx) x [2, 2 + ncy), tiles of
// kStatsTileX x kStatsTileY cells per block, nine partials per block.
constexpr int kStatsTileX = 64, kStatsTileY = 16, kStatsValues = 9;
struct StatsParams {
    int ncx, ncy, P;                 // reduced cells; row pitch of w, hu and BA
    int hv_sx, hv_sy;                // hv element (x, y) at hv[x * hv_sx + y * hv_sy]
    int tiles_x, nblocks;
    const real *w, *hu, *hv, *BA;
    real g, one_over_dx, one_over_dy, epsilon;
    double *partial;   // [block * kStatsValues + v]
    double *out;       // kStatsValues
};

// Visual outputs of one step (Pass1 aux + Pass2 decay): the state at the start of the
// step, its face bottoms, and the three auxiliary planes updated in place.
struct AuxParams {
    int nx, ny, P;
    const real *w, *BX, *BY;
    real *depth, *slope, *white;
    real two_theta, one_over_dx, one_over_dy;
    real threshold;    // dissipation threshold on the slope angle
    real decay;        // |whiteWaterDecayRate|^dt of the step
};

// addSolitaryWave over the whole ghost-inclusive grid.
struct SolitonParams {
    int X, Y, P, PT;
    real *w, *hu, *hvT;
    real *w_lo, *hu_lo, *hvT_lo;     // compensated state, else null
    const real *BA;
    real g, dx, dy, H, cos_t, sin_t, x0, y0;
};

// The left-mouse water brush: fan vertices (0..11 rim, 12 centre) in cell-index units,
// snapped to 1/256, and the cell box [x0, x0 + bw) x [y0, y0 + bh) that can be covered.
struct WaterBrushParams {
    double vx[13], vy[13];
    int x0, y0, bw, bh, P, PT;
    real *w, *hu, *hvT;
    real *w_lo, *hu_lo, *hvT_lo;     // compensated state, else null
    const real *BA;
    float disp_A, disp_B;
};

// Normals of the current state (Pass1's visual target), planes of X x Y.
struct NormalParams {
    int X, Y, P;
    const real *w;
    real one_over_dx, one_over_dy;
    real *out;         // 3 planes, row-major X x Y
};

}  // namespace CEL_NS

#endif
