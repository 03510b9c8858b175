// engine.hpp -- host side of the MI355X Celeris step: ShallowWaterEngine.
//
// Mirrors the numerical interface of the reference class (engine.hpp:58-97):
// construction + remesh, resetTimestep (fixed, caller-chosen or adaptive dt),
// timestep(), boundary changes (createBoundaryShaders), plus state upload/download.  All
// device resources are owned here; the D3D11 device/context pair of the reference is
// replaced by one HIP stream per engine.  Included once per working precision (see
// params.hpp): cel::ShallowWaterEngine works in FP32, cel64::ShallowWaterEngine in FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/celeris.h"
#include "engine_api.hpp"
#include "params.hpp"

#if (defined(CEL_FP64) && !defined(CEL_ENGINE_HPP_FP64)) || (!defined(CEL_FP64) && !defined(CEL_ENGINE_HPP_FP32))
#ifdef CEL_FP64
#define CEL_ENGINE_HPP_FP64
#else
#define CEL_ENGINE_HPP_FP32
#endif

namespace CEL_NS {

#ifdef CEL_FP64
using cel::Error;
#endif

hipError_t launch_explicit(int scheme, bool comp, int rows, bool fast, const ExplicitParams &p, hipStream_t st);
hipError_t launch_solve(bool comp, int shape_x, int shape_y, const LineParams &p, const IrregularAll &irr,
                        hipStream_t st);
int default_shape(int n);
int shape_capacity(int shape);
hipError_t launch_boundary(const BoundaryParams &p, hipStream_t st);
hipError_t launch_static_fields(const StaticParams &p, hipStream_t st);
// out[j * pout + i] = in[i * pin + j] for i < rows, j < cols
hipError_t launch_transpose(const real *in, int pin, real *out, int pout, int rows, int cols, hipStream_t st);
hipError_t launch_export_coefficients(const StaticParams &p, real *ax, real *ay, hipStream_t st);
hipError_t launch_stats(const StatsParams &p, hipStream_t st);
hipError_t launch_gather(const real *w, const real *hu, const real *hvT, int P, int PT, const int *xy, int n,
                         real *out, hipStream_t st);
hipError_t launch_aux(const AuxParams &p, hipStream_t st);
hipError_t launch_normals(const NormalParams &p, hipStream_t st);
hipError_t launch_soliton(const SolitonParams &p, hipStream_t st);
hipError_t launch_water_brush(const WaterBrushParams &p, hipStream_t st);

// st_getCountOfMatrices, engine.cpp:741-748
int count_of_matrices(int n);

// Host copy of one side's BoundarySetting plus the per-strip-cell tables built from it.
struct SideState {
    cel_boundary cfg{};
    std::vector<float> waves;     // irregular table (amp, period, theta, phase)
    std::vector<float> uts;       // (t, eta, hu, hv)
    int uts_iter = 1;             // UniformTimeSeries::iter
    real *d_tab = nullptr;        // sponge gamma / sine table / irregular table
    real *d_irr_out = nullptr;    // irregular strip values
    real *d_irr_wave = nullptr;   // amp, omega, phase per component
    unsigned char *d_wet = nullptr;
    real sine_omega = 0;
    int irr_nw = 0;
};

// Everything one step issues (kernel parameter blocks, predictor copies, step clock).
struct StepLaunch {
    struct Copy { real *dst; const real *src; size_t bytes; };
    int scheme;
    int rows, shape_x, shape_y;     // explicit band rows, line-solve shapes
    int ncopy;
    Copy copy[8];
    ExplicitParams ep;
    LineParams lp;
    IrregularAll irr;
    BoundaryParams bp;
    bool aux;                       // visual outputs on
    AuxParams ap;
    real clock[kClockStride];
};

constexpr int kMaxGraphSteps = 1024;

class ShallowWaterEngine final : public cel::EngineApi {
public:
    ShallowWaterEngine(const cel_settings &s, const cel_boundary bnd[4], const float *bottom3, const float *state3);
    ~ShallowWaterEngine();
    ShallowWaterEngine(const ShallowWaterEngine &) = delete;
    ShallowWaterEngine &operator=(const ShallowWaterEngine &) = delete;

    void setBoundary(int side, const cel_boundary &b) override;   // createBoundaryShaders()
    void resetTimestep(float dt) override;                         // resetTimestep() with a fixed dt
    void timestep();                                               // ShallowWaterEngine::timestep()
    void step(int n) override;
    void sync() override;
    void getState(float *state3) override;
    void getStateF64(double *state3) override;
    void setState(const float *state3) override;
    double time() const { return (double)total_time_; }
    double timeValue() const override { return (double)total_time_; }
    int scheme() const override { return time_scheme_now_; }
    float timedSteps(int n) override;
    void setProfiling(bool on) override;
    bool kernelTime(const std::string &name, float *ms, int *launches) override;
    size_t deviceBytes() const override { return arena_bytes_ + table_bytes_; }
    bool compensated() const { return comp_; }
    const cel_settings &settings() const { return s_; }
    // the explicit kernel's fast variant (shortened exact division and square root) applies
    bool fastArith() const { return fast_mesh_ && fast_allowed_; }
    // device bytes an engine with these settings holds after creation (admission control)
    static size_t estimateBytes(const cel_settings &s, const cel_boundary bnd[4]);
    void coefficients(float *ax, float *ay) override;
    void stats(cel_stats *out) override;
    void resetTimestepAdaptive(const cel_adaptive &a, cel_stats *out) override;
    void setGraphSteps(int k) override;                       // steps per captured block (0: off)
    // plans the next step (kernel parameters from the current host state) and, if advance,
    // moves the host state on as the step will
    void planStep(StepLaunch &L, bool advance = true);
    // tile shape and line-solve shapes: timed choice, current choice, caller's choice
    void autotune(int repeats, cel_tuning *out) override;
    void getTuning(cel_tuning *out) const override;
    void setTuning(const cel_tuning &t) override;
    // line-solve block shapes (indices into kShapeC / kShapeW) as a batch chooses them
    int lineLength() const { return nx_ > ny_ ? nx_ : ny_; }     // the longer line direction
    int lineLengthX() const { return nx_; }
    int lineLengthY() const { return ny_; }
    int shapeX() const { return shape_x_; }
    int shapeY() const { return shape_y_; }
    void setLineShapes(int sx, int sy);
    // the numeric sentinel: nonzero once a step produced a non-finite value (waits for the
    // stream); setState clears it
    int numericStatus() override;
    // everything the next steps read (state, history ring, corrector copies, visual record,
    // scheme counters, dt history, time, time-series cursors) as one buffer: writes it into
    // buf (if buf and cap suffice) and returns its size; restore continues bit-identically
    size_t checkpoint(void *buf, size_t cap) override;
    void restore(const void *buf, size_t bytes) override;
    int *statusWord() const { return d_status_; }
    void sample(const int *xy, int n, float *out3) override;
    int timestepCount() const override { return timestep_count_; }
    void setVisual(bool on, float threshold, float decay) override;
    void getAux(float *aux3) override;
    void getNormals(float *normal3) override;
    void addSolitaryWave(float H, float theta_deg, float x0, float y0) override;
    void setBottom(const float *bottom3, bool keep_depth) override;
    void raiseTerrain(float world_x, float world_y, float radius, float displacement) override;
    void shiftTerrain(float shift_up) override;
    void editWater(float world_x, float world_y, float radius, float disp_A, float disp_B) override;

private:
    void allocate();
    void uploadBottom(const float *bottom3);
    void uploadHostBottom();
    // one host plane of X x Y values (contiguous) into a device plane: one contiguous upload
    // into the x-line right-hand-side buffer (unused between steps), then a pitched copy
    // (stagePlane) or a transpose (stagePlaneT, column pitch PT) on the device
    void stagePlane(real *dst, const real *src);
    void stagePlaneT(real *dstT, const real *src);
    void bottomChanged();
    void downloadW(std::vector<real> &w);
    void uploadW(const std::vector<real> &w);
    void resetScheme();
    void buildSideTables(int side);
    void freeSideTables(int side);
    void validateBoundaries() const;
    void fillExplicit(ExplicitParams &p) const;
    int default_band_rows(int device) const;
    void fillLines(LineParams &p) const;
    void fillStatic(StaticParams &p) const;
    void check(hipError_t e, const char *what) const;
    // the device planes a checkpoint holds, in its order
    std::vector<std::pair<real *, size_t>> checkpointPlanes() const;
    void issueStep(const StepLaunch &L, const real *dclock, bool profile);

    // captured multi-step blocks
    struct GraphKey {
        int v[11] = {};
        bool operator==(const GraphKey &o) const { return std::equal(v, v + 11, o.v); }
    };
    struct GraphEntry { GraphKey key; hipGraphExec_t exec; };
    std::vector<GraphEntry> graphs_;
    int graph_steps_ = 0;
    real *d_clock_ = nullptr;
    real *h_clock_[2] = {};
    hipEvent_t clock_ev_[2] = {};
    bool clock_used_[2] = {};
    int clock_slot_ = 0;
    GraphKey discreteKey() const;
    void graphBlock();
    void clearGraphs();
    void timeSeries(int side, real t, real &eta, real &hu, real &hv);

    cel_settings s_;
    SideState side_[4];
    int nx_, ny_, X_, Y_, P_, PT_, Lx_, Ly_;
    real dx_, dy_, PI_;
    hipStream_t stream_ = nullptr;
    int device_ = 0;

    // device arena
    char *arena_ = nullptr;
    size_t arena_bytes_ = 0, table_bytes_ = 0;
    real *w_[2], *hu_[2], *hvT_[2];
    bool comp_ = false;                                       // compensated two-float state
    real *w_lo_[2] = {}, *hu_lo_[2] = {}, *hvT_lo_[2] = {};
    real *BY_, *BX_, *BA_;
    real *BAT_;                         // BA transposed (the y-lines' coefficients)
    real *Dx_, *DyT_;
    // Adams-Bashforth ring: per slot a quad plane (dU/dt of w, hu, hv and F*, four reals per
    // cell) and a G* plane; the predictor's copy of a slot in the same layout
    real *H_[3], *H1_[3];
    real *Hp_, *Hp1_;
    real *Cw_, *Chu_, *ChvT_;
    std::vector<real> hostBA_;
    std::vector<real> hostBottom_;      // BY, BX, BA planes, ghost-inclusive (terrain edits)

    // GetStats buffers and resetTimestep's adaptive-dt state (engine.cpp:1698-1726)
    double *d_stats_ = nullptr;     // partials, then the nine results
    double *h_stats_ = nullptr;     // pinned
    int stats_blocks_ = 0;
    float uniform_dt_ = 0.f;
    bool dt_initialized_ = false, target_cfl_set_ = false;
    float target_cfl_ = 0.f;
    void runStats(double *out9);
    void deriveStats(const double *raw, cel_stats *out) const;

    // visual outputs: auxiliary planes (max depth, slope angle, whitewater), own allocation
    bool aux_on_ = false;
    real *aux_mem_ = nullptr;
    real aux_threshold_ = 0, aux_rate_ = 0;      // threshold, whiteWaterDecayRate
    void fillAux(AuxParams &p) const;
    void zeroAux();

    // explicit band rows and line-solve shapes (indices into kShapeC / kShapeW)
    int rows_ = kDefaultBandRows, shape_x_ = 0, shape_y_ = 0;
    float tuned_us_[3] = {};
    // the explicit kernel's FAST variant (shortened exact division and square root): the
    // mesh and settings admit it (static_fields' check, fastSettings) and the caller allows it
    int *d_flag_ = nullptr;
    int *d_status_ = nullptr;        // numeric sentinel of the explicit kernel (ExplicitParams::status)
    bool fast_mesh_ = false, fast_allowed_ = true;
    bool fastSettings() const;

    // scheme state (engine.cpp:3254-3269, 1476-1519)
    int cur_ = 0;
    int ring_old_ = 0, ring_oo_ = 1, ring_scr_ = 2;
    int time_scheme_now_ = CEL_EULER;
    int timestep_count_ = 1;
    int correction_steps_count_ = 0;
    real total_time_ = 0;
    real current_timestep_, old_timestep_, old_old_timestep_;

    // profiling: five events per step recorded without host synchronisation and
    // reduced after the stream drains (collectProfile)
    bool profiling_ = false;
    struct KTime { float ms = 0.f; int n = 0; };
    KTime kt_[3];
    hipEvent_t ev_[8] = {};
    std::vector<hipEvent_t> pool_;
    size_t pool_used_ = 0;
    hipEvent_t nextEvent();
    void collectProfile();
};

}  // namespace CEL_NS

#endif
