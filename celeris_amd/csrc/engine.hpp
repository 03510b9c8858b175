// engine.hpp -- host side of the MI355X Celeris step: ShallowWaterEngine.
//
// Mirrors the numerical interface of the reference class (engine.hpp:58-97):
// construction + remesh, resetTimestep (fixed or caller-chosen dt), timestep(),
// boundary changes (createBoundaryShaders), plus state upload/download.  All device
// resources are owned here; the D3D11 device/context pair of the reference is replaced
// by one HIP stream per engine.
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/celeris.h"
#include "params.hpp"

namespace cel {

struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string &m) : std::runtime_error(m), status(s) {}
};

hipError_t launch_explicit(int scheme, const ExplicitParams &p, hipStream_t st);
hipError_t launch_solve(const LineParams &p, const IrregularAll &irr, hipStream_t st);
hipError_t launch_boundary(const BoundaryParams &p, hipStream_t st);
hipError_t launch_static_fields(const StaticParams &p, hipStream_t st);
hipError_t launch_export_coefficients(const StaticParams &p, float *ax, float *ay, hipStream_t st);

// st_getCountOfMatrices, engine.cpp:741-748
int count_of_matrices(int n);

// Host copy of one side's BoundarySetting plus the per-strip-cell tables built from it.
struct SideState {
    cel_boundary cfg{};
    std::vector<float> waves;     // irregular table (amp, period, theta, phase)
    std::vector<float> uts;       // (t, eta, hu, hv)
    int uts_iter = 1;             // UniformTimeSeries::iter
    float *d_tab = nullptr;       // sponge gamma / sine table / irregular table
    float *d_irr_out = nullptr;   // irregular strip values
    float *d_irr_wave = nullptr;  // amp, omega, phase per component
    unsigned char *d_wet = nullptr;
    float sine_omega = 0.f;
    int irr_nw = 0;
};

class ShallowWaterEngine {
public:
    ShallowWaterEngine(const cel_settings &s, const cel_boundary bnd[4], const float *bottom3, const float *state3);
    ~ShallowWaterEngine();
    ShallowWaterEngine(const ShallowWaterEngine &) = delete;
    ShallowWaterEngine &operator=(const ShallowWaterEngine &) = delete;

    void setBoundary(int side, const cel_boundary &b);       // createBoundaryShaders()
    void resetTimestep(float dt);                             // resetTimestep() with a fixed dt
    void timestep();                                          // ShallowWaterEngine::timestep()
    void step(int n);
    void sync();
    void getState(float *state3);
    void setState(const float *state3);
    double time() const { return total_time_; }
    int scheme() const { return time_scheme_now_; }
    float timedSteps(int n);
    void setProfiling(bool on);
    bool kernelTime(const std::string &name, float *ms, int *launches);
    size_t deviceBytes() const { return arena_bytes_ + table_bytes_; }
    void coefficients(float *ax, float *ay);

private:
    void allocate();
    void uploadBottom(const float *bottom3);
    void resetScheme();
    void buildSideTables(int side);
    void freeSideTables(int side);
    void validateBoundaries() const;
    void fillExplicit(ExplicitParams &p) const;
    void fillLines(LineParams &p) const;
    void fillStatic(StaticParams &p) const;
    void check(hipError_t e, const char *what) const;
    void timeSeries(int side, float t, float &eta, float &hu, float &hv);

    cel_settings s_;
    SideState side_[4];
    int nx_, ny_, X_, Y_, P_, PT_, Lx_, Ly_;
    float dx_, dy_, PI_;
    hipStream_t stream_ = nullptr;
    int device_ = 0;

    // device arena
    char *arena_ = nullptr;
    size_t arena_bytes_ = 0, table_bytes_ = 0;
    float *w_[2], *hu_[2], *hvT_[2];
    float *BY_, *BX_, *BA_;
    float *Kx_, *Qx_, *Ky_, *Qy_, *DDx_, *DDy_, *KyT_, *QyT_;   // static per-cell fields
    float *Dx_, *DyT_;
    float *G_[3][3];          // ring slot, component
    float *FG_[3][2];
    float *Gp_[3], *FGp_[2];
    float *Cw_, *Chu_, *ChvT_;
    std::vector<float> hostBA_;

    // scheme state (engine.cpp:3254-3269, 1476-1519)
    int cur_ = 0;
    int ring_old_ = 0, ring_oo_ = 1, ring_scr_ = 2;
    int time_scheme_now_ = CEL_EULER;
    int timestep_count_ = 1;
    int correction_steps_count_ = 0;
    float total_time_ = 0.f;
    float current_timestep_, old_timestep_, old_old_timestep_;

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

}  // namespace cel
