// engine_api.hpp -- the precision-independent face of an engine, used by the C ABI.
//
// ShallowWaterEngine exists once per working precision (namespace cel: FP32, namespace
// cel64: FP64); both implement EngineApi, so the C entry points drive either through one
// pointer.  Exchanged arrays are FP32 (the reference's textures) except the FP64 state
// read-back.
#pragma once

#include <stdexcept>
#include <string>

#include "../../include/celeris.h"

namespace cel {

struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string &m) : std::runtime_error(m), status(s) {}
};

class EngineApi {
public:
    virtual ~EngineApi() = default;
    virtual void setBoundary(int side, const cel_boundary &b) = 0;
    virtual void resetTimestep(float dt) = 0;
    virtual void step(int n) = 0;
    virtual void sync() = 0;
    virtual void getState(float *state3) = 0;
    virtual void getStateF64(double *state3) = 0;
    virtual void setState(const float *state3) = 0;
    virtual double timeValue() const = 0;
    virtual int scheme() const = 0;
    virtual float timedSteps(int n) = 0;
    virtual void setProfiling(bool on) = 0;
    virtual bool kernelTime(const std::string &name, float *ms, int *launches) = 0;
    virtual size_t deviceBytes() const = 0;
    virtual void coefficients(float *ax, float *ay) = 0;
    virtual void stats(cel_stats *out) = 0;
    virtual void resetTimestepAdaptive(const cel_adaptive &a, cel_stats *out) = 0;
    virtual void setGraphSteps(int k) = 0;
    virtual void autotune(int repeats, cel_tuning *out) = 0;
    virtual void getTuning(cel_tuning *out) const = 0;
    virtual void setTuning(const cel_tuning &t) = 0;
    virtual void sample(const int *xy, int n, float *out3) = 0;
    virtual int timestepCount() const = 0;
    virtual void setVisual(bool on, float threshold, float decay) = 0;
    virtual void getAux(float *aux3) = 0;
    virtual void getNormals(float *normal3) = 0;
    virtual void addSolitaryWave(float H, float theta_deg, float x0, float y0) = 0;
    virtual void setBottom(const float *bottom3, bool keep_depth) = 0;
    virtual void raiseTerrain(float world_x, float world_y, float radius, float displacement) = 0;
    virtual void shiftTerrain(float shift_up) = 0;
    virtual void editWater(float world_x, float world_y, float radius, float disp_A, float disp_B) = 0;
};

}  // namespace cel
