// engine_api.hpp -- the precision-independent face of an engine, used by the C ABI.
//
// ShallowWaterEngine exists once per working precision (namespace cel: FP32, namespace
// cel64: FP64); both implement EngineApi, so the C entry points drive either through one
// pointer.  Exchanged arrays are FP32 (the reference's textures) except the FP64 state
// read-back.
#pragma once

#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

#include "../../include/celeris.h"

namespace cel {

// resetTimestep's post-processing of GetStats' nine reductions (engine.cpp:1686-1696), in
// FP64: density 1000 kg/m^3, cell area dx dy; dt is the next step's.
inline void derive_stats(const cel_settings &s, double dt, const double *r, cel_stats *o)
{
    std::memset(o, 0, sizeof(*o));
    o->sum_h = r[0]; o->sum_bh_h2 = r[1]; o->sum_hu = r[2]; o->sum_hv = r[3]; o->sum_hu2v2 = r[4];
    o->max_u2v2 = r[5]; o->max_h = r[6]; o->max_cfl_rate = r[7]; o->max_u2v2_h = r[8];
    const double density = 1000.0;
    const double area = (double)s.width / (double)(s.nx - 1) * ((double)s.length / (double)(s.ny - 1));
    const double g = s.gravity;
    o->mass = r[0] * density * area;
    o->x_momentum = r[2] * density * area;
    o->y_momentum = r[3] * density * area;
    o->kinetic_energy = 0.5 * density * area * r[4];
    o->potential_energy = g * density * area * r[1];
    o->total_energy = o->kinetic_energy + o->potential_energy;
    o->max_speed = std::sqrt(r[5]);
    o->max_depth = r[6];
    o->max_froude = std::sqrt(r[8] / g);
    o->timestep = (float)dt;
    o->cfl_number = r[7] * dt;
    o->cells = 16 * (s.nx / 4) * (s.ny / 4);
}

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
    virtual int numericStatus() = 0;
    virtual size_t checkpoint(void *buf, size_t cap) = 0;
    virtual void restore(const void *buf, size_t bytes) = 0;
};

}  // namespace cel
