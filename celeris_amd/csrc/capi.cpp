// capi.cpp -- extern "C" entry points (include/celeris.h).  Exceptions never cross the ABI:
// each call maps cel::Error / std::bad_alloc to a status and records the message.
#include <algorithm>
#include <cstring>
#include <new>
#include <string>

#include "batch.hpp"
#include "engine.hpp"
#define CEL_FP64
#include "engine.hpp"
#undef CEL_FP64

namespace cel {
void bottom_from_corners(int nx, int ny, const float *z, float *bottom3);
void initial_state(const cel_settings &s, const float *bottom3, const cel_soliton *sol, int nsol, float *state3,
                   const float *base3);
float probe_bandwidth(int kind, size_t bytes, int launches);
float probe_stencil_pattern(int nx, int ny, int rows, int layout, int launches);
float probe_solve_pattern(int lines, int elim, int launches);
hipError_t stats_noalloc(int nx, int ny, int pitch, const float *w, const float *hu, const float *hv, const float *BA,
                         float g, float one_over_dx, float one_over_dy, float epsilon, double *out9, hipStream_t st);
}

struct cel_engine {
    cel::EngineApi *impl;
};

struct cel_batch {
    cel::Batch *impl;
};

namespace {
thread_local std::string g_last_error;

template <class F>
int guarded(F &&f)
{
    try {
        g_last_error.clear();
        f();
        return CEL_OK;
    } catch (const cel::Error &e) {
        g_last_error = e.what();
        return e.status;
    } catch (const std::bad_alloc &) {
        g_last_error = "host allocation failed";
        return CEL_ENOMEM;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return CEL_EDEVICE;
    } catch (...) {
        g_last_error = "unknown error";
        return CEL_EDEVICE;
    }
}

cel::EngineApi &eng(cel_engine *e)
{
    if (!e || !e->impl) throw cel::Error(CEL_EINVAL, "null engine");
    return *e->impl;
}

cel::Batch &bat(cel_batch *b)
{
    if (!b || !b->impl) throw cel::Error(CEL_EINVAL, "null batch");
    return *b->impl;
}
}  // namespace

extern "C" {

const char *cel_last_error(void) { return g_last_error.c_str(); }

int cel_bottom_from_corners(int nx, int ny, const float *z, float *bottom3)
{
    return guarded([&] {
        if (nx < 4 || ny < 4 || !z || !bottom3) throw cel::Error(CEL_EINVAL, "bad arguments");
        cel::bottom_from_corners(nx, ny, z, bottom3);
    });
}

int cel_initial_state(const cel_settings *s, const float *bottom3, const cel_soliton *solitons, int n_solitons,
                      float *state3)
{
    return guarded([&] {
        if (!s || !bottom3 || !state3 || n_solitons < 0 || (n_solitons > 0 && !solitons))
            throw cel::Error(CEL_EINVAL, "bad arguments");
        if (s->nx < 4 || s->ny < 4) throw cel::Error(CEL_EINVAL, "nx and ny must be at least 4");
        cel::initial_state(*s, bottom3, solitons, n_solitons, state3, nullptr);
    });
}

int cel_hot_start_state(const cel_settings *s, const float *bottom3, const float *hot3, const cel_soliton *solitons,
                        int n_solitons, float *state3)
{
    return guarded([&] {
        if (!s || !bottom3 || !hot3 || !state3 || n_solitons < 0 || (n_solitons > 0 && !solitons))
            throw cel::Error(CEL_EINVAL, "bad arguments");
        if (s->nx < 4 || s->ny < 4) throw cel::Error(CEL_EINVAL, "nx and ny must be at least 4");
        cel::initial_state(*s, bottom3, solitons, n_solitons, state3, hot3);
    });
}

int cel_engine_sample(cel_engine *e, const int *xy, int n, float *out3)
{
    return guarded([&] { eng(e).sample(xy, n, out3); });
}

int cel_engine_timestep_count(cel_engine *e, int *count)
{
    return guarded([&] {
        if (!count) throw cel::Error(CEL_EINVAL, "null output");
        *count = eng(e).timestepCount();
    });
}

int cel_engine_set_visual(cel_engine *e, int enable, float dissipation_threshold, float whitewater_decay)
{
    return guarded([&] { eng(e).setVisual(enable != 0, dissipation_threshold, whitewater_decay); });
}

int cel_engine_get_aux(cel_engine *e, float *aux3)
{
    return guarded([&] { eng(e).getAux(aux3); });
}

int cel_engine_get_normals(cel_engine *e, float *normal3)
{
    return guarded([&] { eng(e).getNormals(normal3); });
}

int cel_engine_add_solitary_wave(cel_engine *e, float H, float theta_deg, float x0, float y0)
{
    return guarded([&] { eng(e).addSolitaryWave(H, theta_deg, x0, y0); });
}

int cel_engine_set_bottom(cel_engine *e, const float *bottom3, int keep_depth)
{
    return guarded([&] { eng(e).setBottom(bottom3, keep_depth != 0); });
}

int cel_engine_raise_terrain(cel_engine *e, float x, float y, float radius, float displacement)
{
    return guarded([&] { eng(e).raiseTerrain(x, y, radius, displacement); });
}

int cel_engine_shift_terrain(cel_engine *e, float shift_up)
{
    return guarded([&] { eng(e).shiftTerrain(shift_up); });
}

int cel_engine_edit_water(cel_engine *e, float x, float y, float radius, float disp_A, float disp_B)
{
    return guarded([&] { eng(e).editWater(x, y, radius, disp_A, disp_B); });
}

int cel_engine_create(const cel_settings *s, const cel_boundary bnd[4], const float *bottom3, const float *state3,
                      cel_engine **out)
{
    return guarded([&] {
        if (!s || !out) throw cel::Error(CEL_EINVAL, "null settings or output");
        *out = nullptr;
        auto *e = new cel_engine{nullptr};
        try {
            if (s->state_precision == CEL_STATE_FP64) e->impl = new cel64::ShallowWaterEngine(*s, bnd, bottom3, state3);
            else e->impl = new cel::ShallowWaterEngine(*s, bnd, bottom3, state3);
        } catch (...) {
            delete e;
            throw;
        }
        *out = e;
    });
}

int cel_engine_destroy(cel_engine *e)
{
    return guarded([&] {
        if (!e) return;
        delete e->impl;
        delete e;
    });
}

int cel_engine_set_boundary(cel_engine *e, int side, const cel_boundary *b)
{
    return guarded([&] {
        if (!b) throw cel::Error(CEL_EINVAL, "null boundary");
        eng(e).setBoundary(side, *b);
    });
}

int cel_engine_set_dt(cel_engine *e, float dt) { return guarded([&] { eng(e).resetTimestep(dt); }); }
int cel_engine_step(cel_engine *e, int nsteps) { return guarded([&] { eng(e).step(nsteps); }); }
int cel_engine_sync(cel_engine *e) { return guarded([&] { eng(e).sync(); }); }
int cel_engine_get_state(cel_engine *e, float *state3) { return guarded([&] { eng(e).getState(state3); }); }
int cel_engine_get_state_f64(cel_engine *e, double *state3) { return guarded([&] { eng(e).getStateF64(state3); }); }
int cel_engine_set_state(cel_engine *e, const float *state3) { return guarded([&] { eng(e).setState(state3); }); }

int cel_engine_time(cel_engine *e, double *t)
{
    return guarded([&] {
        if (!t) throw cel::Error(CEL_EINVAL, "null output");
        *t = eng(e).timeValue();
    });
}

int cel_engine_scheme(cel_engine *e, int *scheme)
{
    return guarded([&] {
        if (!scheme) throw cel::Error(CEL_EINVAL, "null output");
        *scheme = eng(e).scheme();
    });
}

int cel_engine_timed_steps(cel_engine *e, int nsteps, float *ms)
{
    return guarded([&] {
        if (!ms) throw cel::Error(CEL_EINVAL, "null output");
        *ms = eng(e).timedSteps(nsteps);
    });
}

int cel_engine_set_profiling(cel_engine *e, int on) { return guarded([&] { eng(e).setProfiling(on != 0); }); }

int cel_engine_kernel_time(cel_engine *e, const char *name, float *ms, int *launches)
{
    return guarded([&] {
        if (!name) throw cel::Error(CEL_EINVAL, "null name");
        if (!eng(e).kernelTime(name, ms, launches)) throw cel::Error(CEL_EINVAL, std::string("unknown kernel ") + name);
    });
}

int cel_engine_device_bytes(cel_engine *e, size_t *bytes)
{
    return guarded([&] {
        if (!bytes) throw cel::Error(CEL_EINVAL, "null output");
        *bytes = eng(e).deviceBytes();
    });
}

int cel_engine_coefficients(cel_engine *e, float *abc_x3, float *abc_y3)
{
    return guarded([&] {
        if (!abc_x3 || !abc_y3) throw cel::Error(CEL_EINVAL, "null output");
        eng(e).coefficients(abc_x3, abc_y3);
    });
}

int cel_estimate_device_bytes(const cel_settings *s, const cel_boundary bnd[4], size_t *bytes)
{
    return guarded([&] {
        if (!s || !bnd || !bytes) throw cel::Error(CEL_EINVAL, "null argument");
        *bytes = s->state_precision == CEL_STATE_FP64 ? cel64::ShallowWaterEngine::estimateBytes(*s, bnd)
                                                      : cel::ShallowWaterEngine::estimateBytes(*s, bnd);
    });
}

int cel_batch_create(cel_batch **out)
{
    return guarded([&] {
        if (!out) throw cel::Error(CEL_EINVAL, "null output");
        *out = nullptr;
        auto *b = new cel_batch{nullptr};
        try {
            b->impl = new cel::Batch();
        } catch (...) {
            delete b;
            throw;
        }
        *out = b;
    });
}

int cel_batch_destroy(cel_batch *b)
{
    return guarded([&] {
        if (!b) return;
        delete b->impl;
        delete b;
    });
}

int cel_batch_add(cel_batch *b, const cel_settings *s, const cel_boundary bnd[4], const float *bottom3,
                  const float *state3, int *id)
{
    return guarded([&] {
        if (!s || !id) throw cel::Error(CEL_EINVAL, "null settings or output");
        *id = bat(b).add(*s, bnd, bottom3, state3);
    });
}

int cel_batch_remove(cel_batch *b, int id) { return guarded([&] { bat(b).remove(id); }); }
int cel_batch_set_paused(cel_batch *b, int id, int paused) { return guarded([&] { bat(b).setPaused(id, paused != 0); }); }
int cel_batch_step(cel_batch *b, int nsteps) { return guarded([&] { bat(b).step(nsteps); }); }
int cel_batch_set_graph_steps(cel_batch *b, int steps) { return guarded([&] { bat(b).setGraphSteps(steps); }); }
int cel_batch_sync(cel_batch *b) { return guarded([&] { bat(b).sync(); }); }

int cel_batch_timed_steps(cel_batch *b, int nsteps, float *ms)
{
    return guarded([&] {
        if (!ms) throw cel::Error(CEL_EINVAL, "null output");
        *ms = bat(b).timedSteps(nsteps);
    });
}

int cel_batch_count(cel_batch *b, int *n)
{
    return guarded([&] {
        if (!n) throw cel::Error(CEL_EINVAL, "null output");
        *n = bat(b).count();
    });
}

int cel_batch_device_bytes(cel_batch *b, size_t *bytes)
{
    return guarded([&] {
        if (!bytes) throw cel::Error(CEL_EINVAL, "null output");
        *bytes = bat(b).deviceBytes();
    });
}

int cel_engine_checkpoint(cel_engine *e, void *buf, size_t *bytes)
{
    return guarded([&] {
        if (!bytes) throw cel::Error(CEL_EINVAL, "null size");
        const size_t need = eng(e).checkpoint(buf, buf ? *bytes : 0);
        if (buf && *bytes < need) {
            *bytes = need;
            throw cel::Error(CEL_EINVAL, "checkpoint buffer too small");
        }
        *bytes = need;
    });
}

int cel_engine_restore(cel_engine *e, const void *buf, size_t bytes)
{
    return guarded([&] { eng(e).restore(buf, bytes); });
}

int cel_engine_check_numeric(cel_engine *e)
{
    return guarded([&] {
        if (eng(e).numericStatus() != 0)
            throw cel::Error(CEL_ENUMERIC, "a step produced a non-finite surface elevation or line right-hand side");
    });
}

int cel_batch_check_numeric(cel_batch *b, int *ids, int max_ids, int *n_failed)
{
    return guarded([&] {
        if (max_ids < 0 || (max_ids > 0 && !ids)) throw cel::Error(CEL_EINVAL, "bad id buffer");
        const std::vector<int> f = bat(b).failed();
        for (int i = 0; i < (int)f.size() && i < max_ids; ++i) ids[i] = f[i];
        if (n_failed) *n_failed = (int)f.size();
    });
}

int cel_batch_adopt(cel_batch *b, cel_engine *e, int *id)
{
    return guarded([&] {
        cel::Batch &bt = bat(b);
        if (!e || !e->impl) throw cel::Error(CEL_EINVAL, "null engine");
        auto *fp32 = dynamic_cast<cel::ShallowWaterEngine *>(e->impl);
        if (!fp32) throw cel::Error(CEL_EINVAL, "batched scenarios run in FP32 (plain or compensated)");
        if (fp32->settings().correction_steps_num != 0)
            throw cel::Error(CEL_EINVAL, "batched scenarios do not take corrector steps");
        std::unique_ptr<cel::ShallowWaterEngine> up(fp32);
        int v = 0;
        try {
            v = bt.adopt(up);
        } catch (...) {
            up.release();                        // still the caller's
            throw;
        }
        e->impl = nullptr;                       // the batch owns it from here on
        if (id) *id = v;
    });
}

int cel_batch_set_state(cel_batch *b, int id, const float *state3)
{
    return guarded([&] {
        if (!state3) throw cel::Error(CEL_EINVAL, "null state");
        bat(b).setState(id, state3);
    });
}

int cel_batch_get_state(cel_batch *b, int id, float *state3)
{
    return guarded([&] { bat(b).scenario(id).getState(state3); });
}

int cel_batch_get_state_f64(cel_batch *b, int id, double *state3)
{
    return guarded([&] { bat(b).scenario(id).getStateF64(state3); });
}

int cel_batch_stats(cel_batch *b, int id, cel_stats *out)
{
    return guarded([&] { bat(b).scenario(id).stats(out); });
}

int cel_batch_time(cel_batch *b, int id, double *t)
{
    return guarded([&] {
        if (!t) throw cel::Error(CEL_EINVAL, "null output");
        *t = bat(b).scenario(id).time();
    });
}

int cel_engine_autotune(cel_engine *e, int repeats, cel_tuning *out)
{
    return guarded([&] { eng(e).autotune(repeats, out); });
}

int cel_engine_get_tuning(cel_engine *e, cel_tuning *out) { return guarded([&] { eng(e).getTuning(out); }); }

int cel_engine_set_tuning(cel_engine *e, const cel_tuning *t)
{
    return guarded([&] {
        if (!t) throw cel::Error(CEL_EINVAL, "null tuning");
        eng(e).setTuning(*t);
    });
}

int cel_batch_autotune(cel_batch *b, int repeats, int tune_lines, int *band_rows)
{
    return guarded([&] {
        const int r = bat(b).autotune(repeats, tune_lines != 0);
        if (band_rows) *band_rows = r;
    });
}

int cel_engine_set_graph_steps(cel_engine *e, int steps) { return guarded([&] { eng(e).setGraphSteps(steps); }); }

int cel_engine_stats(cel_engine *e, cel_stats *out) { return guarded([&] { eng(e).stats(out); }); }

int cel_engine_reset_timestep(cel_engine *e, const cel_adaptive *a, cel_stats *out)
{
    return guarded([&] {
        if (!a) throw cel::Error(CEL_EINVAL, "null adaptive settings");
        if (!(a->ema_alpha >= 0.f && a->ema_alpha <= 1.f)) throw cel::Error(CEL_EINVAL, "ema_alpha must lie in [0, 1]");
        if (!(a->min_timestep_ratio > 0.f)) throw cel::Error(CEL_EINVAL, "min_timestep_ratio must be positive");
        eng(e).resetTimestepAdaptive(*a, out);
    });
}

int cel_stats_noalloc(const cel_settings *s, const float *w, const float *hu, const float *hv, const float *BA,
                      int pitch, double *sums9, void *stream)
{
    return guarded([&] {
        if (!s || !w || !hu || !hv || !BA || !sums9) throw cel::Error(CEL_EINVAL, "null argument");
        if (s->nx < 4 || s->ny < 4 || pitch < s->nx + 4) throw cel::Error(CEL_EINVAL, "bad grid size or pitch");
        if (!(s->width > 0.f) || !(s->length > 0.f) || !(s->gravity > 0.f))
            throw cel::Error(CEL_EINVAL, "width, length and gravity must be positive");
        // the engine's constants (fillConstantBuffers, engine.cpp:4099-4136; CalcEpsilon, engine.cpp:290-299)
        const float dx = s->width / (float)(s->nx - 1), dy = s->length / (float)(s->ny - 1);
        const float m = std::max(dx, dy);
        const float eps = std::min(s->epsilon, (m * m) * (m * m));
        const hipError_t e = cel::stats_noalloc(s->nx, s->ny, pitch, w, hu, hv, BA, s->gravity, 1.f / dx, 1.f / dy, eps,
                                                sums9, static_cast<hipStream_t>(stream));
        if (e == hipErrorInvalidValue)
            throw cel::Error(CEL_EINVAL, "grid above the 4096 x 4096 statistics scratch (16384 tiles of 64 x 16 cells)");
        if (e != hipSuccess) throw cel::Error(CEL_EDEVICE, std::string("stats: ") + hipGetErrorString(e));
    });
}

int cel_stats_from_sums(const cel_settings *s, float dt, const double *sums9, cel_stats *out)
{
    return guarded([&] {
        if (!s || !sums9 || !out) throw cel::Error(CEL_EINVAL, "null argument");
        if (s->nx < 4 || s->ny < 4) throw cel::Error(CEL_EINVAL, "bad grid size");
        cel::derive_stats(*s, (double)dt, sums9, out);
    });
}

int cel_probe_bandwidth(int kind, size_t bytes, int launches, float *ms)
{
    return guarded([&] {
        const float t = cel::probe_bandwidth(kind, bytes, launches);
        if (ms) *ms = t;
    });
}

int cel_probe_stencil_pattern(int nx, int ny, int band_rows, int layout, int launches, float *ms)
{
    return guarded([&] {
        const float t = cel::probe_stencil_pattern(nx, ny, band_rows, layout, launches);
        if (ms) *ms = t;
    });
}

int cel_probe_solve_pattern(int lines, int elim, int launches, float *ms)
{
    return guarded([&] {
        const float t = cel::probe_solve_pattern(lines, elim, launches);
        if (ms) *ms = t;
    });
}

}  // extern "C"
