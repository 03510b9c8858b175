/*
 * celeris_oracle.h -- CPU restatement of Celeris's Boussinesq time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle for the
 * HIP implementation under celeris_amd/.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product never links it.
 *
 * It is a scalar transliteration of the pixel-shader passes in
 * msvc/Release/shaders/compute.hlsl and of the pass sequencing in
 * ShallowWaterEngine::timestep() (engine.cpp:849-1522), instantiated in two
 * precisions: FP32 (same operation order as the HLSL, built with
 * -ffp-contract=off; the parity judge) and FP64 (the accuracy judge).
 * Every texture of the reference is kept as an RGBA array of the same size,
 * so quirks of the reference (ghost zeroing, strip copy-back order, the
 * gradient ring rotation) are reproduced literally.
 *
 * Array layout of every plane exchanged through this header: row-major,
 * ghost-inclusive, (nx+4) columns by (ny+4) rows, x fastest.
 */
#ifndef CELERIS_ORACLE_H
#define CELERIS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Boundary types, settings.hpp:90-96 (BoundaryType). */
enum { ORC_SOLID = 0, ORC_SPONGE = 1, ORC_SINEWAVE = 2, ORC_IRREGULAR = 3, ORC_UTS = 4 };
/* Sides, settings.hpp (BoundarySide). */
enum { ORC_WEST = 0, ORC_EAST = 1, ORC_SOUTH = 2, ORC_NORTH = 3 };
/* Time schemes, settings.hpp:304 (TimeIntegrationScheme). */
enum { ORC_EULER = 0, ORC_PREDICTOR = 1, ORC_CORRECTOR = 2, ORC_PREDICTOR_ADAPTIVE = 3 };

/* BoundarySetting, settings.hpp:149-159. */
typedef struct {
    int type;
    int width;
    float water_level;
    float amplitude, period, theta;   /* SineWaveSetting */
    int n_waves;                      /* irregular table: (amp, period, theta, phase) x n_waves */
    const float *waves;
    int n_uts;                        /* uniform time series rows: (t, eta, hu, hv) x n_uts */
    const float *uts;
} orc_boundary;

/* The InitSetting fields the step reads (settings.hpp:324-351). */
typedef struct {
    int nx, ny;
    float width, length;              /* W (x), L (y); dx = W/(nx-1) */
    float still_water_elevation;
    int is_boussinesq;
    float epsilon;
    float theta;                      /* minmod theta; TWO_THETA = 2*theta */
    float friction;
    int is_manning;
    int correction_steps_num;
    float timestep;                   /* fixed dt ("uniform time step") */
    int time_scheme;                  /* ORC_PREDICTOR or ORC_PREDICTOR_ADAPTIVE */
    float gravity;
    orc_boundary bnd[4];              /* indexed by ORC_WEST..ORC_NORTH */
} orc_config;

/* OpenMP threads of the row loops (timing only). */
void orc_set_threads(int n);
int orc_get_threads(void);

typedef struct orc_sim_f32 orc_sim_f32;
typedef struct orc_sim_f64 orc_sim_f64;

/* bottom: three planes BY, BX, BA (each (nx+4)*(ny+4)).
 * state : three planes w, hu, hv.  Returns NULL on bad input. */
orc_sim_f32 *orc_f32_create(const orc_config *cfg, const float *bottom, const float *state);
void orc_f32_destroy(orc_sim_f32 *s);
void orc_f32_step(orc_sim_f32 *s, int nsteps);
/* Sets dt for the following steps (resetTimestep with an externally chosen dt). */
void orc_f32_set_dt(orc_sim_f32 *s, float dt);
void orc_f32_get_state(const orc_sim_f32 *s, float *state3);
double orc_f32_time(const orc_sim_f32 *s);
int orc_f32_scheme(const orc_sim_f32 *s);
/* Level-0 tridiagonal coefficients (A, B, C planes) for x and y, engine.cpp:2638-2694, 2781-2826. */
void orc_f32_coefficients(const orc_sim_f32 *s, float *abc_x3, float *abc_y3);

/* GetStats (compute.hlsl:1319-1378) of the current state: sums in double over the
 * reference's 4x4-block domain; see oracle_impl.inc for the nine values. */
void orc_f32_stats(const orc_sim_f32 *s, double *out9);
/* resetTimestep's dt choice (engine.cpp:1698-1729); returns the new dt. */
float orc_f32_update_timestep(orc_sim_f32 *s, float ema_alpha, float min_ratio, float target_cfl, float uniform_dt,
                              double *stats9);

/* Visual outputs of Pass1/Pass2 (compute.hlsl:437-457, 535-540).  set_visual sets the
 * dissipation threshold and the "Whitewater Decay" setting (defaults 0.25 and 0.1, the
 * CML defaults); get_aux returns m_psAuxiliary1Texture as three planes (max inundated
 * depth, slope angle, whitewater), get_normals the normal texture (nx, ny, nz). */
void orc_f32_set_visual(orc_sim_f32 *s, float dissipation_threshold, float whitewater_decay);
void orc_f32_get_aux(const orc_sim_f32 *s, float *out3);
void orc_f32_get_normals(const orc_sim_f32 *s, float *out3);
/* addSolitaryWave (engine.cpp:2154-2220); refresh != 0 re-runs Pass1 on the new state. */
void orc_f32_add_solitary_wave(orc_sim_f32 *s, float H, float theta_deg, float x0, float y0, int refresh);
/* Pass1 of the current state (as the bootstrap Pass1), refreshing H, U, V and the normal. */
void orc_f32_refresh(orc_sim_f32 *s);
/* Terrain and water edits: fillTerrainTexture(Lite), raiseLowerTerrain, shiftTerrainSlider
 * and the left-mouse water brush (see oracle_impl.inc).  refresh != 0 re-runs Pass1 on the
 * edited state; the reference does not, so its next step's fluxes still see the old state. */
void orc_f32_set_bottom(orc_sim_f32 *s, const float *bottom3, int keep_depth);
void orc_f32_raise_terrain(orc_sim_f32 *s, float x, float y, float radius, float displacement, int refresh);
void orc_f32_shift_terrain(orc_sim_f32 *s, float shift_up, int refresh);
void orc_f32_edit_water(orc_sim_f32 *s, float x, float y, float radius, float disp_A, float disp_B, int refresh);
/* The bottom planes (BY, BX, BA) as the oracle holds them. */
void orc_f32_get_bottom(const orc_sim_f32 *s, float *bottom3);

orc_sim_f64 *orc_f64_create(const orc_config *cfg, const float *bottom, const float *state);
void orc_f64_destroy(orc_sim_f64 *s);
void orc_f64_step(orc_sim_f64 *s, int nsteps);
void orc_f64_set_dt(orc_sim_f64 *s, float dt);
void orc_f64_get_state(const orc_sim_f64 *s, double *state3);
double orc_f64_time(const orc_sim_f64 *s);
int orc_f64_scheme(const orc_sim_f64 *s);
void orc_f64_coefficients(const orc_sim_f64 *s, double *abc_x3, double *abc_y3);
void orc_f64_stats(const orc_sim_f64 *s, double *out9);
float orc_f64_update_timestep(orc_sim_f64 *s, float ema_alpha, float min_ratio, float target_cfl, float uniform_dt,
                              double *stats9);
void orc_f64_set_visual(orc_sim_f64 *s, float dissipation_threshold, float whitewater_decay);
void orc_f64_get_aux(const orc_sim_f64 *s, double *out3);
void orc_f64_get_normals(const orc_sim_f64 *s, double *out3);
void orc_f64_add_solitary_wave(orc_sim_f64 *s, float H, float theta_deg, float x0, float y0, int refresh);
void orc_f64_refresh(orc_sim_f64 *s);
void orc_f64_set_bottom(orc_sim_f64 *s, const float *bottom3, int keep_depth);
void orc_f64_raise_terrain(orc_sim_f64 *s, float x, float y, float radius, float displacement, int refresh);
void orc_f64_shift_terrain(orc_sim_f64 *s, float shift_up, int refresh);
void orc_f64_edit_water(orc_sim_f64 *s, float x, float y, float radius, float disp_A, float disp_B, int refresh);
void orc_f64_get_bottom(const orc_sim_f64 *s, double *bottom3);

/* Terrain preprocessing, terrain_heightfield.cpp:244-298: corner heights
 * z[(ny+5)*(nx+5)] -> bottom planes BY, BX, BA. */
void orc_bottom_from_corners(int nx, int ny, const float *z, float *bottom3);

/* Tridiagonal cyclic reduction exactly as the reference performs it
 * (coefficient pyramid engine.cpp:2696-2749 + CyclicReduceDx/CyclicSubstituteXx,
 * compute.hlsl:835-918) for ONE line of n interior unknowns; A[0] and C[n-1]
 * are taken as given.  Used by tests to pin the solver against a direct solve. */
void orc_f32_cr_solve(int n, const float *a, const float *b, const float *c, const float *d, float *x);
void orc_f64_cr_solve(int n, const double *a, const double *b, const double *c, const double *d, double *x);

/* Per-function entry points used by the known-answer tests (compute.hlsl helpers). */
float orc_f32_minmod(float a, float b, float c);
void orc_f32_reconstruct(float two_theta, float west, float here, float east, float *out_west, float *out_east);
void orc_f32_correct_w(float B_west, float B_east, float w_bar, float *w_west, float *w_east);
float orc_f32_divide_by_h(float h, float epsilon);
float orc_f32_numerical_flux(float aplus, float aminus, float Fplus, float Fminus, float Udiff);
float orc_f32_sponge(float L, float x);
float orc_f32_wavenumber(float omega, float d, float g);
void orc_f32_ab_coef(float dt, float dt1, float dt2, float *coef3);
/* Time integration (compute.hlsl:569-605, 755-798): the AB3 and AM4 increments at
 * constant dt and the three second-order derivative estimates at unequal steps. */
float orc_f32_ab3_increment(float dt, float f0, float fm1, float fm2);
double orc_f64_ab3_increment(double dt, double f0, double fm1, double fm2);
float orc_f32_am4_increment(float dt, float f1, float f0, float fm1, float fm2);
double orc_f64_am4_increment(double dt, double f1, double f0, double fm1, double fm2);
void orc_f32_second_order_derivatives(float y_n, float y_n1, float y_n2, float dt_n1, float dt_n2, float *der3);
void orc_f64_second_order_derivatives(double y_n, double y_n1, double y_n2, double dt_n1, double dt_n2, double *der3);
/* Uniform time series (timeSeriesHelper, engine.cpp:1887-1904) over n rows (t, eta, hu,
 * hv) with the caller's cursor. */
void orc_f32_time_series(const float *rows4, int n, int *iter, float t, float *out3);
void orc_f64_time_series(const float *rows4, int n, int *iter, double t, double *out3);
/* The (w, hu, hv) planes one irregular-wave side writes at the simulation's time. */
void orc_f32_irregular_strip(orc_sim_f32 *s, int side, float *out3);
void orc_f64_irregular_strip(orc_sim_f64 *s, int side, double *out3);

#ifdef __cplusplus
}
#endif
#endif
