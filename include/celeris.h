/*
 * celeris.h -- C ABI of the MI355X-native Celeris Boussinesq time step.
 *
 * Drop-in for the numerical path of ShallowWaterEngine (engine.hpp:58-97,
 * engine.cpp:849-1522): the step, its settings and its boundary-condition
 * interface.  Plain pointers and sizes only; nothing throws across this ABI.
 * Every call returns a cel_status; cel_last_error() describes the last failure
 * on the calling thread.
 *
 * Array convention (host side): every plane is row-major, ghost-inclusive,
 * (nx+4) columns by (ny+4) rows, x fastest -- the reference's texture layout
 * (engine.cpp:3024-3034).  State planes are w, hu, hv; bottom planes are
 * BY (north face), BX (east face), BA (centre), terrain_heightfield.cpp:254-256.
 */
#ifndef CELERIS_H
#define CELERIS_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  Replace Coercri::DXError throws (engine.cpp:119-121, 345) and
 * std::abort on unknown settings (settings.cpp:289). */
typedef enum {
    CEL_OK = 0,
    CEL_EINVAL = 1,    /* bad argument or unsupported setting */
    CEL_ENOMEM = 2,    /* device or host allocation failed */
    CEL_EDEVICE = 3,   /* HIP runtime error */
    CEL_ENUMERIC = 4,  /* non-finite state detected */
    CEL_ESTATE = 5     /* call not valid in the engine's current state */
} cel_status;

/* BoundaryType, settings.hpp:90-96. */
typedef enum {
    CEL_SOLID = 0,
    CEL_SPONGE = 1,
    CEL_SINEWAVE = 2,
    CEL_IRREGULARWAVES = 3,
    CEL_UNIFORMTIMESERIES = 4
} cel_boundary_type;

/* BoundarySide, settings.hpp:98-103. */
typedef enum { CEL_WEST = 0, CEL_EAST = 1, CEL_SOUTH = 2, CEL_NORTH = 3 } cel_side;

/* TimeIntegrationScheme, settings.hpp:304. */
typedef enum {
    CEL_EULER = 0,
    CEL_PREDICTOR = 1,
    CEL_CORRECTOR = 2,
    CEL_PREDICTOR_ADAPTIVE = 3
} cel_time_scheme;

/* State storage and arithmetic.  CEL_STATE_FP32 is the reference's arithmetic.
 * CEL_STATE_COMPENSATED keeps (w, hu, hv) as two-float pairs (hi + lo): the explicit
 * update adds its increment with an exact two-sum, and the line systems are solved for
 * the increment of hu and hv, which is then added the same way, so increments far below
 * the state's FP32 resolution accumulate instead of rounding away.  Every other
 * operation reads the hi words in FP32.  Not combined with corrector steps.
 * CEL_STATE_FP64 runs every kernel and every host-side table in FP64 (the same
 * operations in double precision); arrays still cross the ABI as FP32 except through
 * cel_engine_get_state_f64.  Not available in batches.  No counterpart in the reference. */
typedef enum { CEL_STATE_FP32 = 0, CEL_STATE_COMPENSATED = 1, CEL_STATE_FP64 = 2 } cel_state_precision;


/* Maximum irregular-wave components per side: MAX_NUM_OF_IRREGULAR_WAVES,
 * engine.hpp:216 / compute.hlsl:111-114. */
#define CEL_MAX_IRREGULAR_WAVES 1000
/* Longest supported tridiagonal line (cells per side). */
#define CEL_MAX_LINE 8192

/* BoundarySetting, settings.hpp:149-159.  Tables are copied by the engine. */
typedef struct {
    int type;                 /* cel_boundary_type */
    int width;                /* strip width in cells (reference default 2) */
    float water_level;        /* waterLevel */
    float amplitude;          /* SineWaveSetting.amplitude */
    float period;             /* SineWaveSetting.period */
    float theta;              /* SineWaveSetting.theta (radians from +x) */
    int n_waves;              /* IrregularWaves: rows of (amplitude, period, theta, phase) */
    const float *waves;
    int n_uts;                /* UniformTimeSeries: rows of (t, eta, hu, hv) */
    const float *uts;
} cel_boundary;

/* The InitSetting fields the step reads (settings.hpp:324-351), g from the
 * "gravity" setting (gui_manager.cpp:107). */
typedef struct {
    int nx, ny;                   /* interior cells; arrays are (nx+4) x (ny+4) */
    float width, length;          /* W (x extent) and L (y extent): dx = W/(nx-1), dy = L/(ny-1) */
    float still_water_elevation;
    int is_boussinesq;
    float epsilon;
    float theta;                  /* generalized-minmod theta (1..2) */
    float friction;
    int is_manning;
    int correction_steps_num;
    float timestep;               /* fixed dt, "uniform time step" */
    int time_scheme;              /* CEL_PREDICTOR or CEL_PREDICTOR_ADAPTIVE */
    float gravity;
    int state_precision;          /* cel_state_precision; 0 = the reference's FP32 */
} cel_settings;

/* Soliton, settings.hpp:180-192. */
typedef struct {
    float param_h, theta, length, xc, yc;
} cel_soliton;

typedef struct cel_engine cel_engine;

/* The same initial condition starting from a hot-start state (engine.cpp:2943-2970): w,
 * hu, hv from hot3 (the file's values, ghost-inclusive planes), then the solitary waves
 * added on top. */
int cel_hot_start_state(const cel_settings *s, const float *bottom3, const float *hot3, const cel_soliton *solitons,
                        int n_solitons, float *state3);

/* Thread-local description of the last failed call (empty if none). */
const char *cel_last_error(void);

/* UpdateTerrainHeightfield, terrain_heightfield.cpp:244-298.  z holds the corner
 * heights, (nx+5) x (ny+5), corner (i, j) at x = (i-2.5)/(nx-1)*W - W/2,
 * y = (j-2.5)/(ny-1)*L.  Writes the three bottom planes BY, BX, BA. */
int cel_bottom_from_corners(int nx, int ny, const float *z, float *bottom3);

/* Initial condition of createSimTextures, engine.cpp:2940-3016 (no hot-start
 * file): w = max(SWL, BA) with clamped ghosts, plus solitary waves
 * (engine.cpp:103-112, taper 2978-2999).  Writes the three state planes. */
int cel_initial_state(const cel_settings *s, const float *bottom3, const cel_soliton *solitons,
                      int n_solitons, float *state3);

/* ShallowWaterEngine ctor + remesh(R_MESH) (engine.cpp:709-740): allocates the
 * device state, uploads bottom and state, builds the per-cell boundary tables and
 * resets the time scheme (Euler bootstrap, engine.cpp:3254-3269).
 * bnd[4] is indexed by cel_side. */
int cel_engine_create(const cel_settings *s, const cel_boundary bnd[4], const float *bottom3,
                      const float *state3, cel_engine **out);
int cel_engine_destroy(cel_engine *e);

/* BoundarySetting change + createBoundaryShaders (engine.cpp:2405-2453). */
int cel_engine_set_boundary(cel_engine *e, int side, const cel_boundary *b);

/* resetTimestep with a caller-chosen dt (engine.cpp:1714-1729, 1750). */
int cel_engine_set_dt(cel_engine *e, float dt);

/* nsteps x { resetTimestep (fixed dt); timestep() } (main.cpp:396-425). Asynchronous
 * on the engine's HIP stream. */
int cel_engine_step(cel_engine *e, int nsteps);

/* Multi-step capture: with steps > 0 (at most 1024), cel_engine_step runs whole blocks
 * of that many steps as one hipGraph launch each (captured once per starting state of
 * the ping-pong buffers, history ring and scheme counters, then replayed with that
 * block's step clock -- t^n and time-series values -- copied in beforehand), and the
 * remainder step by step.  Results are identical to step-by-step execution.  0 turns
 * capture off.  Not used while profiling is on.  No counterpart in the reference, whose
 * frame loop issues every pass from the host. */
int cel_engine_set_graph_steps(cel_engine *e, int steps);

/* Blocks until queued steps finish. */
int cel_engine_sync(cel_engine *e);

/* Copies the current state (w, hu, hv planes) to the host (synchronises). */
int cel_engine_get_state(cel_engine *e, float *state3);
/* The state as FP64 planes: hi + lo for a compensated engine, hi otherwise. */
int cel_engine_get_state_f64(cel_engine *e, double *state3);
/* (w, hu, hv) at n cells (xy[2k], xy[2k+1] = ghost-inclusive column and row) of the
 * current state, out3[3k..3k+2]; synchronises.  The gauge reads of the logger
 * (st_DumpToFile, engine.cpp:628-675) without copying the whole state. */
int cel_engine_sample(cel_engine *e, const int *xy, int n, float *out3);
/* Hot start (engine.cpp:2943-2970): replaces the state; the AB history and the
 * scheme counters are reset exactly as after remesh. */
int cel_engine_set_state(cel_engine *e, const float *state3);

/* Virtual time total_time (engine.cpp:1518) and the scheme of the next step. */
int cel_engine_time(cel_engine *e, double *t);
int cel_engine_scheme(cel_engine *e, int *scheme);
/* The reference's timestep_count (engine.cpp:1491-1495): 1 after creation or a hot start,
 * incremented by every step after which the next scheme is Euler or a predictor.  The
 * logger's schedule (a record when it is a multiple of logStep) is keyed on it. */
int cel_engine_timestep_count(cel_engine *e, int *count);

/* Visual outputs of the step (compute.hlsl:437-457, 535-540).  When enabled, every step
 * updates the auxiliary texture m_psAuxiliary1Texture from the step's starting state on
 * the cells [1, nx+2) x [1, ny+2): r = maximum over steps of the mean corrected face
 * depth (the inundation record dumped by dumpInundationToFile, engine.cpp:781-817),
 * b = atan of the largest w-slope deviation per metre, a = whitewater (1 where b exceeds
 * dissipation_threshold, decayed by (1 / (1 + 2 whitewater_decay))^dt every step).  The
 * reference defaults are 0.25 and 0.1 (main.cpp:971-977).  Enabling (or a hot start)
 * starts the record from zero; disabled engines do no extra work. */
int cel_engine_set_visual(cel_engine *e, int enable, float dissipation_threshold, float whitewater_decay);
/* The auxiliary planes (max inundated depth, slope angle, whitewater), each
 * (nx+4) x (ny+4) ghost-inclusive; CEL_EINVAL when visual outputs are off. */
int cel_engine_get_aux(cel_engine *e, float *aux3);
/* The normal texture of the current state (Pass1, compute.hlsl:437-445): planes nx, ny,
 * nz of normalize((w_W - w_E)/dx, (w_S - w_N)/dy, 2) on [1, nx+3) x [1, ny+3), 0 outside. */
int cel_engine_get_normals(cel_engine *e, float *normal3);
/* addSolitaryWave (engine.cpp:2154-2220, compute.hlsl:947-976): adds the solitary wave
 * (H, direction theta_deg in degrees, crest through (x0, y0) m) to the current state over
 * the whole ghost-inclusive grid, with the local depth max(w - BA, 0).  The next step's
 * fluxes see the new state (the reference's next Pass2 still reads the H, U, V of the
 * Pass1 that ran before the wave was added). */
int cel_engine_add_solitary_wave(cel_engine *e, float H, float theta_deg, float x0, float y0);

/* Terrain changes.  cel_engine_set_bottom replaces the bottom planes (BY, BX, BA as from
 * cel_bottom_from_corners): fillTerrainTexture (engine.cpp:3313-3379) with keep_depth != 0
 * (w becomes (w - BA_old) + BA_new everywhere, so water depths stay), fillTerrainTextureLite
 * (engine.cpp:3381-3412) with keep_depth == 0 (the water stays).  Both clear the visual
 * record; the line-solve coefficients and the depth-dependent boundary tables follow the
 * new bottom; the time-stepping history is kept. */
int cel_engine_set_bottom(cel_engine *e, const float *bottom3, int keep_depth);
/* raiseLowerTerrain (engine.cpp:4685-4843): the Gaussian brush 2 exp(-z / (2 (1/3)^2)),
 * z = |p|^2 / (2 radius)^2 <= 1, times displacement, sampled at the cell corners, is added to
 * w and BA (the mean of the four corners), BX and BY (the means of the face's two corners)
 * over the brush's index box.  World coordinates: x about the domain centre
 * (index (nx-1)/W x + (nx+4)/2), y from the south (index (ny-1)/L y + 2.5), metres.
 * Negative displacement lowers. */
int cel_engine_raise_terrain(cel_engine *e, float x, float y, float radius, float displacement);
/* shiftTerrainSlider (engine.cpp:4845-4924), as resetTimestep applies a Tide/Surge/SLR change
 * (engine.cpp:1544-1553): cells with w - BA < 0.01 m become still water (BA below the still
 * water level) or move with the terrain; BY, BX, BA move by shift_up; the visual record is
 * cleared. */
int cel_engine_shift_terrain(cel_engine *e, float shift_up);
/* The left-mouse water brush (applyMouseShader, engine.cpp:4561-4683; left_mouse.hlsl:59-103)
 * at world (x, y) with the given radius: the cells whose centres the reference's 12-triangle
 * fan covers (D3D11 rasterisation: vertices snapped to 1/256 of a cell, top-left rule) take
 * dh = disp_A + disp_B r, r the interpolated radius (0 at the centre, 1 on the rim):
 * w' = max(B, w + dh), (hu, hv)' = (w' - B) (hu, hv) / max(0.5, w - B).  The reference's
 * actions at strength s (= "left_mouse_strength" x dt): add water (s, 0), remove water
 * (-s, 0), stir (-s, 1.5 s). */
int cel_engine_edit_water(cel_engine *e, float x, float y, float radius, float disp_A, float disp_B);

/* Runs nsteps and returns the elapsed device time in ms, measured with HIP events
 * on the engine's stream (inputs are already resident in HBM). */
int cel_engine_timed_steps(cel_engine *e, int nsteps, float *ms);

/* Per-kernel device time (ms) accumulated by the steps run while profiling is on
 * (profiling synchronises after every step); names: "explicit_step", "solve_lines",
 * "boundary_apply". */
int cel_engine_set_profiling(cel_engine *e, int on);
int cel_engine_kernel_time(cel_engine *e, const char *name, float *ms, int *launches);

/* Device bytes held by the engine. */
int cel_engine_device_bytes(cel_engine *e, size_t *bytes);

/* GetStats + resetTimestep's reductions (compute.hlsl:1319-1378, engine.cpp:1633-1696).
 * Per-cell values are the shader's FP32 ones; sums are reduced in FP64 on the device
 * over the cells the reference's 4x4 blocks cover (nx/4 by ny/4 blocks from cell
 * (2, 2)).  Derived values use density 1000 kg/m^3 and the cell area dx*dy, as the
 * reference does, but in FP64. */
typedef struct {
    double sum_h, sum_bh_h2, sum_hu, sum_hv, sum_hu2v2;      /* raw sums */
    double max_u2v2, max_h, max_cfl_rate, max_u2v2_h;        /* raw maxima */
    double mass, x_momentum, y_momentum;
    double kinetic_energy, potential_energy, total_energy;
    double max_speed, max_depth, max_froude;
    double cfl_number;           /* max_cfl_rate * dt */
    float timestep;              /* dt of the next step */
    int cells;                   /* cells reduced */
} cel_stats;

/* The settings resetTimestep reads for the adaptive dt (settings.hpp:333-337 and the
 * "target cfl" slider, settings.cpp:120 / terrain_heightfield.cpp:207). */
typedef struct {
    float ema_alpha;             /* exponential_moving_average_alpha (reference 0.5) */
    float min_timestep_ratio;    /* min_timestep_ratio (reference 0.05) */
    float target_cfl;            /* "target cfl" until the first adaptive call sets it */
} cel_adaptive;

/* GetStats of the current state (synchronises). */
int cel_engine_stats(cel_engine *e, cel_stats *out);
/* Checkpoint and resume.  cel_engine_checkpoint writes into buf (*bytes its capacity) all
 * that the next steps read: both state buffers (and a compensated state's low words), the
 * Adams-Bashforth ring of dU/dt and (F*, G*), the corrector's copies, the visual record,
 * the scheme counters, the dt history, the time and the time-series cursors; with buf NULL
 * it only sets *bytes to the size.  cel_engine_restore loads such a buffer into an engine
 * created with the same settings, boundaries and bottom (and the same visual setting); it
 * then continues bit-identically to the engine the checkpoint was taken from.  The
 * reference's hot start (engine.cpp:2943-2970, cel_engine_set_state here) keeps only
 * (w, hu, hv) and restarts the scheme with two Euler steps (engine.cpp:3267-3269). */
int cel_engine_checkpoint(cel_engine *e, void *buf, size_t *bytes);
int cel_engine_restore(cel_engine *e, const void *buf, size_t bytes);
/* GetStats without allocation: the nine FP64 sums and maxima of cel_engine_stats (the same
 * kernels, tiles and reduction orders, so the same bits) over caller-owned DEVICE planes
 * w, hu, hv, BA (row-major, ghost-inclusive, (nx+4) x (ny+4) cells, `pitch` floats per
 * row), written to the DEVICE array sums9 (9 doubles) on `stream` (a hipStream_t; NULL for
 * the default stream), asynchronously.  Its partial sums live in a file-scope device array
 * of 16384 tiles (1.2 MB per GPU, held until the process ends), so grids above
 * 4096 x 4096 cells are rejected (CEL_EINVAL); the call makes no hipMalloc, hipFree or
 * other allocating call, for callers that cannot allocate (HIP graph capture, serving
 * loops, harnesses that forbid it).  Calls that overlap on one GPU share that array and
 * race.  No counterpart in the reference. */
int cel_stats_noalloc(const cel_settings *s, const float *w, const float *hu, const float *hv, const float *BA,
                      int pitch, double *sums9, void *stream);
/* resetTimestep's post-processing (engine.cpp:1686-1696) of nine host-side sums from
 * cel_stats_noalloc, with dt the next step's: the cel_stats cel_engine_stats reports. */
int cel_stats_from_sums(const cel_settings *s, float dt, const double *sums9, cel_stats *out);
/* Numeric sentinel: CEL_ENUMERIC once a step has produced a non-finite surface elevation
 * or line right-hand side (the explicit kernel raises a device flag; a non-finite value
 * anywhere in the state reaches one of them within a step), else CEL_OK.  Waits for the
 * engine's work; cel_engine_set_state clears it.  The reference has no such check
 * (h >= 0 is kept only numerically, compute.hlsl:317-328, 637). */
int cel_engine_check_numeric(cel_engine *e);

/* resetTimestep (engine.cpp:1536-1756) without the GUI parts: GetStats, then the dt of
 * the next step -- adaptive (max(target/rate, timestep*min_ratio), moving average when
 * growing) while the scheme is CEL_PREDICTOR_ADAPTIVE, else the uniform dt (the
 * settings' timestep, or the last cel_engine_set_dt).  out may be NULL. */
int cel_engine_reset_timestep(cel_engine *e, const cel_adaptive *a, cel_stats *out);

/* Device bytes an engine with these settings and boundaries holds once created
 * (admission control against an HBM budget). */
int cel_estimate_device_bytes(const cel_settings *s, const cel_boundary bnd[4], size_t *bytes);

/* ---- Batched scenarios.  No counterpart in the reference (one engine per process).
 * A batch owns independent scenarios of any grid size and advances them together: one
 * launch per kernel per step over a work-list spanning every scenario (explicit tiles,
 * line-solve blocks per line-length class, boundary cells), optionally replayed from
 * hipGraphs of several steps.  Each scenario's results are those of a cel_engine with
 * the same inputs.  Scenarios cannot take corrector steps. */
typedef struct cel_batch cel_batch;
int cel_batch_create(cel_batch **out);
int cel_batch_destroy(cel_batch *b);
/* Adds a scenario (as cel_engine_create) and returns its id (> 0). */
int cel_batch_add(cel_batch *b, const cel_settings *s, const cel_boundary bnd[4], const float *bottom3,
                  const float *state3, int *id);
int cel_batch_remove(cel_batch *b, int id);
/* A paused scenario keeps its state and is skipped by cel_batch_step until resumed
 * (paused = 0): a scheduler runs a chosen subset of the resident scenarios. */
int cel_batch_set_paused(cel_batch *b, int id, int paused);
/* Advances every scenario that is not paused by nsteps (asynchronous). */
int cel_batch_step(cel_batch *b, int nsteps);
/* Steps per captured block (0: direct launches). */
int cel_batch_set_graph_steps(cel_batch *b, int steps);
/* Times the batched explicit launch with band heights 4, 8 and 16 over the current
 * scenarios and keeps the fastest (*band_rows); with tune_lines, autotunes each
 * scenario's line-solve shapes as cel_engine_autotune does.  Results are unchanged. */
int cel_batch_autotune(cel_batch *b, int repeats, int tune_lines, int *band_rows);
int cel_batch_sync(cel_batch *b);
/* Runs nsteps and returns the elapsed device time (HIP events on the batch stream). */
int cel_batch_timed_steps(cel_batch *b, int nsteps, float *ms);
int cel_batch_count(cel_batch *b, int *n);
int cel_batch_device_bytes(cel_batch *b, size_t *bytes);
/* Per-scenario access (synchronises the batch first). */
/* Moves an engine (FP32 or compensated, no corrector steps) into the batch as a new member,
 * *id its id: the engine can be created on another host thread while the batch runs.  On
 * success the handle is left empty (cel_engine_destroy still frees it); on failure it
 * stays the caller's.  No counterpart in the reference. */
int cel_batch_adopt(cel_batch *b, cel_engine *e, int *id);
/* Hot start of one member: its state from state3 (as cel_engine_set_state), clearing its
 * numeric sentinel.  No counterpart in the reference. */
int cel_batch_set_state(cel_batch *b, int id, const float *state3);
/* Numeric sentinel of the members: ids (up to max_ids of them) of the members whose steps
 * produced a non-finite surface elevation or line right-hand side since they were added
 * or last given a state, and their number.  Waits for the batch's work.  The other
 * members are unaffected: a caller evicts the failed ones (cel_batch_remove).  No
 * counterpart in the reference (which has no NaN check, engine.cpp:1536-1756). */
int cel_batch_check_numeric(cel_batch *b, int *ids, int max_ids, int *n_failed);
int cel_batch_get_state(cel_batch *b, int id, float *state3);
int cel_batch_get_state_f64(cel_batch *b, int id, double *state3);
int cel_batch_stats(cel_batch *b, int id, cel_stats *out);
int cel_batch_time(cel_batch *b, int id, double *t);

/* Work decomposition of an engine's kernels.  Results do not depend on it: every cell and
 * every elimination of the line solves is computed the same way whatever the choice.  No
 * counterpart in the reference. */
typedef struct {
    int band_rows;              /* rows per wave in the explicit stencil: 2, 4, 8, 12 or 16 */
    int x_chunk, x_waves;       /* x-line solve: unknowns per thread (1..16), waves per line (1..8) */
    int y_chunk, y_waves;       /* y-line solve */
    float explicit_us, x_lines_us, y_lines_us;   /* timings behind an autotuned choice (else 0) */
    /* Shortened exact division and square root in the explicit stencil (same bits as the
     * library operations on the arguments it is used for).  set: 0 forbids it, nonzero lets
     * the engine use it wherever the mesh and settings admit it (the default); get: 1 while
     * it is in use. */
    int fast_arith;
} cel_tuning;

/* Times every band height and every line-solve block shape that fits the engine's line
 * lengths on the current state (each candidate launched `repeats` times; the kernels
 * only write buffers the next step overwrites, so the state is unchanged), keeps the
 * fastest and reports it. */
int cel_engine_autotune(cel_engine *e, int repeats, cel_tuning *out);
int cel_engine_get_tuning(cel_engine *e, cel_tuning *out);
/* Applies a tuning (e.g. one cached from an earlier run on the same grid). */
int cel_engine_set_tuning(cel_engine *e, const cel_tuning *t);

/* Level-0 tridiagonal coefficients as the engine uses them (engine.cpp:2638-2694,
 * 2781-2826): three planes A, B, C for x and for y, computed on the device. */
int cel_engine_coefficients(cel_engine *e, float *abc_x3, float *abc_y3);

/* HBM streaming probe: kind 0 reads, 1 writes, 2 copies `bytes` (a multiple of 16 KiB, at
 * most 16 GiB) in one launch of 16 KiB per workgroup; `launches` timed launches after one
 * warm-up, *ms = average launch time.  Gives the device's achievable streaming rates and a
 * known-size launch for calibrating the HBM byte counters.  Allocates its own buffers.  No
 * counterpart in the reference. */
int cel_probe_bandwidth(int kind, size_t bytes, int launches, float *ms);

/* The explicit stencil's memory-access pattern without its arithmetic, on an nx x ny grid
 * (4..8192 per side) with band_rows-row bands (8, 12, 16): the same tiles, row loads and
 * stores, transposed hv staging and D_y stores.  layout 0: the stencil's (60-column strips);
 * 1: the same with hv and D_y row-major (no transposed accesses); 2: 64-column strips
 * aligned to 128-byte lines.  `launches` timed launches after one warm-up, *ms = average
 * launch time: how long the stencil's data movement alone takes.  Allocates its own planes.
 * No counterpart in the reference. */
int cel_probe_stencil_pattern(int nx, int ny, int band_rows, int layout, int launches, float *ms);

/* The line solves' structure without their coefficient arithmetic: `lines` x-lines and as
 * many y-lines (1..8192) of 2048 unknowns, one 256-thread block each, as the engine's line
 * solve lays them out: 16-byte loads of the right-hand side and bed, three chunk levels with
 * a neighbour exchange each, 17 cross-thread levels in LDS (a barrier each), 8-byte stores.
 * elim 0 combines the equations with additions only; 1 with cyclic reduction's eliminations
 * (diagonally dominant system).  `launches` timed launches after one warm-up, *ms = average
 * launch time.  Allocates its own planes.  No counterpart in the reference. */
int cel_probe_solve_pattern(int lines, int elim, int launches, float *ms);

#ifdef __cplusplus
}
#endif
#endif
