/*
 * oracle.c -- CPU oracle for Celeris's Boussinesq step (TEST INFRASTRUCTURE ONLY;
 * loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
 * never by the product).  Instantiates oracle_impl.inc in FP32 and FP64.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "celeris_oracle.h"
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- FP32 instance ---- */
typedef float real;
#define FN(x) orc_f32_##x
#define SIM struct orc_sim_f32
#define SQRT sqrtf
#define SIN sinf
#define COS cosf
#define TANH tanhf
#define POW powf
#define FABS fabsf
#define ATAN atanf
#define COSH coshf
#define EXP expf
#include "oracle_impl.inc"
#undef FN
#undef SIM
#undef SQRT
#undef SIN
#undef COS
#undef TANH
#undef POW
#undef FABS
#undef ATAN
#undef COSH
#undef EXP
#undef MAXI
#undef MINI
#undef NUM_TEX
#undef MAXLEV
#undef CURRENT_STATE
#define real double
/* ---- FP64 instance ---- */
#define FN(x) orc_f64_##x
#define SIM struct orc_sim_f64
#define SQRT sqrt
#define SIN sin
#define COS cos
#define TANH tanh
#define POW pow
#define FABS fabs
#define ATAN atan
#define COSH cosh
#define EXP exp
#include "oracle_impl.inc"
#undef real

/* OpenMP threads of the oracle's row loops (timing only; results do not depend on it). */
void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
int orc_get_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* UpdateTerrainHeightfield, terrain_heightfield.cpp:244-298: corner heights at
 * ((i-2.5), (j-2.5)) for i<nx+5, j<ny+5 reduced to the per-cell face/centre values
 * in the same accumulation order as the reference loop. */
void orc_bottom_from_corners(int nx, int ny, const float *z, float *bottom3)
{
    const int pitch = nx + 4, zp = nx + 5;
    const size_t N = (size_t)pitch * (ny + 4);
    float *BY = bottom3, *BX = bottom3 + N, *BA = bottom3 + 2 * N;
    for (int j = 0; j < ny + 5; ++j) {
        for (int i = 0; i < nx + 5; ++i) {
            const float zz = z[j * zp + i];
            if (i < nx + 4) {
                if (j < ny + 4) BA[j * pitch + i] = 0.25f * zz;
                if (j > 0) {
                    BA[(j - 1) * pitch + i] += 0.25f * zz;
                    BY[(j - 1) * pitch + i] = 0.5f * zz;
                }
            }
            if (i > 0) {
                if (j < ny + 4) {
                    BA[j * pitch + i - 1] += 0.25f * zz;
                    BX[j * pitch + i - 1] = 0.5f * zz;
                }
                if (j > 0) {
                    BA[(j - 1) * pitch + i - 1] += 0.25f * zz;
                    BX[(j - 1) * pitch + i - 1] += 0.5f * zz;
                    BY[(j - 1) * pitch + i - 1] += 0.5f * zz;
                }
            }
        }
    }
}

/* Known-answer-test entry points for the compute.hlsl helpers. */
float orc_f32_minmod(float a, float b, float c) { return orc_f32_MinMod(a, b, c); }
void orc_f32_reconstruct(float two_theta, float west, float here, float east, float *ow, float *oe)
{
    orc_f32_Reconstruct(two_theta, west, here, east, ow, oe);
}
void orc_f32_correct_w(float B_west, float B_east, float w_bar, float *w_west, float *w_east)
{
    orc_f32_CorrectW(B_west, B_east, w_bar, w_west, w_east);
}
float orc_f32_divide_by_h(float h, float epsilon) { return orc_f32_DivideByH(h, epsilon); }
float orc_f32_numerical_flux(float ap, float am, float Fp, float Fm, float dU) { return orc_f32_NumericalFlux(ap, am, Fp, Fm, dU); }
float orc_f32_sponge(float L, float x) { return orc_f32_SpongeFunction(4.0f * atanf(1.0f), L, x); }
float orc_f32_wavenumber(float omega, float d, float g) { return orc_f32_Wavenumber(omega, d, g); }
void orc_f32_ab_coef(float dt, float dt1, float dt2, float *coef3) { orc_f32_ABCoef(dt, dt1, dt2, coef3); }
