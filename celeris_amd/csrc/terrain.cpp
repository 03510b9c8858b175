// terrain.cpp -- host preprocessing on the path: bottom planes from corner heights and
// the initial condition, in FP32 with the reference's accumulation order.
#include <algorithm>
#include <cmath>

#include "engine.hpp"

namespace cel {

// UpdateTerrainHeightfield, terrain_heightfield.cpp:260-298.  Corner (i, j) lies at
// ((i - 2.5) dx - W/2, (j - 2.5) dy); each cell's BA is the mean of its four corners,
// BX the mean of its two east corners, BY the mean of its two north corners, summed in
// the order the reference's corner loop visits them.
void bottom_from_corners(int nx, int ny, const float *z, float *bottom3)
{
    const int pitch = nx + 4, zp = nx + 5;
    const size_t N = (size_t)pitch * (ny + 4);
    float *BY = bottom3, *BX = bottom3 + N, *BA = bottom3 + 2 * N;
    for (int j = 0; j < ny + 4; ++j)
        for (int i = 0; i < nx + 4; ++i) {
            const float z00 = z[(size_t)j * zp + i], z10 = z[(size_t)j * zp + i + 1];
            const float z01 = z[(size_t)(j + 1) * zp + i], z11 = z[(size_t)(j + 1) * zp + i + 1];
            const size_t c = (size_t)j * pitch + i;
            BA[c] = ((0.25f * z00 + 0.25f * z10) + 0.25f * z01) + 0.25f * z11;
            BX[c] = 0.5f * z10 + 0.5f * z11;
            BY[c] = 0.5f * z01 + 0.5f * z11;
        }
}

namespace {
// solitaryWave, engine.cpp:103-112; pow(x, n) with an integer n follows MSVC's
// repeated squaring: x^3 = x * (x * x), c^2 = c * c.
void solitary_wave(float g, float depth, float H, float theta, float x0, float y0, float x, float y, float &eta,
                   float &hu, float &hv)
{
    const float K = std::sqrt(0.75f * std::fabs(H) / (depth * (depth * depth)));
    const float C = std::sqrt(g * (H + depth));
    const float ch = std::cosh(K * ((x - x0) * std::cos(theta) + (y - y0) * std::sin(theta)));
    const float e = H * 1.0f / (ch * ch);
    eta = e;
    hu = (C * std::cos(theta) * e);
    hv = (C * std::sin(theta) * e);
}
}  // namespace

// createSimTextures initial condition, engine.cpp:2945-3016: w = max(SWL, B), hu = hv = 0
// (or the hot-start values base3 when given), plus every solitary wave.
void initial_state(const cel_settings &s, const float *bottom3, const cel_soliton *sol, int nsol, float *state3,
                   const float *base3)
{
    const int nx = s.nx, ny = s.ny, X = nx + 4, Y = ny + 4;
    const size_t N = (size_t)X * Y;
    const float *BA = bottom3 + 2 * N;
    const float W = s.width, L = s.length;
    const float delta_x = W / (nx - 1), delta_y = L / (ny - 1);
    const float PI = 4.0f * std::atan(1.0f);
    for (int j = 0; j < Y; ++j)
        for (int i = 0; i < X; ++i) {
            const int ii = std::max(2, std::min(nx + 1, i));
            const int jj = std::max(2, std::min(ny + 1, j));
            const float B = BA[(size_t)X * jj + ii];
            const float waterDepth = std::max(0.0f, s.still_water_elevation - B);
            const size_t c0 = (size_t)j * X + i;
            float w = base3 ? base3[c0] : std::max(s.still_water_elevation, B);
            float hu = base3 ? base3[N + c0] : 0.f, hv = base3 ? base3[2 * N + c0] : 0.f;
            for (int k = 0; k < nsol; ++k) {
                const cel_soliton &t = sol[k];
                if (t.param_h == 0) continue;
                const float vert = std::fabs(std::sin(t.theta) * (delta_x * i - t.xc)) +
                                   std::fabs(std::cos(t.theta) * (delta_y * j - t.yc));
                float e, u, v;
                if (vert > t.length / 2.0f) {
                    float th = 0;
                    const float smoothing = PI / std::sqrt(0.75f * t.param_h / (waterDepth * (waterDepth * waterDepth)));
                    if (vert - t.length / 2.0f < smoothing)
                        th = t.param_h * (0.5f * std::cos(PI / smoothing * (vert - t.length / 2.0f)) + 0.5f);
                    else
                        th = 0;
                    solitary_wave(s.gravity, waterDepth, th, t.theta, t.xc, t.yc, delta_x * i, delta_y * j, e, u, v);
                } else {
                    solitary_wave(s.gravity, waterDepth, t.param_h, t.theta, t.xc, t.yc, delta_x * i, delta_y * j, e, u, v);
                }
                w += e; hu += u; hv += v;
            }
            const size_t c = (size_t)j * X + i;
            state3[c] = w; state3[N + c] = hu; state3[2 * N + c] = hv;
        }
}

}  // namespace cel
