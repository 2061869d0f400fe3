// Device-resident drift and kick (predict.c:10-41, timestep.c:223-226 and
// :253-255).  Positions and velocities stay FP64 on the device.  Compiled
// with -ffp-contract=off: x += v*d and v += a*k round exactly as the C code.
//
// The *_table variants read their factor from a device array indexed by a
// device step counter, so one captured hipGraph serves every step.
#include "internal.h"

namespace cuco {

namespace {

// predict.c:31-36.  A particle moves far less than a box length per step;
// the bounded loops cannot spin on a non-finite value.
__device__ inline double wrap_coord(double x, double L)
{
    for (int k = 0; k < 16 && x < 0; ++k) x += L;
    for (int k = 0; k < 16 && x >= L; ++k) x -= L;
    return x;
}

__global__ void k_drift(double4 *__restrict__ pos, const double4 *__restrict__ vel,
                        uint32_t n, double L, const double *__restrict__ factors,
                        const uint32_t *__restrict__ step, double drift)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double d = factors ? factors[3 * step[0] + 1] : drift;
    double4 p = pos[i];
    const double4 v = vel[i];
    p.x += v.x * d;
    p.y += v.y * d;
    p.z += v.z * d;
    p.x = wrap_coord(p.x, L);
    p.y = wrap_coord(p.y, L);
    p.z = wrap_coord(p.z, L);
    pos[i] = p;
}

template <typename A>
__global__ void k_kick(double4 *__restrict__ vel, const A *__restrict__ app,
                       const A *__restrict__ apm, uint32_t n,
                       const double *__restrict__ factors,
                       const uint32_t *__restrict__ step, int which, double kpp,
                       double kpm)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (factors) {
        kpp = factors[3 * step[0] + which];
        kpm = kpp;
    }
    double4 v = vel[i];
    if (app) {
        const A a = app[i];
        v.x += (double)a.x * kpp;
        v.y += (double)a.y * kpp;
        v.z += (double)a.z * kpp;
    }
    if (apm) {
        const A a = apm[i];
        v.x += (double)a.x * kpm;
        v.y += (double)a.y * kpm;
        v.z += (double)a.z * kpm;
    }
    vel[i] = v;
}

__global__ void k_step_advance(uint32_t *step) { step[0] += 1u; }

}  // namespace

int launch_drift(double4 *pos, const double4 *vel, uint32_t n, double L,
                 double drift, hipStream_t s)
{
    if (n == 0) return CUCO_OK;
    k_drift<<<grid_for(n, 256), 256, 0, s>>>(pos, vel, n, L, nullptr, nullptr, drift);
    CUCO_LAUNCH_CHECK();
    return CUCO_OK;
}

int launch_drift_table(double4 *pos, const double4 *vel, uint32_t n, double L,
                       const double *factors, const uint32_t *step,
                       hipStream_t s)
{
    if (n == 0) return CUCO_OK;
    k_drift<<<grid_for(n, 256), 256, 0, s>>>(pos, vel, n, L, factors, step, 0.0);
    CUCO_LAUNCH_CHECK();
    return CUCO_OK;
}

int launch_kick64(double4 *vel, const double4 *app, const double4 *apm,
                  uint32_t n, double kpp, double kpm, hipStream_t s)
{
    if (n == 0) return CUCO_OK;
    k_kick<double4><<<grid_for(n, 256), 256, 0, s>>>(vel, app, apm, n, nullptr,
                                                     nullptr, 0, kpp, kpm);
    CUCO_LAUNCH_CHECK();
    return CUCO_OK;
}

int launch_kick32(double4 *vel, const float4 *app, const float4 *apm,
                  uint32_t n, double kpp, double kpm, hipStream_t s)
{
    if (n == 0) return CUCO_OK;
    k_kick<float4><<<grid_for(n, 256), 256, 0, s>>>(vel, app, apm, n, nullptr,
                                                    nullptr, 0, kpp, kpm);
    CUCO_LAUNCH_CHECK();
    return CUCO_OK;
}

int launch_kick_table(double4 *vel, const void *app, const void *apm, int fp32,
                      uint32_t n, const double *factors, const uint32_t *step,
                      int which, hipStream_t s)
{
    if (n == 0) return CUCO_OK;
    if (fp32)
        k_kick<float4><<<grid_for(n, 256), 256, 0, s>>>(
            vel, static_cast<const float4 *>(app), static_cast<const float4 *>(apm),
            n, factors, step, which, 0.0, 0.0);
    else
        k_kick<double4><<<grid_for(n, 256), 256, 0, s>>>(
            vel, static_cast<const double4 *>(app), static_cast<const double4 *>(apm),
            n, factors, step, which, 0.0, 0.0);
    CUCO_LAUNCH_CHECK();
    return CUCO_OK;
}

int launch_step_advance(uint32_t *step, hipStream_t s)
{
    k_step_advance<<<1, 1, 0, s>>>(step);
    CUCO_LAUNCH_CHECK();
    return CUCO_OK;
}

}  // namespace cuco
