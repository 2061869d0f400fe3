// Device-resident drift and kick (predict.c:10-41, timestep.c:51-59 and
// :86-92).  Positions and velocities stay FP64 on the device.  Compiled
// with -ffp-contract=off: x += v*d and v += a*k round exactly as the C code.
//
// The *_table variants read their factor from a device array indexed by a
// device step counter, so one captured hipGraph serves every step.
#include "internal.h"

namespace cuco {

namespace {

// ---- drift and kick tables (driftfac.c:14-42) on the device
// Entry i is the integral from a_begin to a_i = exp(ln a_begin + (ln a_max -
// ln a_begin)/1000 (i+1)) of da/(H a^3) (drift) and da/(H a^2) (kick),
// H = hubble sqrt(om/a^3 + (1-om-ol)/a^2 + ol).  Thread i integrates
// [a_{i-1}, a_i] (a_{-1} = a_begin) with the 16-point Gauss-Legendre rule
// (exact for polynomials of degree 31; an interval spans a factor
// (a_max/a_begin)^(1/1000) in a, so the smooth integrand is resolved to
// rounding), then one thread adds the pieces in table order.
constexpr double kGl16[8][2] = {  // Legendre roots x > 0 and weights
    {0.09501250983763745, 0.18945061045506859}, {0.2816035507792589, 0.1826034150449236},
    {0.45801677765722737, 0.16915651939500262}, {0.6178762444026438, 0.14959598881657676},
    {0.755404408355003, 0.12462897125553403},   {0.8656312023878318, 0.09515851168249259},
    {0.9445750230732326, 0.062253523938647706}, {0.9894009349916499, 0.027152459411754037}};

__device__ inline void drift_kick_integrands(double a, double om, double ol, double hubble,
                                             double &fd, double &fk)
{
    // driftfac.c:48-52 and :59-63
    double h = om / (a * a * a) + (1 - om - ol) / (a * a) + ol;
    h = hubble * sqrt(h);
    fd = 1 / (h * a * a * a);
    fk = 1 / (h * a * a);
}

__global__ __launch_bounds__(1024) void k_drift_tables(double a_begin, double lb, double lm,
                                                       double om, double ol, double hubble,
                                                       double *__restrict__ drift,
                                                       double *__restrict__ kick)
{
    __shared__ double sd[kDriftTable], sk[kDriftTable];
    for (int i = threadIdx.x; i < kDriftTable; i += blockDim.x) {
        const double lo = i == 0 ? a_begin : exp(lb + ((lm - lb) / kDriftTable) * i);
        const double hi = exp(lb + ((lm - lb) / kDriftTable) * (i + 1));
        const double c = 0.5 * (lo + hi), r = 0.5 * (hi - lo);
        double d = 0, k = 0;
        for (int q = 0; q < 8; ++q) {
            double fd, fk, gd, gk;
            drift_kick_integrands(c - r * kGl16[q][0], om, ol, hubble, fd, fk);
            drift_kick_integrands(c + r * kGl16[q][0], om, ol, hubble, gd, gk);
            d += kGl16[q][1] * (fd + gd);
            k += kGl16[q][1] * (fk + gk);
        }
        sd[i] = d * r;
        sk[i] = k * r;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double d = 0, k = 0;
        for (int i = 0; i < kDriftTable; ++i) {
            d += sd[i];
            k += sk[i];
            drift[i] = d;
            kick[i] = k;
        }
    }
}

// predict.c:31-36.  A particle moves far less than a box length per step;
// the bounded loops cannot spin on a non-finite value.
__device__ inline double wrap_coord(double x, double L)
{
    for (int k = 0; k < 16 && x < 0; ++k) x += L;
    for (int k = 0; k < 16 && x >= L; ++k) x -= L;
    return x;
}

__global__ void k_drift(double4 *__restrict__ pos, const double4 *__restrict__ vel,
                        BoxView v, const double *__restrict__ factors,
                        const uint32_t *__restrict__ step, double drift)
{
    int bi;
    uint32_t lb;
    if (!box_block(v, bi, lb)) return;
    const BoxDev &bx = v.box[bi];
    const uint32_t il = lb * blockDim.x + threadIdx.x;
    if (il >= bx.n) return;
    const uint32_t i = bx.p0 + il;
    const double L = bx.g.L;
    const double d = factors ? factors[3 * step[0] + 1] : drift;
    double4 p = pos[i];
    const double4 q = vel[i];
    p.x += q.x * d;
    p.y += q.y * d;
    p.z += q.z * d;
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

// ------------------------------------------------------------ integer timeline
// get_gravkick_factor (driftfac.c:96-126): the kick table interpolated
// linearly in ln a at two integer times; the u < 1 branch as there.
__device__ inline double table_at(const double *__restrict__ table, double u)
{
    int i = (int)u;
    if (i >= kDriftTable) i = kDriftTable - 1;
    if (i < 1) return u * table[0];
    return table[i - 1] + (table[i] - table[i - 1]) * (u - i);
}

__device__ inline double kick_factor(const TimelineDev &t, const double *__restrict__ table,
                                     int time0, int time1)
{
    const double a1 = t.log_begin + time0 * t.interval;
    const double a2 = t.log_begin + time1 * t.interval;
    const double u1 = (a1 - t.log_begin) / (t.log_max - t.log_begin) * kDriftTable;
    const double u2 = (a2 - t.log_begin) / (t.log_max - t.log_begin) * kDriftTable;
    return table_at(table, u2) - table_at(table, u1);
}

// advance_and_find_timesteps, particle part (timestep.c:30-70), with
// get_timestep (timestep.c:96-138): for every particle whose step ends now,
// a new power-of-two step from the acceleration criterion, then the kick
// with its short-range acceleration from the middle of the old step to the
// middle of the new one.  A step that falls off the integer timeline sets
// bit 0 of flags and leaves the particle unchanged.
template <typename A>
__global__ void k_advance(double4 *__restrict__ vel, const A *__restrict__ app,
                          const A *__restrict__ apm, int32_t *__restrict__ ti_beg,
                          int32_t *__restrict__ ti_end, uint32_t n, TimelineDev t,
                          const double *__restrict__ kick_table,
                          int32_t *__restrict__ flags)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t end = ti_end[i];
    if (end != t.ti_current) return;
    const int32_t beg = ti_beg[i];
    const A g = app[i];
    const A pm = apm[i];
    // get_timestep (timestep.c:96-138)
    double ax = t.fac1 * g.x, ay = t.fac1 * g.y, az = t.fac1 * g.z;
    ax += t.fac1 * pm.x;
    ay += t.fac1 * pm.y;
    az += t.fac1 * pm.z;
    double ac = sqrt(ax * ax + ay * ay + az * az);
    if (ac == 0) ac = 1.0e-30;
    double dt = sqrt(2 * t.err_tol * t.atime * t.soft_table / ac);
    dt *= t.hubble_a;
    if (dt >= t.dt_displacement) dt = t.dt_displacement;
    if (dt < t.min_dt) dt = t.min_dt;
    const double steps = dt / t.interval;
    if (!(steps > 0 && steps < (double)kTimebase)) {
        atomicOr(flags, 1);
        return;
    }
    int ti_step = (int)steps;
    if (!(ti_step > 0 && ti_step < kTimebase)) {
        atomicOr(flags, 1);
        return;
    }
    // timestep.c:32-49
    int ti_min = kTimebase;
    while (ti_min > ti_step) ti_min >>= 1;
    ti_step = ti_min;
    if (ti_step > (end - beg)) {
        if (((kTimebase - end) % ti_step) > 0) ti_step = end - beg;
    }
    if (t.ti_current == kTimebase) ti_step = 0;
    if ((kTimebase - t.ti_current) < ti_step) ti_step = kTimebase - t.ti_current;
    const int tstart = (beg + end) / 2;
    const int tend = end + ti_step / 2;
    const double k = kick_factor(t, kick_table, tstart, tend);
    ti_beg[i] = end;
    ti_end[i] = end + ti_step;
    double4 v = vel[i];
    v.x += (double)g.x * k;
    v.y += (double)g.y * k;
    v.z += (double)g.z * k;
    vel[i] = v;
}

// Reductions over Ti_endstep: a fixed grid of 256-thread workgroups strides
// over int4 loads (ti_end is a whole device allocation, so 16-byte
// aligned), reduces in registers, across the wave and across the four waves
// in LDS, and makes one atomic per workgroup.  One atomic per wave on a
// single address serialises in L2 (3 ms at 16.7M particles).
constexpr int kRedThreads = 256;
constexpr uint32_t kRedMaxBlocks = 1024;

inline uint32_t red_blocks(uint32_t n)
{
    const uint32_t b = ((n >> 2) + kRedThreads - 1) / kRedThreads;
    return b < 1 ? 1 : b > kRedMaxBlocks ? kRedMaxBlocks : b;
}

template <bool COUNT>
__global__ __launch_bounds__(kRedThreads) void k_ti_reduce(
    const int32_t *__restrict__ ti_end, uint32_t n, int32_t ti,
    int32_t *__restrict__ out)
{
    __shared__ int32_t sh[kRedThreads / 64];
    auto op = [](int32_t a, int32_t b) { return COUNT ? a + b : min(a, b); };
    auto elem = [ti](int32_t x) { return COUNT ? (int32_t)(x == ti) : x; };
    int32_t v = COUNT ? 0 : 0x7fffffff;
    const uint32_t n4 = n >> 2, stride = gridDim.x * kRedThreads;
    const int4 *t4 = reinterpret_cast<const int4 *>(ti_end);
    for (uint32_t i = blockIdx.x * kRedThreads + threadIdx.x; i < n4; i += stride) {
        const int4 q = t4[i];
        v = op(v, op(op(elem(q.x), elem(q.y)), op(elem(q.z), elem(q.w))));
    }
    for (uint32_t i = 4 * n4 + blockIdx.x * kRedThreads + threadIdx.x; i < n; i += stride)
        v = op(v, elem(ti_end[i]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = op(v, __shfl_xor(v, o, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        v = op(op(sh[0], sh[1]), op(sh[2], sh[3]));
        if (COUNT) atomicAdd(out, v); else atomicMin(out, v);
    }
}

// Pair-work estimate of an active set: every particle weighs the number of
// particles in its aligned block of 2^shift Morton keys (shift 12: 16^3
// cells, about the size of the short-range sphere, so the weight tracks the
// candidates the particle meets); out[0] += weights of all particles,
// out[1] += weights of the active ones (sorted order: skeys, perm; off is
// the CSR cell table).
__global__ __launch_bounds__(kRedThreads) void k_active_work(
    const uint32_t *__restrict__ skeys, const uint32_t *__restrict__ perm,
    const uint32_t *__restrict__ off, uint32_t n, int shift,
    const int32_t *__restrict__ ti_end, int32_t ti,
    unsigned long long *__restrict__ out)
{
    __shared__ unsigned long long sh[2][kRedThreads / 64];
    unsigned long long wa = 0, wx = 0;
    const uint32_t stride = gridDim.x * kRedThreads;
    for (uint32_t j = blockIdx.x * kRedThreads + threadIdx.x; j < n; j += stride) {
        const uint32_t b0 = (skeys[j] >> shift) << shift;
        const uint32_t w = off[b0 + (1u << shift)] - off[b0];
        wa += w;
        if (ti_end[perm[j]] == ti) wx += w;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        wa += __shfl_xor(wa, o, 64);
        wx += __shfl_xor(wx, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        sh[0][threadIdx.x >> 6] = wa;
        sh[1][threadIdx.x >> 6] = wx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd(&out[0], sh[0][0] + sh[0][1] + sh[0][2] + sh[0][3]);
        atomicAdd(&out[1], sh[1][0] + sh[1][1] + sh[1][2] + sh[1][3]);
    }
}

}  // namespace

int launch_active_work(const uint32_t *skeys, const uint32_t *perm, const uint32_t *off,
                       uint32_t n, int shift, const int32_t *ti_end, int32_t ti_current,
                       unsigned long long *out, hipStream_t s)
{
    int rc = launch_fill_u32(reinterpret_cast<uint32_t *>(out), 0u, 4, s);
    if (rc) return rc;
    if (n == 0) return CUCO_OK;
    const uint32_t blocks = std::min<uint32_t>(kRedMaxBlocks, (n + kRedThreads - 1) / kRedThreads);
    k_active_work<<<blocks, kRedThreads, 0, s>>>(skeys, perm, off, n, shift, ti_end,
                                                  ti_current, out);
    CUCO_LAUNCH_CHECK();
    return CUCO_OK;
}

int launch_drift_tables(double a_begin, double a_max, double omega0, double omega_lambda,
                        double hubble, double *drift, double *kick, hipStream_t s)
{
    k_drift_tables<<<1, 1024, 0, s>>>(a_begin, log(a_begin), log(a_max), omega0, omega_lambda,
                                      hubble, drift, kick);
    CUCO_LAUNCH_CHECK();
    return CUCO_OK;
}

int launch_drift(double4 *pos, const double4 *vel, const BoxSet &bs,
                 double drift, hipStream_t s)
{
    if (bs.total[kMapPart] == 0) return CUCO_OK;
    k_drift<<<xcd_grid(bs.total[kMapPart]), 256, 0, s>>>(
        pos, vel, bs.view(0, kMapPart), nullptr, nullptr, drift);
    CUCO_LAUNCH_CHECK();
    return CUCO_OK;
}

int launch_drift_table(double4 *pos, const double4 *vel, const BoxSet &bs,
                       const double *factors, const uint32_t *step,
                       hipStream_t s)
{
    if (bs.total[kMapPart] == 0) return CUCO_OK;
    k_drift<<<xcd_grid(bs.total[kMapPart]), 256, 0, s>>>(
        pos, vel, bs.view(0, kMapPart), factors, step, 0.0);
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

int launch_advance(double4 *vel, const void *app, const void *apm, int fp32,
                   int32_t *ti_beg, int32_t *ti_end, uint32_t n,
                   const TimelineDev &t, const double *kick_table,
                   int32_t *flags, hipStream_t s)
{
    if (n == 0) return CUCO_OK;
    if (fp32)
        k_advance<float4><<<grid_for(n, 256), 256, 0, s>>>(
            vel, static_cast<const float4 *>(app), static_cast<const float4 *>(apm),
            ti_beg, ti_end, n, t, kick_table, flags);
    else
        k_advance<double4><<<grid_for(n, 256), 256, 0, s>>>(
            vel, static_cast<const double4 *>(app), static_cast<const double4 *>(apm),
            ti_beg, ti_end, n, t, kick_table, flags);
    CUCO_LAUNCH_CHECK();
    return CUCO_OK;
}

int launch_ti_min(const int32_t *ti_end, uint32_t n, int32_t *out, hipStream_t s)
{
    int rc = launch_fill_u32(reinterpret_cast<uint32_t *>(out), 0x7fffffffu, 1, s);
    if (rc) return rc;
    if (n == 0) return CUCO_OK;
    k_ti_reduce<false><<<red_blocks(n), kRedThreads, 0, s>>>(ti_end, n, 0, out);
    CUCO_LAUNCH_CHECK();
    return CUCO_OK;
}

int launch_count_active(const int32_t *ti_end, uint32_t n, int32_t ti_current,
                        int32_t *out, hipStream_t s)
{
    int rc = launch_fill_u32(reinterpret_cast<uint32_t *>(out), 0u, 1, s);
    if (rc) return rc;
    if (n == 0) return CUCO_OK;
    k_ti_reduce<true><<<red_blocks(n), kRedThreads, 0, s>>>(ti_end, n, ti_current, out);
    CUCO_LAUNCH_CHECK();
    return CUCO_OK;
}

}  // namespace cuco
