// FP64 short-range pair force -- the plain path.
//
// One thread per sorted target particle, visiting neighbour cells in the
// reference's order: ix, iy, iz ascending over
//   [(int)(fac*(x-rcut)+M)-M, (int)(fac*(x+rcut))]  (gravlist.cpp:233-243),
// and inside a cell the particles in stable-sorted order
// (gravlist_kernel.cu:195-213).  With x-major keys the cells of one (ix,iy)
// column are contiguous in the sorted array, so each column is one or two
// contiguous j ranges (two when the z range wraps), read straight from the
// CSR offsets.  Columns and z cells that cannot hold a particle within rcut
// are skipped; that only removes pairs whose r^2 exceeds rcut^2, so the
// summation order of the surviving terms equals the reference loop's.
//
// The pair law is gravlist.cpp:279-327 (with the source mass), evaluated in
// the same operation order as the C source; this file is compiled with
// -ffp-contract=off, so no multiply-add is fused and the FP64 result is
// bit-identical to the CPU restatement.
#include "internal.h"

namespace cuco {

namespace {

constexpr int kThreads = 256;

// sqrt(x) and n / d for the ranges pair terms take, as the operation
// sequences of the compiler's correctly rounded f64 sqrt and divide on
// gfx950 without their extreme-exponent scaling and special-value selects
// (inactive there: the same operations on the same values, so the same
// bits).  sqrt_pair: x = 0 or 2^-767 <= x < 2^1023.  div_pair: normal n and
// d whose quotient and reciprocal stay normal (masses over cubed
// distances, the spline's 0.06667 / u^3 with u >= 0.5).
__device__ inline double sqrt_pair(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    double g = x * y;
    double h = y * 0.5;
    const double e = fma(-h, g, 0.5);
    g = fma(g, e, g);
    h = fma(h, e, h);
    double d = fma(-g, g, x);
    g = fma(d, h, g);
    d = fma(-g, g, x);
    g = fma(d, h, g);
    return x == 0.0 ? x : g;
}

__device__ inline double div_pair(double n, double d)
{
    double y = __builtin_amdgcn_rcp(d);
    double e = fma(-d, y, 1.0);
    y = fma(y, e, y);
    e = fma(-d, y, 1.0);
    y = fma(y, e, y);
    const double q = n * y;
    const double r = fma(-d, q, n);
    return fma(r, y, q);
}

// Minimum image.  WRAP: the reference's per-pair tests.  Otherwise the
// caller passes the image shift of j's cell (sx, sy, sz: -L for a cell
// index below 0, +L above M-1, else 0): when rcut plus one cell stays below
// half the box (cell_images), a pair in an image cell is exactly one the
// tests would wrap and no other pair is, and dx + (-L) is the same IEEE
// operation as dx -= L (dx + 0.0 leaves dx unchanged), so the result is
// bit-identical.
template <bool WRAP>
__device__ inline void pair64(const PairParams &pp, const double *__restrict__ tab,
                              double px, double py, double pz, double4 q,
                              double sx, double sy, double sz,
                              double &ax, double &ay, double &az,
                              uint32_t &in_count)
{
    double dx = q.x - px;
    double dy = q.y - py;
    double dz = q.z - pz;
    if (WRAP) {
        if (dx > pp.boxhalf) dx -= pp.L;
        if (dx < -pp.boxhalf) dx += pp.L;
        if (dy > pp.boxhalf) dy -= pp.L;
        if (dy < -pp.boxhalf) dy += pp.L;
        if (dz > pp.boxhalf) dz -= pp.L;
        if (dz < -pp.boxhalf) dz += pp.L;
    } else {
        dx += sx;
        dy += sy;
        dz += sz;
    }
    double r2 = dx * dx + dy * dy + dz * dz;
    if (r2 > pp.rcut2) return;
    ++in_count;
    double r = sqrt_pair(r2);
    double fac;
    if (r >= pp.h) {
        fac = div_pair(q.w, r2 * r);
    } else {
        double u = r * pp.h_inv;
        if (u < 0.5)
            fac = q.w * pp.h3_inv * (10.66667 + u * u * (32.0 * u - 38.4));
        else
            fac = q.w * pp.h3_inv *
                  (21.33333 - 48.0 * u + 38.4 * u * u - 10.66667 * u * u * u -
                   div_pair(0.06667, u * u * u));  // u >= 0.5
    }
    int tabindex = (int)(pp.asmthfac * r);
    if (tabindex < kNTab) {
        fac *= tab[tabindex];
        ax += dx * fac;
        ay += dy * fac;
        az += dz * fac;
    }
}

// Cell pruning only has to keep every cell that can hold a pair inside
// rcut: cells it keeps beyond rcut contribute nothing (their pairs fail the
// r^2 test) and do not change the order of the others.  So it runs in FP32
// with margins far above FP32 rounding of coordinates below 2^12 cells.
//
// Squared distance in cell units from u to the cell interval [c, c+1), made
// smaller by the margin.
__device__ inline float gap_cells(float u, int c)
{
    float g = fmaxf(0.0f, fmaxf((float)c - u, u - (float)(c + 1)));
    g = fmaxf(0.0f, g - 1e-2f);
    return g * g;
}

__device__ inline int wave_min_i(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return __builtin_amdgcn_readfirstlane(v);
}

__device__ inline int wave_max_i(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return __builtin_amdgcn_readfirstlane(v);
}

// z cells of column (ix, iy) that can hold a pair inside rcut for the target
// at cell coordinate uz: [za, zb] inside the reference range [zl, zr], or
// za > zb.  gxy: squared gap to the column in cells (both widened by the
// margin).
__device__ inline void z_window(float uz, float gxy, float rc2, int zl, int zr, int &za, int &zb)
{
    const float zext = sqrtf(fmaxf(rc2 - gxy, 0.0f)) + 1e-2f;
    za = max(zl, (int)floorf(uz - zext));
    zb = min(zr, (int)floorf(uz + zext));
}

// Image shift of a neighbour cell index c (unwrapped) on a mesh of M.
__device__ inline double image_shift(int c, int M, double L)
{
    return c < 0 ? -L : (c >= M ? L : 0.0);
}

// Eight waves per SIMD (at most 64 VGPRs, a few spilled): each lane's pair
// loop is a long dependent FP64 chain (correctly rounded sqrt and divide),
// so more resident waves hide more of it (441 ms at the register
// allocation's natural 5 waves, 400 ms at 8, halo256).
template <bool COUNT, bool WRAP>
__global__ __launch_bounds__(kThreads, 8) void k_pp64(
    const double4 *__restrict__ spos, const uint32_t *__restrict__ perm,
    const uint32_t *__restrict__ off, uint32_t i0, uint32_t n, PairParams pp,
    const float *__restrict__ table, double4 *__restrict__ acc,
    unsigned long long *__restrict__ counter,
    const int32_t *__restrict__ active_end, int32_t ti_current)
{
    __shared__ double tab[kNTab];
    for (int k = threadIdx.x; k < kNTab; k += kThreads) tab[k] = (double)table[k];
    __syncthreads();
    const uint32_t il = blockIdx.x * kThreads + threadIdx.x;
    const uint32_t i = i0 + min(il, n - 1);
    // Lanes past the end or outside the active set compute nothing but stay
    // for the wave-wide decisions below.
    bool live = il < n;
    if (live && active_end && active_end[perm[i]] != ti_current) live = false;

    const int M = pp.M;
    const double fac = pp.fac;
    const double4 p = spos[i];
    if (live && !(p.x >= 0 && p.x < pp.L && p.y >= 0 && p.y < pp.L && p.z >= 0 && p.z < pp.L)) {
        // Positions are wrapped on upload and by every drift; anything else
        // is left without a short-range force rather than indexing wildly.
        acc[perm[i]] = make_double4(0.0, 0.0, 0.0, 0.0);
        live = false;
    }
    int xl = (int)(fac * (p.x - pp.rcut) + M);
    xl -= M;
    int xr = (int)(fac * (p.x + pp.rcut));
    int yl = (int)(fac * (p.y - pp.rcut) + M);
    yl -= M;
    int yr = (int)(fac * (p.y + pp.rcut));
    int zl = (int)(fac * (p.z - pp.rcut) + M);
    zl -= M;
    int zr = (int)(fac * (p.z + pp.rcut));

    const float ux = (float)(fac * p.x), uy = (float)(fac * p.y), uz = (float)(fac * p.z);
    const float rc2 = (float)(pp.rcut * fac * pp.rcut * fac) * (1.0f + 1e-4f);

    double ax = 0, ay = 0, az = 0;
    uint32_t in_count = 0;

    // Wave-uniform walk when the wave's targets are close together (dense
    // regions: a handful of z cells of one column): the wave visits the union
    // of its lanes' cell ranges in the reference order (ix, iy, z ascending,
    // unwrapped) with uniform loop bounds and uniform j loads, and every lane
    // takes only the cells of its own range -- the pairs and their order for
    // each target are exactly those of the per-lane walk below.
    constexpr int kBig = 1 << 30;
    const int XL = wave_min_i(live ? xl : kBig), XR = wave_max_i(live ? xr : -kBig);
    const int YL = wave_min_i(live ? yl : kBig), YR = wave_max_i(live ? yr : -kBig);
    const int ZL = wave_min_i(live ? zl : kBig), ZR = wave_max_i(live ? zr : -kBig);
    if (XL > XR) return;  // no live lane
    const bool compact = XR - XL <= (xr - xl) + 1 && YR - YL <= (yr - yl) + 1 &&
                         ZR - ZL <= (zr - zl) + 8 && XR - XL < M && YR - YL < M && ZR - ZL < M;
    if (__builtin_amdgcn_readfirstlane((uint32_t)(__ballot(live && !compact) == 0ull))) {
        for (int ix = XL; ix <= XR; ++ix) {
            const float gx = gap_cells(ux, ix);
            const bool in_x = live && ix >= xl && ix <= xr && gx <= rc2;
            const int iix = wrap_index(ix, M);
            const double sx = image_shift(ix, M, pp.L);
            for (int iy = YL; iy <= YR; ++iy) {
                const double sy = image_shift(iy, M, pp.L);
                const float gxy = gx + gap_cells(uy, iy);
                int za = 1, zb = 0;
                if (in_x && iy >= yl && iy <= yr && gxy <= rc2) z_window(uz, gxy, rc2, zl, zr, za, zb);
                const int ZA = wave_min_i(za <= zb ? za : kBig);
                const int ZB = wave_max_i(za <= zb ? zb : -kBig);
                if (ZA > ZB) continue;
                const uint32_t col = ((uint32_t)iix * M + wrap_index(iy, M)) * (uint32_t)M;
                for (int z = ZA; z <= ZB; ++z) {
                    const bool in_z = z >= za && z <= zb;
                    const double sz = image_shift(z, M, pp.L);
                    uint32_t js, je;
                    load_range(off, col + wrap_index(z, M), js, je);
                    js = __builtin_amdgcn_readfirstlane(js);
                    je = __builtin_amdgcn_readfirstlane(je);
                    for (uint32_t j = js; j < je; ++j) {
                        const double4 q = spos[j];
                        if (in_z && j != i)
                            pair64<WRAP>(pp, tab, p.x, p.y, p.z, q, sx, sy, sz, ax, ay, az,
                                         in_count);
                    }
                }
            }
        }
    } else if (live) {
        for (int ix = xl; ix <= xr; ++ix) {
            const float gx = gap_cells(ux, ix);
            if (gx > rc2) continue;
            const int iix = wrap_index(ix, M);
            const double sx = image_shift(ix, M, pp.L);
            for (int iy = yl; iy <= yr; ++iy) {
                const float gxy = gx + gap_cells(uy, iy);
                if (gxy > rc2) continue;
                const int iiy = wrap_index(iy, M);
                const double sy = image_shift(iy, M, pp.L);
                int za, zb;
                z_window(uz, gxy, rc2, zl, zr, za, zb);
                if (za > zb) continue;
                const uint32_t col = ((uint32_t)iix * M + iiy) * (uint32_t)M;
                // Up to two contiguous runs in reference order (wrapped part
                // first when the range starts below 0): z cells za..zb, the
                // second run starting at the wrap.
                for (int z0 = za; z0 <= zb;) {
                    const int zw = wrap_index(z0, M);
                    const int z1 = min(zb, z0 + (M - 1 - zw));  // last z before a wrap
                    const double sz = image_shift(z0, M, pp.L);
                    const uint32_t js = off[col + zw];
                    const uint32_t je = off[col + zw + (z1 - z0) + 1];
                    for (uint32_t j = js; j < je; ++j) {
                        if (j == i) continue;
                        pair64<WRAP>(pp, tab, p.x, p.y, p.z, spos[j], sx, sy, sz, ax, ay,
                                     az, in_count);
                    }
                    z0 = z1 + 1;
                }
            }
        }
    }
    if (!live) return;
    acc[perm[i]] = make_double4(ax * pp.G, ay * pp.G, az * pp.G, 0.0);
    if (COUNT) atomicAdd(counter, (unsigned long long)in_count);
}

}  // namespace

// One box: its sorted particles are [i0, i0+n) and off points at the cell
// table entry of the box's first key (entries are global sorted indices).
int launch_pp64(const double4 *spos, const uint32_t *perm, const uint32_t *off,
                uint32_t i0, uint32_t n, const PairParams &pp, const float *table,
                double4 *acc, unsigned long long *pair_counter,
                const int32_t *active_end, int32_t ti_current, hipStream_t s)
{
    if (n == 0) return CUCO_OK;
    // Cell-level image shifts are exact when a neighbour cell's pairs all
    // lie within rcut + one cell of the target, below half the box by a
    // margin far above rounding; otherwise the per-pair tests run.
    const bool cell_images = pp.rcut + 1.001 * (pp.L / pp.M) < pp.boxhalf;
#define CUCO_PP64(COUNT, WRAP, CTR)                                                      \
    k_pp64<COUNT, WRAP><<<grid_for(n, kThreads), kThreads, 0, s>>>(                      \
        spos, perm, off, i0, n, pp, table, acc, CTR, active_end, ti_current)
    if (pair_counter) {
        if (cell_images) CUCO_PP64(true, false, pair_counter);
        else CUCO_PP64(true, true, pair_counter);
    } else {
        if (cell_images) CUCO_PP64(false, false, nullptr);
        else CUCO_PP64(false, true, nullptr);
    }
#undef CUCO_PP64
    CUCO_LAUNCH_CHECK();
    return CUCO_OK;
}

}  // namespace cuco
