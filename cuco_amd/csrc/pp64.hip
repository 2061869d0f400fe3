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

__device__ inline void pair64(const PairParams &pp, const double *__restrict__ tab,
                              double px, double py, double pz, double4 q,
                              double &ax, double &ay, double &az,
                              uint32_t &in_count)
{
    double dx = q.x - px;
    double dy = q.y - py;
    double dz = q.z - pz;
    if (dx > pp.boxhalf) dx -= pp.L;
    if (dx < -pp.boxhalf) dx += pp.L;
    if (dy > pp.boxhalf) dy -= pp.L;
    if (dy < -pp.boxhalf) dy += pp.L;
    if (dz > pp.boxhalf) dz -= pp.L;
    if (dz < -pp.boxhalf) dz += pp.L;
    double r2 = dx * dx + dy * dy + dz * dz;
    if (r2 > pp.rcut2) return;
    ++in_count;
    double r = sqrt(r2);
    double fac;
    if (r >= pp.h) {
        fac = q.w / (r2 * r);
    } else {
        double u = r * pp.h_inv;
        if (u < 0.5)
            fac = q.w * pp.h3_inv * (10.66667 + u * u * (32.0 * u - 38.4));
        else
            fac = q.w * pp.h3_inv *
                  (21.33333 - 48.0 * u + 38.4 * u * u - 10.66667 * u * u * u -
                   0.06667 / (u * u * u));
    }
    int tabindex = (int)(pp.asmthfac * r);
    if (tabindex < kNTab) {
        fac *= tab[tabindex];
        ax += dx * fac;
        ay += dy * fac;
        az += dz * fac;
    }
}

// Squared distance in cell units from u to the cell interval [c, c+1),
// shrunk by a tolerance far above FP64 rounding of fac*x.
__device__ inline double gap_cells(double u, int c)
{
    double g = fmax(0.0, fmax((double)c - u, u - (double)(c + 1)));
    g = fmax(0.0, g - 1e-6);
    return g * g;
}

template <bool COUNT>
__global__ __launch_bounds__(kThreads) void k_pp64(
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
    if (il >= n) return;
    const uint32_t i = i0 + il;
    if (active_end && active_end[perm[i]] != ti_current) return;

    const int M = pp.M;
    const double fac = pp.fac;
    const double4 p = spos[i];
    if (!(p.x >= 0 && p.x < pp.L && p.y >= 0 && p.y < pp.L && p.z >= 0 && p.z < pp.L)) {
        // Positions are wrapped on upload and by every drift; anything else
        // is left without a short-range force rather than indexing wildly.
        acc[perm[i]] = make_double4(0.0, 0.0, 0.0, 0.0);
        return;
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

    const double ux = fac * p.x, uy = fac * p.y, uz = fac * p.z;
    const double rc = pp.rcut * fac;
    const double rc2 = rc * rc;

    double ax = 0, ay = 0, az = 0;
    uint32_t in_count = 0;
    for (int ix = xl; ix <= xr; ++ix) {
        const double gx = gap_cells(ux, ix);
        if (gx > rc2) continue;
        const int iix = wrap_index(ix, M);
        for (int iy = yl; iy <= yr; ++iy) {
            const double gxy = gx + gap_cells(uy, iy);
            if (gxy > rc2) continue;
            const int iiy = wrap_index(iy, M);
            const double zext = sqrt(rc2 - gxy);
            const int za = max(zl, (int)floor(uz - zext - 1e-6));
            const int zb = min(zr, (int)floor(uz + zext + 1e-6));
            if (za > zb) continue;
            const uint32_t col = ((uint32_t)iix * M + iiy) * (uint32_t)M;
            // Up to two contiguous runs in reference order (wrapped part first
            // when the range starts below 0).
            int ra[2], rb[2], nr = 0;
            if (za < 0 && zb >= 0) {
                ra[0] = za + M; rb[0] = M - 1;
                ra[1] = 0; rb[1] = zb; nr = 2;
            } else if (za < M && zb >= M) {
                ra[0] = za; rb[0] = M - 1;
                ra[1] = 0; rb[1] = zb - M; nr = 2;
            } else {
                ra[0] = wrap_index(za, M); rb[0] = wrap_index(zb, M); nr = 1;
            }
            for (int q = 0; q < nr; ++q) {
                const uint32_t js = off[col + ra[q]];
                const uint32_t je = off[col + rb[q] + 1];
                for (uint32_t j = js; j < je; ++j) {
                    if (j == i) continue;
                    pair64(pp, tab, p.x, p.y, p.z, spos[j], ax, ay, az, in_count);
                }
            }
        }
    }
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
    if (pair_counter)
        k_pp64<true><<<grid_for(n, kThreads), kThreads, 0, s>>>(
            spos, perm, off, i0, n, pp, table, acc, pair_counter, active_end,
            ti_current);
    else
        k_pp64<false><<<grid_for(n, kThreads), kThreads, 0, s>>>(
            spos, perm, off, i0, n, pp, table, acc, nullptr, active_end,
            ti_current);
    CUCO_LAUNCH_CHECK();
    return CUCO_OK;
}

}  // namespace cuco
