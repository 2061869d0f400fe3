/*
 * cuco_oracle.c -- FP64 CPU restatement of cuco's gravity path.
 *
 * TEST INFRASTRUCTURE ONLY (see cuco_oracle.h): the parity checker and the
 * timed CPU baseline.  The product never links or calls this file.
 *
 * Build flags matter: the PP pair law is compiled with -ffp-contract=off so
 * that every multiply and add rounds separately, in the order the reference
 * C source writes them (gravlist.cpp:279-337).  The engine's FP64 kernel is
 * compiled the same way, which makes the two bit-identical.
 */
#include "cuco_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

static double g_pp_setup_s = 0, g_pp_eval_s = 0;

static double wall_seconds(void)
{
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0.0;
#endif
}

void orc_pp_last_timing(double *setup_s, double *eval_s)
{
    *setup_s = g_pp_setup_s;
    *eval_s = g_pp_eval_s;
}

static int orc_threads(const orc_params *p)
{
#ifdef _OPENMP
    return p->nthreads > 0 ? p->nthreads : omp_get_max_threads();
#else
    (void)p;
    return 1;
#endif
}

/* ------------------------------------------------------------------ */
/* Parameters and the short-range table                                */
/* ------------------------------------------------------------------ */

/* gravlist.cpp:56-59: table[i] = erfc(u) + 2u/sqrt(pi) exp(-u^2),
 * u = 3/NTAB*(i+0.5), stored as float (gravlist.cpp:15). */
void orc_shortrange_table(float *table)
{
    for (int i = 0; i < ORC_NTAB; ++i) {
        double u = 3.0 / ORC_NTAB * (i + 0.5);
        table[i] = (float)(erfc(u) + 2.0 * u / sqrt(M_PI) * exp(-u * u));
    }
}

/* gravlist.cpp:28-45 (softening), :56-70 (table, factors). */
void orc_set_params(orc_params *p, int32_t M, double L, double G,
                    double asmth, double rcut, double soft_h)
{
    int32_t nthreads = p->nthreads;
    memset(p, 0, sizeof(*p));
    p->nthreads = nthreads;
    p->M = M;
    p->L = L;
    p->G = G;
    p->asmth = asmth;
    p->rcut = rcut;
    p->rcut2 = rcut * rcut;
    p->asmthfac = 0.5 / asmth * (ORC_NTAB / 3.0);
    p->to_grid_fac = M / L;
    p->boxsize = L;
    p->boxhalf = L / 2.0;
    p->h = soft_h;
    p->h_inv = 1.0 / soft_h;
    p->h3_inv = p->h_inv * p->h_inv * p->h_inv;
    orc_shortrange_table(p->table);
}

/* ------------------------------------------------------------------ */
/* Cell keys, stable sort, cell table                                  */
/* ------------------------------------------------------------------ */

static inline int32_t cell_coord(double fac, double x, int32_t M)
{
    int32_t c = (int32_t)(fac * x);   /* gravlist_kernel.cu:111-113 */
    if (c >= M) c = M - 1;            /* clamp as pm_periodic.c:100-101 */
    if (c < 0) c = 0;
    return c;
}

/* gravlist_kernel.cu:92-118 */
void orc_calc_keys(const orc_params *p, const double *pos, int64_t n,
                   uint32_t *keys, uint32_t *index)
{
    const int32_t M = p->M;
    for (int64_t i = 0; i < n; ++i) {
        int32_t ix = cell_coord(p->to_grid_fac, pos[3 * i + 0], M);
        int32_t iy = cell_coord(p->to_grid_fac, pos[3 * i + 1], M);
        int32_t iz = cell_coord(p->to_grid_fac, pos[3 * i + 2], M);
        keys[i] = (uint32_t)(((int64_t)ix * M + iy) * M + iz);
        index[i] = (uint32_t)i;
    }
}

static int cmp_u64(const void *a, const void *b)
{
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return (x > y) - (x < y);
}

/* Stable sort == sort by (key, original position): radixsort.cu's per-digit
 * splits (:338-358) and scatter (:721-737) both preserve input order. */
void orc_sort_pairs(uint32_t *keys, uint32_t *values, int64_t n)
{
    uint64_t *tmp = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(n > 0 ? n : 1));
    uint32_t *vals = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; ++i) {
        tmp[i] = ((uint64_t)keys[i] << 32) | (uint64_t)i;
        vals[i] = values[i];
    }
    qsort(tmp, (size_t)n, sizeof(uint64_t), cmp_u64);
    for (int64_t i = 0; i < n; ++i) {
        uint32_t pos = (uint32_t)(tmp[i] & 0xffffffffu);
        keys[i] = (uint32_t)(tmp[i] >> 32);
        values[i] = vals[pos];
    }
    free(vals);
    free(tmp);
}

/* gravlist_kernel.cu:136-155 (cellStart at a key change, cellEnd at the
 * next change or at n); empty cells carry 0xffffffff (gravlist.cu:126). */
void orc_cell_table(const uint32_t *sorted_keys, int64_t n, int64_t ncells,
                    uint32_t *cell_start, uint32_t *cell_end)
{
    for (int64_t c = 0; c < ncells; ++c) {
        cell_start[c] = 0xffffffffu;
        cell_end[c] = 0xffffffffu;
    }
    for (int64_t i = 0; i < n; ++i) {
        uint32_t k = sorted_keys[i];
        if (i == 0 || k != sorted_keys[i - 1]) {
            cell_start[k] = (uint32_t)i;
            if (i > 0) cell_end[sorted_keys[i - 1]] = (uint32_t)i;
        }
        if (i == n - 1) cell_end[k] = (uint32_t)(i + 1);
    }
}

/* ------------------------------------------------------------------ */
/* PP short-range force                                                */
/* ------------------------------------------------------------------ */

/* One pair, gravlist.cpp:279-327 (minimum image, cut, spline, table).
 * Returns 1 when the pair is inside rcut (counts toward pairs_in). */
static inline int pair_accumulate(const orc_params *p, double pos_x,
                                  double pos_y, double pos_z, double xj,
                                  double yj, double zj, double mass,
                                  double *acc_x, double *acc_y, double *acc_z)
{
    double dx = xj - pos_x;
    double dy = yj - pos_y;
    double dz = zj - pos_z;
    if (dx > p->boxhalf) dx -= p->boxsize;
    if (dx < -p->boxhalf) dx += p->boxsize;
    if (dy > p->boxhalf) dy -= p->boxsize;
    if (dy < -p->boxhalf) dy += p->boxsize;
    if (dz > p->boxhalf) dz -= p->boxsize;
    if (dz < -p->boxhalf) dz += p->boxsize;
    double r2 = dx * dx + dy * dy + dz * dz;
    if (r2 > p->rcut2) return 0;
    double r = sqrt(r2);
    double fac;
    if (r >= p->h) {
        fac = mass / (r2 * r);
    } else {
        double u = r * p->h_inv;
        if (u < 0.5)
            fac = mass * p->h3_inv * (10.66667 + u * u * (32.0 * u - 38.4));
        else
            fac = mass * p->h3_inv *
                  (21.33333 - 48.0 * u + 38.4 * u * u - 10.66667 * u * u * u -
                   0.06667 / (u * u * u));
    }
    int tabindex = (int)(p->asmthfac * r);
    if (tabindex < ORC_NTAB) {
        fac *= p->table[tabindex];
        *acc_x += dx * fac;
        *acc_y += dy * fac;
        *acc_z += dz * fac;
    }
    return 1;
}

int orc_pp_force(const orc_params *p, const double *pos, const double *mass,
                 int64_t n, const int64_t *targets, int64_t ntargets,
                 double *acc, uint64_t *pairs_in)
{
    const int32_t M = p->M;
    const int64_t ncells = (int64_t)M * M * M;
    /* A cell range wider than the mesh would visit a cell twice. */
    if (2.0 * p->rcut * p->to_grid_fac + 2.0 > (double)M) return -2;

    const double t_start = wall_seconds();
    uint32_t *keys = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n ? n : 1));
    uint32_t *index = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n ? n : 1));
    uint32_t *cstart = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)ncells);
    uint32_t *cend = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)ncells);
    double *spos = (double *)malloc(sizeof(double) * 4 * (size_t)(n ? n : 1));
    if (!keys || !index || !cstart || !cend || !spos) {
        free(keys); free(index); free(cstart); free(cend); free(spos);
        return -1;
    }
    orc_calc_keys(p, pos, n, keys, index);
    orc_sort_pairs(keys, index, n);
    orc_cell_table(keys, n, ncells, cstart, cend);
    for (int64_t i = 0; i < n; ++i) {
        uint32_t o = index[i];
        spos[4 * i + 0] = pos[3 * o + 0];
        spos[4 * i + 1] = pos[3 * o + 1];
        spos[4 * i + 2] = pos[3 * o + 2];
        spos[4 * i + 3] = mass[o];
    }

    const double t_setup_done = wall_seconds();
    const int64_t nt = targets ? ntargets : n;
    uint64_t total_in = 0;
    const double fac = p->to_grid_fac, rcut = p->rcut;

#pragma omp parallel for schedule(dynamic, 64) reduction(+ : total_in) num_threads(orc_threads(p))
    for (int64_t t = 0; t < nt; ++t) {
        const int64_t target = targets ? targets[t] : t;
        const double pos_x = pos[3 * target + 0];
        const double pos_y = pos[3 * target + 1];
        const double pos_z = pos[3 * target + 2];
        double acc_x = 0, acc_y = 0, acc_z = 0;
        uint64_t in = 0;
        /* gravlist.cpp:233-243: (int) truncation after shifting by +M. */
        int xl = (int)(fac * (pos_x - rcut) + M); xl -= M;
        int xr = (int)(fac * (pos_x + rcut));
        int yl = (int)(fac * (pos_y - rcut) + M); yl -= M;
        int yr = (int)(fac * (pos_y + rcut));
        int zl = (int)(fac * (pos_z - rcut) + M); zl -= M;
        int zr = (int)(fac * (pos_z + rcut));
        for (int ix = xl; ix <= xr; ++ix) {
            int iix = ix;
            while (iix < 0) iix += M;
            while (iix >= M) iix -= M;
            for (int iy = yl; iy <= yr; ++iy) {
                int iiy = iy;
                while (iiy < 0) iiy += M;
                while (iiy >= M) iiy -= M;
                for (int iz = zl; iz <= zr; ++iz) {
                    int iiz = iz;
                    while (iiz < 0) iiz += M;
                    while (iiz >= M) iiz -= M;
                    int64_t ibox = ((int64_t)iix * M + iiy) * M + iiz;
                    uint32_t s = cstart[ibox];
                    if (s == 0xffffffffu) continue;
                    uint32_t e = cend[ibox];
                    for (uint32_t j = s; j < e; ++j) {
                        if ((int64_t)index[j] == target) continue; /* :274-277 */
                        in += (uint64_t)pair_accumulate(
                            p, pos_x, pos_y, pos_z, spos[4 * j + 0],
                            spos[4 * j + 1], spos[4 * j + 2], spos[4 * j + 3],
                            &acc_x, &acc_y, &acc_z);
                    }
                }
            }
        }
        acc[3 * t + 0] = acc_x * p->G; /* gravlist.cpp:335-337 */
        acc[3 * t + 1] = acc_y * p->G;
        acc[3 * t + 2] = acc_z * p->G;
        total_in += in;
    }
    if (pairs_in) *pairs_in = total_in;
    g_pp_setup_s = t_setup_done - t_start;
    g_pp_eval_s = wall_seconds() - t_setup_done;
    free(keys); free(index); free(cstart); free(cend); free(spos);
    return 0;
}

int orc_pp_force_bruteforce(const orc_params *p, const double *pos,
                            const double *mass, int64_t n, double *acc)
{
#pragma omp parallel for schedule(dynamic, 16) num_threads(orc_threads(p))
    for (int64_t i = 0; i < n; ++i) {
        double ax = 0, ay = 0, az = 0;
        for (int64_t j = 0; j < n; ++j) {
            if (j == i) continue;
            pair_accumulate(p, pos[3 * i], pos[3 * i + 1], pos[3 * i + 2],
                            pos[3 * j], pos[3 * j + 1], pos[3 * j + 2], mass[j],
                            &ax, &ay, &az);
        }
        acc[3 * i + 0] = ax * p->G;
        acc[3 * i + 1] = ay * p->G;
        acc[3 * i + 2] = az * p->G;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* FFT (mixed radix, recursive decimation in time)                     */
/* ------------------------------------------------------------------ */

static int64_t smallest_factor(int64_t n)
{
    if (n % 2 == 0) return 2;
    if (n % 3 == 0) return 3;
    if (n % 5 == 0) return 5;
    for (int64_t f = 7; f * f <= n; f += 2)
        if (n % f == 0) return f;
    return n;
}

/* out[k] = sum_j in[j*s] W^{jk}, W = tw[N/n] where tw[m] = exp(sign 2 pi i m/N). */
static void fft_rec(int64_t n, int64_t s, const cplx *in, cplx *out,
                    const cplx *tw, int64_t N)
{
    if (n == 1) { out[0] = in[0]; return; }
    int64_t pf = smallest_factor(n);
    int64_t m = n / pf;
    for (int64_t q = 0; q < pf; ++q)
        fft_rec(m, s * pf, in + q * s, out + q * m, tw, N);
    const int64_t step = N / n;
    cplx t_stack[64], y_stack[64];
    cplx *t = t_stack, *y = y_stack;
    if (pf > 64) {
        t = (cplx *)malloc(sizeof(cplx) * (size_t)pf * 2);
        y = t + pf;
    }
    for (int64_t k = 0; k < m; ++k) {
        for (int64_t q = 0; q < pf; ++q)
            t[q] = out[q * m + k] * tw[((q * k) % n) * step];
        for (int64_t r = 0; r < pf; ++r) {
            cplx acc = 0;
            for (int64_t q = 0; q < pf; ++q)
                acc += t[q] * tw[((q * r) % pf) * (N / pf)];
            y[r] = acc;
        }
        for (int64_t r = 0; r < pf; ++r) out[k + r * m] = y[r];
    }
    if (pf > 64) free(t);
}

static cplx *make_twiddles(int64_t n, int sign)
{
    cplx *tw = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    for (int64_t j = 0; j < n; ++j) {
        double a = 2.0 * M_PI * (double)j / (double)n;
        tw[j] = cos(a) + (double)sign * sin(a) * I;
    }
    return tw;
}

static void fft_line(int64_t n, cplx *x, const cplx *tw, cplx *work)
{
    fft_rec(n, 1, x, work, tw, n);
    memcpy(x, work, sizeof(cplx) * (size_t)n);
}

void orc_fft1d(int64_t n, double *data, int sign)
{
    cplx *tw = make_twiddles(n, sign);
    cplx *work = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    fft_line(n, (cplx *)data, tw, work);
    free(work);
    free(tw);
}

/* Complex FFT along y (stride Mz) and x (stride M*Mz) of an M x M x Mz array. */
static void fft_yx(int32_t M, cplx *c, int sign, int nthreads)
{
    const int64_t Mz = M / 2 + 1;
    cplx *tw = make_twiddles(M, sign);
#pragma omp parallel num_threads(nthreads)
    {
        cplx *line = (cplx *)malloc(sizeof(cplx) * (size_t)M);
        cplx *work = (cplx *)malloc(sizeof(cplx) * (size_t)M);
#pragma omp for collapse(2) schedule(static)
        for (int64_t x = 0; x < M; ++x)
            for (int64_t k = 0; k < Mz; ++k) {
                for (int64_t y = 0; y < M; ++y) line[y] = c[(x * M + y) * Mz + k];
                fft_line(M, line, tw, work);
                for (int64_t y = 0; y < M; ++y) c[(x * M + y) * Mz + k] = line[y];
            }
#pragma omp for collapse(2) schedule(static)
        for (int64_t y = 0; y < M; ++y)
            for (int64_t k = 0; k < Mz; ++k) {
                for (int64_t x = 0; x < M; ++x) line[x] = c[(x * M + y) * Mz + k];
                fft_line(M, line, tw, work);
                for (int64_t x = 0; x < M; ++x) c[(x * M + y) * Mz + k] = line[x];
            }
        free(work);
        free(line);
    }
    free(tw);
}

int orc_fft3d_r2c(int32_t M, const double *in, double *out)
{
    if (M < 2 || M % 2) return -1;
    const int64_t Mz = M / 2 + 1;
    cplx *c = (cplx *)out;
    cplx *tw = make_twiddles(M, -1);
    int nthreads = 1;
#ifdef _OPENMP
    nthreads = omp_get_max_threads();
#endif
#pragma omp parallel num_threads(nthreads)
    {
        cplx *line = (cplx *)malloc(sizeof(cplx) * (size_t)M);
        cplx *work = (cplx *)malloc(sizeof(cplx) * (size_t)M);
#pragma omp for schedule(static)
        for (int64_t r = 0; r < (int64_t)M * M; ++r) {
            for (int64_t z = 0; z < M; ++z) line[z] = in[r * M + z];
            fft_line(M, line, tw, work);
            for (int64_t k = 0; k < Mz; ++k) c[r * Mz + k] = line[k];
        }
        free(work);
        free(line);
    }
    free(tw);
    fft_yx(M, c, -1, nthreads);
    return 0;
}

int orc_fft3d_c2r(int32_t M, const double *in, double *out)
{
    if (M < 2 || M % 2) return -1;
    const int64_t Mz = M / 2 + 1;
    const int64_t nc = (int64_t)M * M * Mz;
    cplx *c = (cplx *)malloc(sizeof(cplx) * (size_t)nc);
    if (!c) return -1;
    memcpy(c, in, sizeof(cplx) * (size_t)nc);
    int nthreads = 1;
#ifdef _OPENMP
    nthreads = omp_get_max_threads();
#endif
    fft_yx(M, c, +1, nthreads);
    cplx *tw = make_twiddles(M, +1);
#pragma omp parallel num_threads(nthreads)
    {
        cplx *line = (cplx *)malloc(sizeof(cplx) * (size_t)M);
        cplx *work = (cplx *)malloc(sizeof(cplx) * (size_t)M);
#pragma omp for schedule(static)
        for (int64_t r = 0; r < (int64_t)M * M; ++r) {
            for (int64_t k = 0; k < Mz; ++k) line[k] = c[r * Mz + k];
            for (int64_t k = Mz; k < M; ++k) line[k] = conj(c[r * Mz + (M - k)]);
            fft_line(M, line, tw, work);
            for (int64_t z = 0; z < M; ++z) out[r * M + z] = creal(line[z]);
        }
        free(work);
        free(line);
    }
    free(tw);
    free(c);
    return 0;
}

/* ------------------------------------------------------------------ */
/* PM long-range force                                                 */
/* ------------------------------------------------------------------ */

/* pm_periodic.c:99-121: slab index, fraction, right neighbour. */
static inline void cic_axis(double fac, double x, int32_t M, int32_t *s,
                            int32_t *ss, double *d)
{
    int32_t slab = (int32_t)(fac * x);
    if (slab >= M) slab = M - 1;
    *d = fac * x - slab;
    int32_t slab2 = slab + 1;
    if (slab2 >= M) slab2 -= M;
    *s = slab;
    *ss = slab2;
}

/* pm_periodic.c:94-132 */
void orc_pm_deposit(const orc_params *p, const double *pos, const double *mass,
                    int64_t n, double *rho)
{
    const int64_t M = p->M;
    memset(rho, 0, sizeof(double) * (size_t)(M * M * M));
    const double fac = p->to_grid_fac;
    for (int64_t i = 0; i < n; ++i) {
        int32_t sx, sxx, sy, syy, sz, szz;
        double dx, dy, dz;
        cic_axis(fac, pos[3 * i + 0], (int32_t)M, &sx, &sxx, &dx);
        cic_axis(fac, pos[3 * i + 1], (int32_t)M, &sy, &syy, &dy);
        cic_axis(fac, pos[3 * i + 2], (int32_t)M, &sz, &szz, &dz);
        const double m = mass[i];
        rho[(sx * M + sy) * M + sz] += m * (1.0 - dx) * (1.0 - dy) * (1.0 - dz);
        rho[(sx * M + syy) * M + sz] += m * (1.0 - dx) * dy * (1.0 - dz);
        rho[(sx * M + sy) * M + szz] += m * (1.0 - dx) * (1.0 - dy) * dz;
        rho[(sx * M + syy) * M + szz] += m * (1.0 - dx) * dy * dz;
        rho[(sxx * M + sy) * M + sz] += m * dx * (1.0 - dy) * (1.0 - dz);
        rho[(sxx * M + syy) * M + sz] += m * dx * dy * (1.0 - dz);
        rho[(sxx * M + sy) * M + szz] += m * dx * (1.0 - dy) * dz;
        rho[(sxx * M + syy) * M + szz] += m * dx * dy * dz;
    }
}

/* pm_periodic.c:82-83 and :146-197 on the r2c output. */
static void apply_greens(const orc_params *p, cplx *c)
{
    const int64_t M = p->M, Mz = M / 2 + 1;
    double asmth2 = (2 * M_PI) * p->asmth / p->L;
    asmth2 *= asmth2;
#pragma omp parallel for schedule(static) num_threads(orc_threads(p))
    for (int64_t x = 0; x < M; ++x) {
        for (int64_t y = 0; y < M; ++y) {
            for (int64_t z = 0; z < Mz; ++z) {
                double kx = x > M / 2 ? (double)(x - M) : (double)x;
                double ky = y > M / 2 ? (double)(y - M) : (double)y;
                double kz = z > M / 2 ? (double)(z - M) : (double)z;
                double k2 = kx * kx + ky * ky + kz * kz;
                if (k2 > 0) {
                    double smth = -exp(-k2 * asmth2) / k2;
                    double fx = 1, fy = 1, fz = 1;
                    if (kx != 0) { fx = (M_PI * kx) / M; fx = sin(fx) / fx; }
                    if (ky != 0) { fy = (M_PI * ky) / M; fy = sin(fy) / fy; }
                    if (kz != 0) { fz = (M_PI * kz) / M; fz = sin(fz) / fz; }
                    double ff = 1 / (fx * fy * fz);
                    smth *= ff * ff * ff * ff;
                    c[(x * M + y) * Mz + z] *= smth;
                }
            }
        }
    }
    c[0] = 0; /* pm_periodic.c:197 */
}

int orc_pm_potential(const orc_params *p, const double *rho, double *phi)
{
    const int64_t M = p->M, Mz = M / 2 + 1;
    cplx *c = (cplx *)malloc(sizeof(cplx) * (size_t)(M * M * Mz));
    if (!c) return -1;
    int rc = orc_fft3d_r2c((int32_t)M, rho, (double *)c);
    if (rc == 0) {
        apply_greens(p, c);
        rc = orc_fft3d_c2r((int32_t)M, (const double *)c, phi);
    }
    free(c);
    return rc;
}

static inline int32_t wrapi(int32_t a, int32_t M)
{
    if (a >= M) a -= M;
    if (a < 0) a += M;
    return a;
}

/* pm_periodic.c:85-86 (fac), :205-271 (gradient), :272-309 (interpolation). */
int orc_pm_interpolate(const orc_params *p, const double *phi,
                       const double *pos, int64_t n, double *acc)
{
    const int32_t M = p->M;
    const int64_t M3 = (int64_t)M * M * M;
    double fac = p->G / (M_PI * p->L);
    fac *= 1 / (2 * p->L / M);
    double *fg = (double *)malloc(sizeof(double) * (size_t)M3);
    if (!fg) return -1;
    for (int dim = 0; dim < 3; ++dim) {
#pragma omp parallel for schedule(static) num_threads(orc_threads(p))
        for (int32_t x = 0; x < M; ++x) {
            for (int32_t y = 0; y < M; ++y) {
                for (int32_t z = 0; z < M; ++z) {
                    int32_t xl = x, xr = x, xll = x, xrr = x;
                    int32_t yl = y, yr = y, yll = y, yrr = y;
                    int32_t zl = z, zr = z, zll = z, zrr = z;
                    if (dim == 0) { xr = wrapi(x + 1, M); xrr = wrapi(x + 2, M); xl = wrapi(x - 1, M); xll = wrapi(x - 2, M); }
                    if (dim == 1) { yr = wrapi(y + 1, M); yrr = wrapi(y + 2, M); yl = wrapi(y - 1, M); yll = wrapi(y - 2, M); }
                    if (dim == 2) { zr = wrapi(z + 1, M); zrr = wrapi(z + 2, M); zl = wrapi(z - 1, M); zll = wrapi(z - 2, M); }
                    fg[((int64_t)x * M + y) * M + z] =
                        fac * ((4.0 / 3) * (phi[((int64_t)xl * M + yl) * M + zl] -
                                            phi[((int64_t)xr * M + yr) * M + zr]) -
                               (1.0 / 6) * (phi[((int64_t)xll * M + yll) * M + zll] -
                                            phi[((int64_t)xrr * M + yrr) * M + zrr]));
                }
            }
        }
#pragma omp parallel for schedule(static) num_threads(orc_threads(p))
        for (int64_t i = 0; i < n; ++i) {
            int32_t sx, sxx, sy, syy, sz, szz;
            double dx, dy, dz;
            cic_axis(p->to_grid_fac, pos[3 * i + 0], M, &sx, &sxx, &dx);
            cic_axis(p->to_grid_fac, pos[3 * i + 1], M, &sy, &syy, &dy);
            cic_axis(p->to_grid_fac, pos[3 * i + 2], M, &sz, &szz, &dz);
#define FG(a, b, c) fg[((int64_t)(a) * M + (b)) * M + (c)]
            double a = FG(sx, sy, sz) * (1.0 - dx) * (1.0 - dy) * (1.0 - dz);
            a += FG(sx, syy, sz) * (1.0 - dx) * dy * (1.0 - dz);
            a += FG(sx, sy, szz) * (1.0 - dx) * (1.0 - dy) * dz;
            a += FG(sx, syy, szz) * (1.0 - dx) * dy * dz;
            a += FG(sxx, sy, sz) * (dx) * (1.0 - dy) * (1.0 - dz);
            a += FG(sxx, syy, sz) * (dx) * dy * (1.0 - dz);
            a += FG(sxx, sy, szz) * (dx) * (1.0 - dy) * dz;
            a += FG(sxx, syy, szz) * (dx) * dy * dz;
#undef FG
            acc[3 * i + dim] = a;
        }
    }
    free(fg);
    return 0;
}

int orc_pm_force(const orc_params *p, const double *pos, const double *mass,
                 int64_t n, double *acc)
{
    const int64_t M3 = (int64_t)p->M * p->M * p->M;
    double *rho = (double *)malloc(sizeof(double) * (size_t)M3);
    double *phi = (double *)malloc(sizeof(double) * (size_t)M3);
    int rc = -1;
    if (rho && phi) {
        orc_pm_deposit(p, pos, mass, n, rho);
        rc = orc_pm_potential(p, rho, phi);
        if (rc == 0) rc = orc_pm_interpolate(p, phi, pos, n, acc);
    }
    free(rho);
    free(phi);
    return rc;
}

/* ------------------------------------------------------------------ */
/* Drift / kick tables                                                 */
/* ------------------------------------------------------------------ */

typedef struct { double om, ol, hubble; int kind; } cosmo_t;

/* driftfac.c:44-64 */
static double integrand(double a, const cosmo_t *c)
{
    double h = c->om / (a * a * a) + (1 - c->om - c->ol) / (a * a) + c->ol;
    h = c->hubble * sqrt(h);
    return c->kind == 0 ? 1 / (h * a * a * a) : 1 / (h * a * a);
}

/* Gauss-Kronrod 7/15 nodes and weights (QUADPACK qk15). */
static const double XGK[8] = {
    0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
    0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
    0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
    0.207784955007898467600689403773245, 0.000000000000000000000000000000000};
static const double WGK[8] = {
    0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
    0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
    0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
    0.204432940075298892414161999234649, 0.209482141084727828012999174891714};
static const double WG[4] = {
    0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
    0.381830050505118944950369775488975, 0.417959183673469387755102040816327};

static double gk15(double a, double b, const cosmo_t *c, double *err)
{
    double center = 0.5 * (a + b), half = 0.5 * (b - a);
    double fc = integrand(center, c);
    double resk = fc * WGK[7], resg = fc * WG[3];
    for (int j = 0; j < 7; ++j) {
        double x = half * XGK[j];
        double f = integrand(center - x, c) + integrand(center + x, c);
        resk += WGK[j] * f;
        if (j % 2 == 1) resg += WG[j / 2] * f;
    }
    *err = fabs((resk - resg) * half);
    return resk * half;
}

static double adapt(double a, double b, const cosmo_t *c, double tol, int depth)
{
    double err;
    double r = gk15(a, b, c, &err);
    if (err <= tol * fabs(r) || depth > 40) return r;
    double m = 0.5 * (a + b);
    return adapt(a, m, c, tol, depth + 1) + adapt(m, b, c, tol, depth + 1);
}

/* driftfac.c:14-42 */
void orc_drift_tables(double a_begin, double a_max, double omega0,
                      double omega_lambda, double hubble, double *drift,
                      double *kick)
{
    double lb = log(a_begin), lm = log(a_max);
    for (int i = 0; i < ORC_DRIFT_TABLE_LENGTH; ++i) {
        double hi = exp(lb + ((lm - lb) / ORC_DRIFT_TABLE_LENGTH) * (i + 1));
        cosmo_t cd = {omega0, omega_lambda, hubble, 0};
        cosmo_t ck = {omega0, omega_lambda, hubble, 1};
        drift[i] = adapt(exp(lb), hi, &cd, 1e-14, 0);
        kick[i] = adapt(exp(lb), hi, &ck, 1e-14, 0);
    }
}

/* driftfac.c:66-95 for one end point. */
static double table_at(const double *t, double lb, double lm, double loga)
{
    double u = (loga - lb) / (lm - lb) * ORC_DRIFT_TABLE_LENGTH;
    int i = (int)u;
    if (i >= ORC_DRIFT_TABLE_LENGTH) i = ORC_DRIFT_TABLE_LENGTH - 1;
    if (i < 1) return u * t[0];
    return t[i - 1] + (t[i] - t[i - 1]) * (u - i);
}

double orc_table_factor(const double *table, double a_begin, double a_max,
                        double loga0, double loga1)
{
    double lb = log(a_begin), lm = log(a_max);
    return table_at(table, lb, lm, loga1) - table_at(table, lb, lm, loga0);
}

/* ------------------------------------------------------------------ */
/* Global KDK                                                          */
/* ------------------------------------------------------------------ */

static void wrap_positions(double *pos, int64_t n, double L)
{
    for (int64_t i = 0; i < 3 * n; ++i) { /* predict.c:31-40 */
        while (pos[i] < 0) pos[i] += L;
        while (pos[i] >= L) pos[i] -= L;
    }
}

int orc_kdk_run(const orc_params *p, double *pos, double *vel,
                const double *mass, int64_t n, int32_t nsteps,
                const double *factors, double *acc_pp, double *acc_pm)
{
    wrap_positions(pos, n, p->L);
    int rc = orc_pp_force(p, pos, mass, n, NULL, 0, acc_pp, NULL);
    if (rc == 0) rc = orc_pm_force(p, pos, mass, n, acc_pm);
    for (int32_t s = 0; s < nsteps && rc == 0; ++s) {
        double k1 = factors[3 * s + 0], d = factors[3 * s + 1], k2 = factors[3 * s + 2];
        for (int64_t i = 0; i < 3 * n; ++i) {
            vel[i] += acc_pp[i] * k1; /* timestep.c:56-59 */
            vel[i] += acc_pm[i] * k1; /* timestep.c:86-92 */
        }
        for (int64_t i = 0; i < 3 * n; ++i) pos[i] += vel[i] * d; /* predict.c:17-21 */
        wrap_positions(pos, n, p->L);
        rc = orc_pp_force(p, pos, mass, n, NULL, 0, acc_pp, NULL);
        if (rc == 0) rc = orc_pm_force(p, pos, mass, n, acc_pm);
        if (rc) break;
        for (int64_t i = 0; i < 3 * n; ++i) {
            vel[i] += acc_pp[i] * k2;
            vel[i] += acc_pm[i] * k2;
        }
    }
    return rc;
}

/* ------------------------------------------------------------------ */
/* Integer timeline (individual power-of-two timesteps)                */
/* ------------------------------------------------------------------ */

#define ORC_TIMEBASE (1 << 28) /* allvars.h:11 */

/* get_drift_factor / get_gravkick_factor at two integer times
 * (driftfac.c:66-126). */
static double ti_factor(const double *table, double lb, double lm,
                        double interval, int time0, int time1)
{
    double a1 = lb + time0 * interval;
    double a2 = lb + time1 * interval;
    return table_at(table, lb, lm, a2) - table_at(table, lb, lm, a1);
}

int orc_timeline_run(const orc_params *p0, const double *tl,
                     const double *drift_tab, const double *kick_tab,
                     double *pos, double *vel, const double *mass, int64_t n,
                     int32_t *ti_beg, int32_t *ti_end, int32_t *state,
                     double *time, double *acc_pp, double *acc_pm,
                     int32_t max_steps, int32_t ti_stop, int32_t *status)
{
    const double time_begin = tl[0], time_max = tl[1], omega0 = tl[2];
    const double omega_lambda = tl[3], hubble = tl[4], err_tol = tl[5];
    const double max_dt = tl[6], min_dt = tl[7], soft_halo = tl[8];
    const double soft_max_phys = tl[9];
    const double lb = log(time_begin), lm = log(time_max);
    const double interval = (log(time_max) - log(time_begin)) / ORC_TIMEBASE; /* init.c:23 */
    orc_params p = *p0;
    int32_t ti_current = state[0], pm_beg = state[1], pm_end = state[2];
    int32_t steps = state[3];
    int finished = 0, at_stop = 0, flags = 0, rc = 0;
    int64_t *active = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    if (!active) return -1;
    for (int32_t k = 0; k < max_steps && !finished && rc == 0; ++k) {
        /* find_next_sync_point_and_drift (run.c:29-64) */
        int32_t min_glob = ti_end[0];
        for (int64_t i = 1; i < n; ++i)
            if (min_glob > ti_end[i]) min_glob = ti_end[i];
        if (min_glob >= pm_end) min_glob = pm_end;
        int32_t target = min_glob;
        if (ti_stop > ti_current && min_glob >= ti_stop) target = ti_stop;
        double d = ti_factor(drift_tab, lb, lm, interval, ti_current, target);
        for (int64_t i = 0; i < 3 * n; ++i) pos[i] += vel[i] * d; /* predict.c:17-21 */
        wrap_positions(pos, n, p.L);
        ti_current = target;
        *time = time_begin * exp(ti_current * interval);
        if (target != min_glob) {
            at_stop = 1;
            break;
        }
        /* set_softenings (gravlist.cpp:28-45) */
        double a = *time;
        double soft = soft_halo * a > soft_max_phys ? soft_max_phys / a : soft_halo;
        orc_set_params(&p, p0->M, p0->L, p0->G, p0->asmth, p0->rcut, 2.8 * soft);
        p.nthreads = p0->nthreads;
        /* compute_accelerations (accel.c:8-23): PM on PM steps; the
         * short-range force of the particles that will be kicked. */
        if (pm_end == ti_current) rc = orc_pm_force(&p, pos, mass, n, acc_pm);
        int64_t na = 0;
        for (int64_t i = 0; i < n; ++i)
            if (ti_end[i] == ti_current) active[na++] = i;
        if (rc == 0 && na > 0) {
            double *tmp = (double *)malloc(sizeof(double) * 3 * (size_t)na);
            if (!tmp) { rc = -1; break; }
            rc = orc_pp_force(&p, pos, mass, n, active, na, tmp, NULL);
            for (int64_t k2 = 0; k2 < na; ++k2)
                for (int dd = 0; dd < 3; ++dd) acc_pp[3 * active[k2] + dd] = tmp[3 * k2 + dd];
            free(tmp);
        }
        if (rc) break;
        /* advance_and_find_timesteps (timestep.c:11-94) */
        double fac1 = 1 / (a * a);
        double hubble_a = omega0 / (a * a * a) + (1 - omega0 - omega_lambda) / (a * a) + omega_lambda;
        hubble_a = hubble * sqrt(hubble_a);
        for (int64_t k2 = 0; k2 < na; ++k2) {
            int64_t i = active[k2];
            const double *g = acc_pp + 3 * i, *pm = acc_pm + 3 * i;
            /* get_timestep (timestep.c:96-138) */
            double ax = fac1 * g[0], ay = fac1 * g[1], az = fac1 * g[2];
            ax += fac1 * pm[0];
            ay += fac1 * pm[1];
            az += fac1 * pm[2];
            double ac = sqrt(ax * ax + ay * ay + az * az);
            if (ac == 0) ac = 1.0e-30;
            double dt = sqrt(2 * err_tol * a * soft / ac);
            dt *= hubble_a;
            if (dt >= max_dt) dt = max_dt;
            if (dt < min_dt) dt = min_dt;
            double st = dt / interval;
            if (!(st > 0 && st < (double)ORC_TIMEBASE)) { flags = 1; continue; }
            int ti_step = (int)st;
            int ti_min = ORC_TIMEBASE;
            while (ti_min > ti_step) ti_min >>= 1;
            ti_step = ti_min;
            int end = ti_end[i], beg = ti_beg[i];
            if (ti_step > (end - beg)) {
                if (((ORC_TIMEBASE - end) % ti_step) > 0) ti_step = end - beg;
            }
            if (ti_current == ORC_TIMEBASE) ti_step = 0;
            if ((ORC_TIMEBASE - ti_current) < ti_step) ti_step = ORC_TIMEBASE - ti_current;
            int tstart = (beg + end) / 2, tend = end + ti_step / 2;
            double kf = ti_factor(kick_tab, lb, lm, interval, tstart, tend);
            ti_beg[i] = end;
            ti_end[i] = end + ti_step;
            for (int dd = 0; dd < 3; ++dd) vel[3 * i + dd] += g[dd] * kf;
        }
        if (pm_end == ti_current) {
            int ti_step = ORC_TIMEBASE;
            while (ti_step > (max_dt / interval)) ti_step >>= 1;
            if (ti_step > (pm_end - pm_beg)) {
                if (((ORC_TIMEBASE - pm_end) % ti_step) > 0) ti_step = pm_end - pm_beg;
            }
            if (ti_current == ORC_TIMEBASE) ti_step = 0;
            int tstart = (pm_beg + pm_end) / 2, tend = pm_end + ti_step / 2;
            double kf = ti_factor(kick_tab, lb, lm, interval, tstart, tend);
            pm_beg = pm_end;
            pm_end = pm_beg + ti_step;
            for (int64_t i = 0; i < 3 * n; ++i) vel[i] += acc_pm[i] * kf;
        }
        steps++;
        if (flags) break;
        if (!(ti_current < ORC_TIMEBASE && *time <= time_max)) finished = 1; /* run.c:24 */
    }
    free(active);
    state[0] = ti_current;
    state[1] = pm_beg;
    state[2] = pm_end;
    state[3] = steps;
    status[0] = finished;
    status[1] = at_stop;
    status[2] = flags;
    return rc;
}
