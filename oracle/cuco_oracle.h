/*
 * cuco_oracle.h -- FP64 CPU restatement of cuco's per-step gravity path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the
 * MI355X force engine in cuco_amd/.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it, and there only to check or
 * to time the CPU baseline.  Nothing in the product links or calls it.
 *
 * Every function restates the reference algorithm from the cited lines of
 * liulei/cuco (paths relative to the reference root).  It is written from the
 * arithmetic those lines perform; no reference source is compiled or copied.
 *
 * Deliberate differences from the reference (each one is a fix or a
 * precision upgrade that the FP64 path of the engine shares):
 *   - grids, FFT and all PM arithmetic in double (reference: float grids and
 *     single-precision cuFFT, pm_periodic.c:19-21,140,201);
 *   - mesh factor M/L in double (reference: float to_slab_fac, pm_periodic.c:23);
 *   - cell keys are clamped to M-1 as the CIC slab is (pm_periodic.c:100-101);
 *     the reference key has no clamp (gravlist_kernel.cu:109-114);
 *   - the pair law uses the SOURCE mass m_j (reference CPU uses the target mass,
 *     gravlist.cpp:231; the GPU uses the global header mass, gravlist.cpp:111);
 *   - particles inside a cell are visited in stable-sorted order, i.e.
 *     ascending original index, which is the GPU path's order
 *     (gravlist_kernel.cu:195-213); the CPU linked list visits them in
 *     descending order (gravlist.cpp:139-140);
 *   - every particle is a target (the reference CPU path evaluates only
 *     active particles, gravlist.cpp:208);
 *   - drift/kick tables are double and built with an adaptive Gauss-Kronrod
 *     rule (reference: float tables from GSL qag GAUSS41, driftfac.c:28-39).
 *
 * Parity status of each function is listed beside its declaration.
 */
#ifndef CUCO_ORACLE_H
#define CUCO_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NTAB 1000               /* allvars.h:27 */
#define ORC_DRIFT_TABLE_LENGTH 1000 /* allvars.h:29 */

typedef struct orc_params {
    int32_t M;           /* PMGRID, mesh cells per dimension (allvars.h:9) */
    int32_t nthreads;    /* OpenMP threads used by the PP/FFT loops (0 = all) */
    double L;            /* All.BoxSize */
    double G;            /* All.G */
    double asmth;        /* All.Asmth = ASMTH*L/M (pm_periodic.c:27) */
    double rcut;         /* All.Rcut = RCUT*Asmth (pm_periodic.c:28) */
    double rcut2;        /* gravlist.cpp:63-64 */
    double asmthfac;     /* 0.5/asmth*(NTAB/3) (gravlist.cpp:66-67) */
    double to_grid_fac;  /* M/L (gravlist.cpp:61) */
    double boxsize;      /* gravlist.cpp:69 */
    double boxhalf;      /* gravlist.cpp:70 */
    double h;            /* ForceSoftening = 2.8*SofteningTable (gravlist.cpp:36-38) */
    double h_inv;        /* gravlist.cpp:39 */
    double h3_inv;       /* gravlist.cpp:40 */
    float table[ORC_NTAB]; /* short-range erfc table (gravlist.cpp:56-59) */
} orc_params;

/* gravlist.cpp:56-59.  Pinned: tests compare against scipy erfc/exp. */
void orc_shortrange_table(float *table);

/* gravlist.cpp:28-45 and :48-115, pm_periodic.c:25-43.  asmth and rcut are
 * lengths; soft_h is the spline softening h (= 2.8*epsilon).
 * Pinned: closed-form checks on every derived scalar in tests. */
void orc_set_params(orc_params *p, int32_t M, double L, double G,
                    double asmth, double rcut, double soft_h);

/* Cell key (ix*M+iy)*M+iz with ix=(int)(M/L*x) clamped to M-1
 * (gravlist_kernel.cu:92-118, clamp from pm_periodic.c:99-101); value is the
 * identity index.  Pinned: brute-force recomputation in tests. */
void orc_calc_keys(const orc_params *p, const double *pos, int64_t n,
                   uint32_t *keys, uint32_t *index);

/* Stable key/value sort: the permutation radixsort.cu's stable LSD sort
 * produces (radixsort.cu:338-409,668-737,1082-1121).  Implemented as a sort
 * of the composite (key<<32|value) with value = original index, which is the
 * closed form of a stable sort.  Pinned: brute force (insertion sort) on
 * small inputs and the sortedness/permutation invariants. */
void orc_sort_pairs(uint32_t *keys, uint32_t *values, int64_t n);

/* Cell start/end table (gravlist_kernel.cu:121-160): start[c] = first sorted
 * index with key c, end[c] = one past the last; 0xffffffff in both for empty
 * cells.  Pinned: invariants sum(end-start)=n, keys monotone. */
void orc_cell_table(const uint32_t *sorted_keys, int64_t n, int64_t ncells,
                    uint32_t *cell_start, uint32_t *cell_end);

/* Short-range PP force (gravlist.cpp:216-338 pair law and cell loop; self
 * skip gravlist.cpp:274-277).  pos is n*3 (x,y,z), mass n.  If targets is
 * NULL every particle is a target and acc is n*3 in original order;
 * otherwise acc is ntargets*3 for the listed original indices.
 * *pairs_in (optional) receives the number of ordered pairs with r2<=rcut2.
 * Pinned: O(N^2) brute-force minimum-image sum, two-particle closed form,
 * Newton's third law, zero force on a perfect lattice. */
int orc_pp_force(const orc_params *p, const double *pos, const double *mass,
                 int64_t n, const int64_t *targets, int64_t ntargets,
                 double *acc, uint64_t *pairs_in);

/* Wall-clock seconds of the last orc_pp_force: cell-list set-up (keys, sort,
 * table, reorder) and the pair loop.  Used to time the CPU baseline. */
void orc_pp_last_timing(double *setup_s, double *eval_s);

/* Brute-force O(N^2) twin of orc_pp_force (every pair, minimum image, same
 * pair law) used to pin the cell-list traversal. */
int orc_pp_force_bruteforce(const orc_params *p, const double *pos,
                            const double *mass, int64_t n, double *acc);

/* PM steps (pm_periodic.c:67-314), double everywhere. */
/* CIC deposit into an unpadded M^3 grid (pm_periodic.c:97-132).
 * Pinned: mass conservation, node-located particle, closed form weights. */
void orc_pm_deposit(const orc_params *p, const double *pos, const double *mass,
                    int64_t n, double *rho);
/* rho (M^3 real) -> potential phi (M^3 real): r2c, Green's function with the
 * erfc long-range filter and CIC^-2 deconvolution, DC zeroed, c2r
 * (pm_periodic.c:138-203).  phi is the unnormalized c2r output, exactly as
 * the reference's forcegrid input.  Pinned: single Fourier mode closed form. */
int orc_pm_potential(const orc_params *p, const double *rho, double *phi);
/* Four-point FD gradient (pm_periodic.c:205-271) and CIC interpolation
 * (pm_periodic.c:272-309) of phi onto the particles: acc is n*3.
 * Pinned: closed form gradient of a sine mode; interpolation symmetry. */
int orc_pm_interpolate(const orc_params *p, const double *phi,
                       const double *pos, int64_t n, double *acc);
/* Whole PM force = deposit + potential + interpolate. */
int orc_pm_force(const orc_params *p, const double *pos, const double *mass,
                 int64_t n, double *acc);

/* Complex FFT of length n (forward sign=-1: exp(-2 pi i jk/n)), unnormalized,
 * any n (mixed radix 2/3/5, direct DFT for other prime factors).
 * data is interleaved re,im.  Pinned: numpy.fft golden vectors. */
void orc_fft1d(int64_t n, double *data, int sign);
/* 3D real-to-complex FFT, in: M^3 real (x-major), out: M*M*(M/2+1) complex
 * interleaved.  Pinned: numpy.fft.rfftn golden vectors, Parseval. */
int orc_fft3d_r2c(int32_t M, const double *in, double *out);
/* 3D complex-to-real inverse (unnormalized: c2r(r2c(x)) = M^3 x).
 * Pinned: numpy.fft.irfftn golden vectors, round trip. */
int orc_fft3d_c2r(int32_t M, const double *in, double *out);

/* Drift/kick tables (driftfac.c:14-64): drift[i] = int_{a0}^{a_i} da/(H a^3),
 * kick[i] = int_{a0}^{a_i} da/(H a^2), a_i = exp(ln a0 + (ln amax-ln a0)/1000*(i+1)),
 * H = hubble*sqrt(om/a^3 + (1-om-ol)/a^2 + ol).
 * Pinned: Einstein-de Sitter closed form, scipy.integrate.quad fixtures. */
void orc_drift_tables(double a_begin, double a_max, double omega0,
                      double omega_lambda, double hubble,
                      double *drift, double *kick);
/* Table difference between ln a0 and ln a1 (driftfac.c:66-126). */
double orc_table_factor(const double *table, double a_begin, double a_max,
                        double loga0, double loga1);

/* Global KDK steps with device-identical arithmetic (predict.c:10-41 drift
 * and wrap, timestep.c:51-59,81-92 kicks, first PP then PM):
 *   for s: v += (a_pp)*k1[s]; v += (a_pm)*k1[s]; x += v*d[s]; wrap;
 *          forces; v += a_pp*k2[s]; v += a_pm*k2[s]
 * The initial forces are computed from the input positions.  pos/vel are
 * updated in place; acc_pp/acc_pm receive the last forces (n*3 each).
 * factors is nsteps*3 (k1, d, k2). */
int orc_kdk_run(const orc_params *p, double *pos, double *vel,
                const double *mass, int64_t n, int32_t nsteps,
                const double *factors, double *acc_pp, double *acc_pm);

/* The reference's main loop with individual power-of-two timesteps:
 * up to max_steps passes of run()'s body (run.c:9-25): next sync point and
 * drift (run.c:29-64, with box wrapping), set_softenings (gravlist.cpp:
 * 28-45), PM on PM steps and the short-range force of the particles whose
 * step ends now (accel.c:8-23), advance_and_find_timesteps with get_timestep
 * (timestep.c:11-138).  tl = {TimeBegin, TimeMax, Omega0, OmegaLambda,
 * Hubble, ErrTolIntAccuracy, MaxSizeTimestep, MinSizeTimestep,
 * SofteningHalo, SofteningHaloMaxPhys}; drift_tab/kick_tab as
 * orc_drift_tables; state = {Ti_Current, PM_Ti_begstep, PM_Ti_endstep,
 * NumCurrentTiStep} in/out; *time = TimeBegin exp(Ti_Current interval)
 * (double, the reference stores float).  When the next sync point reaches
 * ti_stop > Ti_Current the particles are drifted to ti_stop and the call
 * returns (status[1] = 1).  status = {finished, at_stop, flags}; flags = 1
 * when a step left the integer timeline (the reference exits there).
 * acc_pp keeps its old value for particles that were not kicked. */
int orc_timeline_run(const orc_params *p, const double *tl,
                     const double *drift_tab, const double *kick_tab,
                     double *pos, double *vel, const double *mass, int64_t n,
                     int32_t *ti_beg, int32_t *ti_end, int32_t *state,
                     double *time, double *acc_pp, double *acc_pm,
                     int32_t max_steps, int32_t ti_stop, int32_t *status);

#ifdef __cplusplus
}
#endif
#endif
