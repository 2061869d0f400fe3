/*
 * cuco_engine.h -- C ABI of the MI355X force engine for cuco's PM+P3M gravity.
 *
 * Plain C types only: opaque handle, pointers, sizes, int status returns.
 * The library never exits the process; every entry point returns a
 * cuco_status and cuco_last_error() gives the message of the last failure on
 * the calling thread.  One HIP stream per handle; a handle is not
 * thread-safe.  All host arrays are caller-owned; device arrays behind a
 * handle are engine-owned and stay resident across drift/kick/force calls.
 *
 * Each entry point names the reference interface it replaces (file:line in
 * liulei/cuco).  Units are the caller's (the reference uses kpc/h,
 * 1e10 Msun/h, km/s: begrun.c:33-35,65-70).
 */
#ifndef CUCO_ENGINE_H
#define CUCO_ENGINE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum cuco_status {
    CUCO_OK = 0,
    CUCO_ERR_INVALID_ARG = 1,  /* bad size, null pointer, unsupported value */
    CUCO_ERR_HIP = 2,          /* a HIP runtime call failed */
    CUCO_ERR_OUT_OF_MEMORY = 3,
    CUCO_ERR_STATE = 4,        /* e.g. force requested before particles exist */
    CUCO_ERR_UNSUPPORTED = 5   /* e.g. mesh size with a prime factor > 5 */
} cuco_status;

typedef enum cuco_precision {
    CUCO_FP64 = 0, /* plain path: FP64 everywhere, reference cell order */
    CUCO_FP32 = 1  /* chosen path: FP32 cell-relative pairs, FP32 mesh;
                      its pair force takes rcut up to 475 mesh cells
                      (CUCO_ERR_UNSUPPORTED beyond) */
} cuco_precision;

/* Split and mesh parameters (allvars.h:9,24-25; pm_periodic.c:27-28;
 * gravlist.cpp:36-40,61-70).  asmth and rcut are lengths; soft_h is the
 * spline softening length h = 2.8*epsilon (All.ForceSoftening). */
typedef struct cuco_params {
    int32_t pm_grid;   /* PMGRID: mesh cells per dimension (even, 2^a 3^b 5^c) */
    int32_t reserved0; /* must be 0 */
    double box_size;   /* All.BoxSize */
    double G;          /* All.G */
    double asmth;      /* All.Asmth = ASMTH*BoxSize/PMGRID */
    double rcut;       /* All.Rcut = RCUT*Asmth */
    double soft_h;     /* All.ForceSoftening */
} cuco_params;

/* PARTICLE record, byte-compatible with allvars.h:35-45 (60 bytes). */
typedef struct cuco_particle {
    float Pos[3];
    float Mass;
    float Vel[3];
    float GravAccel[3];
    float GravPM[3];
    int32_t Ti_endstep;
    int32_t Ti_begstep;
} cuco_particle;

/* HEADER record of a snapshot / IC file, byte-compatible with
 * allvars.h:47-57 (32 bytes). */
typedef struct cuco_header {
    int32_t NumPart;
    float Mass;
    float Time;
    float Redshift;
    float BoxSize;
    float Omega0;
    float OmegaLambda;
    float HubbleParam;
} cuco_header;

typedef struct cuco_engine cuco_engine;

/* ---------------------------------------------------------------- */
/* Errors                                                            */
/* ---------------------------------------------------------------- */
const char *cuco_status_string(int status);
const char *cuco_last_error(void);

/* ---------------------------------------------------------------- */
/* Lifecycle                                                         */
/* ---------------------------------------------------------------- */

/* Replaces cudaInit (gravlist.cu:25-31), linklist_init's device set-up
 * (gravlist.cpp:48-115), pm_init_periodic (pm_periodic.c:25-43) and the
 * RadixSort constructor (radixsort.h:140-150).  Allocates every device
 * buffer for up to max_particles particles on HIP device `device` (about
 * 265 bytes per particle plus the cell table; the PM meshes and deposit
 * scratch of each precision are allocated on its first use). */
int cuco_create(const cuco_params *params, int64_t max_particles,
                int32_t device, cuco_engine **out);
/* Frees everything the handle owns (the reference never frees). */
int cuco_destroy(cuco_engine *e);
/* set_softenings + setSoftParam (gravlist.cpp:28-45, gravlist.cu:179-181):
 * re-sets the spline softening h (comoving), e.g. each step, in every box. */
int cuco_set_softening(cuco_engine *e, double soft_h);
/* setSimParam (gravlist.cu:173-177) with linklist_init's derived constants
 * (gravlist.cpp:61-70) and pm_init_periodic's split scale
 * (pm_periodic.c:27-28): replaces G, asmth, rcut and soft_h of box `box`
 * (-1: every box).  pm_grid and box_size must equal the values the handle
 * was created with. */
int cuco_set_params(cuco_engine *e, int32_t box, const cuco_params *params);

/* ---------------------------------------------------------------- */
/* Ensembles of independent boxes                                    */
/* ---------------------------------------------------------------- */

/* An ensemble advances nboxes independent periodic boxes together: each
 * stage of a force evaluation or KDK step (keys, sort, cell table, pair
 * forces, deposit, each FFT pass, interpolation, drift, kick) is one launch
 * over all boxes, and cuco_kdk_steps captures the whole step as one graph.
 * Box b has its own params[b] (mesh, box size, split scale, softening) and
 * exactly n_per_box[b] particles; in every particle array of the handle
 * (upload, download, accelerations) its particles follow those of boxes
 * 0..b-1.  Every cuco_* call that takes a handle accepts an ensemble
 * handle; cuco_upload_particles then takes exactly sum(n_per_box)
 * particles.  Each box gives the same forces, bit for bit, as a
 * single-box handle holding the same particles.  The reference runs one box
 * per process; this has no counterpart there. */
int cuco_ensemble_create(const cuco_params *params, const int64_t *n_per_box,
                         int32_t nboxes, int32_t device, cuco_engine **out);
/* Number of boxes and (if n_per_box is not NULL) the particles of each. */
int cuco_ensemble_size(cuco_engine *e, int32_t *nboxes, int64_t *n_per_box);

/* ---------------------------------------------------------------- */
/* Particle state (device resident, FP64)                            */
/* ---------------------------------------------------------------- */

/* Replaces copyPosToDevice (gravlist.cu:56-68): pos and vel are n*3
 * doubles (x,y,z per particle); vel may be NULL (zero); mass may be NULL,
 * in which case every particle has uniform_mass (header.Mass,
 * gravlist.cpp:111).  Positions are wrapped into [0, BoxSize). */
int cuco_upload_particles(cuco_engine *e, int64_t n, const double *pos,
                          const double *vel, const double *mass,
                          double uniform_mass);
/* Reads back positions and velocities (n*3 doubles each; either may be NULL). */
int cuco_download_particles(cuco_engine *e, double *pos, double *vel);
/* Replaces copyAccelFromDevice (gravlist.cu:70-83) and the GravPM read-out
 * (pm_periodic.c:308): n*3 doubles in original particle order, from the
 * last force evaluation of either precision.  Either pointer may be NULL. */
int cuco_download_accel(cuco_engine *e, double *acc_pp, double *acc_pm);

/* ---------------------------------------------------------------- */
/* Forces                                                            */
/* ---------------------------------------------------------------- */

/* Replaces gravity_linklist_cuda (gravlist.cpp:148-193): cell keys,
 * stable radix sort, cell table, reorder, erfc-truncated PP pair force. */
int cuco_pp_force(cuco_engine *e, int32_t precision);
/* Replaces pmforce_periodic (pm_periodic.c:67-314): CIC deposit, r2c FFT,
 * Green's function, c2r FFT, 4-point gradient, CIC interpolation. */
int cuco_pm_force(cuco_engine *e, int32_t precision);
/* Replaces compute_accelerations (accel.c:8-23): box wrap, PM when do_pm,
 * PP always.  One sort serves both. */
int cuco_compute_accelerations(cuco_engine *e, int32_t precision,
                               int32_t do_pm);

/* ---------------------------------------------------------------- */
/* Tuning (no reference counterpart: the reference hard-codes cell = */
/* mesh cell and 256-thread launches, gravlist.cu:22-23)             */
/* ---------------------------------------------------------------- */

typedef struct cuco_tuning {
    int32_t pp_block_shift;       /* FP32 tiles stay inside Morton blocks of
                                     2^shift cells per dimension (2..4) */
    int32_t pp_workgroups_per_cu; /* persistent FP32 pair-kernel workgroups
                                     per compute unit (1..32) */
    int32_t pm_overlap;           /* 1: the PM chain runs on a second stream,
                                     concurrently with the pair force */
    int32_t pp_symmetric;         /* 1 (the default): FP32 pair force
                                     evaluates each pair once for both
                                     particles (full evaluations; active sets
                                     of the integer timeline visit every
                                     ordered pair); 0: every ordered pair */
    int32_t pp_cell_level;        /* candidate enumeration of the FP32
                                     pair kernel: 0 (the default) mesh
                                     cells for tiles narrower than 4 cells,
                                     2x2x2-cell blocks for wider ones; 1 mesh
                                     cells for every tile; 2 two-cell blocks
                                     for every tile (even meshes) */
    int32_t fft_split;            /* FP32 register FFTs of power-of-two
                                     meshes: 0 (the default) 4/8/16/32
                                     threads per line of length <= 32/128/
                                     512/1024; 1 half as many, each holding
                                     twice the values (meshes 96, 160, 192:
                                     one shape either way) */
    int32_t pp_wide_tiles;        /* each-pair-once FP32 kernel: 0 (the
                                     default) tiles of up to 64 targets, one
                                     per lane; 1 tiles of up to 128, two per
                                     lane (each staged particle is read once
                                     for both) */
} cuco_tuning;

int cuco_get_tuning(cuco_engine *e, cuco_tuning *t);
int cuco_set_tuning(cuco_engine *e, const cuco_tuning *t);
/* Times, on the handle's current particles: the FP32 pair force for every
 * Morton block size, workgroup count and pair variant, then its candidate
 * enumeration level and tile width; the FP32 PM chain with each FFT line split; then whole
 * FP32 evaluations with and without PM/PP overlap.  Keeps the fastest and
 * returns it; *best_ms (optional) receives the best pair-kernel time.
 * Every setting gives the same forces up to FP32 rounding.  Leaves the
 * accelerations of the last trial. */
int cuco_autotune(cuco_engine *e, cuco_tuning *chosen, double *best_ms);

/* ---------------------------------------------------------------- */
/* Integrator (device resident)                                      */
/* ---------------------------------------------------------------- */

/* move_particles + do_box_wrapping (predict.c:10-41): x += v*drift; wrap. */
int cuco_drift(cuco_engine *e, double drift_factor);
/* Kicks of advance_and_find_timesteps (timestep.c:51-59, 86-92):
 * v += a_pp*kick_pp; v += a_pm*kick_pm, for all particles. */
int cuco_kick(cuco_engine *e, double kick_pp, double kick_pm);
/* nsteps global KDK steps; factors is nsteps*3 (k1, drift, k2) per step:
 *   v += a*k1; x += v*drift; wrap; forces; v += a*k2
 * (a = a_pp then a_pm).  Forces must be current on entry.  With
 * use_graph != 0 one step is captured once as a hipGraph and replayed. */
int cuco_kdk_steps(cuco_engine *e, int32_t precision, int32_t nsteps,
                   const double *factors, int32_t use_graph);

/* ---------------------------------------------------------------- */
/* Drift / kick factors (host)                                       */
/* ---------------------------------------------------------------- */

#define CUCO_DRIFT_TABLE_LENGTH 1000 /* allvars.h:29 */

/* init_drift_table (driftfac.c:14-42): drift[i] = int da/(H a^3) and
 * kick[i] = int da/(H a^2) from a_begin to
 * exp(ln a_begin + (ln a_max - ln a_begin)/1000*(i+1)),
 * H = hubble*sqrt(om/a^3 + (1-om-ol)/a^2 + ol).  Double tables. */
int cuco_init_drift_table(double a_begin, double a_max, double omega0,
                          double omega_lambda, double hubble, double *drift,
                          double *kick);
/* init_drift_table (driftfac.c:14-42) built on HIP device `device`, the
 * way cuco_timeline_init builds the engine's tables: each interval of the
 * ln a grid integrated with the 16-point Gauss-Legendre rule, the pieces
 * summed in order; the tables are copied into the caller's host arrays
 * (CUCO_DRIFT_TABLE_LENGTH doubles each). */
int cuco_device_drift_table(int32_t device, double a_begin, double a_max,
                            double omega0, double omega_lambda, double hubble,
                            double *drift, double *kick);
/* get_drift_factor / get_gravkick_factor (driftfac.c:66-126) between two
 * values of ln a (the reference takes integer times on the ln a timeline). */
double cuco_table_factor(const double *table, double a_begin, double a_max,
                         double loga0, double loga1);
/* Factors of nsteps global KDK steps uniform in ln a from a_start to a_end:
 * per step (kick to the midpoint, drift over the step, kick from the
 * midpoint), the kicks of timestep.c:48-59 for a global step. */
int cuco_kdk_factors(const double *drift, const double *kick, double a_begin,
                     double a_max, double a_start, double a_end, int32_t nsteps,
                     double *factors);

/* ---------------------------------------------------------------- */
/* Compatibility shims on PARTICLE-layout arrays                     */
/* ---------------------------------------------------------------- */

/* The two shims take n in [1, capacity] (an ensemble: exactly its
 * capacity) and upload positions and masses only: when the handle already
 * holds n particles, their device velocities and integer-timeline state are
 * kept (as the reference's copyPosToDevice leaves velocities on the host);
 * otherwise the device velocities start at zero. */

/* pmforce_periodic() on a caller's P[] (pm_periodic.c:67-314): uploads
 * Pos/Mass, runs the PM force, writes P[i].GravPM. */
int cuco_pmforce_periodic(cuco_engine *e, cuco_particle *P, int64_t n,
                          int32_t precision);
/* gravity_linklist_cuda() on a caller's P[] (gravlist.cpp:148-193):
 * uploads Pos/Mass, runs the PP force, writes P[i].GravAccel. */
int cuco_gravity_linklist(cuco_engine *e, cuco_particle *P, int64_t n,
                          int32_t precision);

/* ---------------------------------------------------------------- */
/* Integer timeline: individual power-of-two timesteps               */
/* ---------------------------------------------------------------- */

#define CUCO_TIMEBASE (1 << 28) /* allvars.h:11 */

/* Run parameters of the reference's integrator (begrun.c:33-56). */
typedef struct cuco_timeline {
    double time_begin;              /* All.TimeBegin (scale factor) */
    double time_max;                /* All.TimeMax */
    double omega0;                  /* All.Omega0 */
    double omega_lambda;            /* All.OmegaLambda */
    double hubble;                  /* All.Hubble (internal units) */
    double err_tol_int_accuracy;    /* All.ErrTolIntAccuracy */
    double max_size_timestep;       /* All.MaxSizeTimestep (in ln a) */
    double min_size_timestep;       /* All.MinSizeTimestep */
    double softening_halo;          /* All.SofteningHalo (comoving) */
    double softening_halo_max_phys; /* All.SofteningHaloMaxPhys */
} cuco_timeline;

/* Integer-timeline state after a cuco_timeline_run call. */
typedef struct cuco_timeline_state {
    int32_t ti_current;       /* All.Ti_Current */
    int32_t pm_ti_begstep;    /* All.PM_Ti_begstep */
    int32_t pm_ti_endstep;    /* All.PM_Ti_endstep */
    int32_t steps;            /* All.NumCurrentTiStep */
    double time;              /* TimeBegin * exp(Ti_Current * Timebase_interval) */
    double timebase_interval; /* All.Timebase_interval */
    int64_t active;           /* particles whose step ended in the last step */
    int32_t finished;         /* Ti_Current reached TIMEBASE or Time > TimeMax */
    int32_t at_stop;          /* the call returned at ti_stop */
} cuco_timeline_state;

/* init (init.c:20-48) and init_drift_table (driftfac.c:14-42): Time =
 * TimeBegin, Ti_Current = 0, every Ti_begstep = Ti_endstep = 0 and
 * PM_Ti_begstep = PM_Ti_endstep = 0; builds the drift and kick tables on
 * the device (as cuco_device_drift_table).  Needs uploaded particles
 * (internal velocities). */
int cuco_timeline_init(cuco_engine *e, const cuco_timeline *t);
/* Up to max_steps passes of run()'s body (run.c:9-25): find the next sync
 * point and drift every particle to it (run.c:29-64); set the softening for
 * the new time (gravlist.cpp:28-45); compute_accelerations (accel.c:8-23:
 * PM on PM steps; the short-range force for the particles whose step ends
 * now -- the reference computes it for every particle but kicks only
 * these); advance_and_find_timesteps (timestep.c:11-94: new power-of-two
 * steps from get_timestep, the short-range kick, the PM kick on PM steps).
 * When the next sync point reaches ti_stop > Ti_Current (the next output
 * time, cuco_find_next_output_time) the particles are drifted to ti_stop and
 * the call returns with at_stop = 1, where the reference calls
 * savepositions (run.c:47-56); the next call continues from there.
 * ti_stop < 0: no stop.  A step that leaves the integer timeline
 * (timestep.c:126-133) returns CUCO_ERR_STATE. */
int cuco_timeline_run(cuco_engine *e, int32_t precision, int32_t max_steps,
                      int32_t ti_stop, cuco_timeline_state *state);
/* Ti_begstep / Ti_endstep of every particle (original order); either may
 * be NULL. */
int cuco_get_timesteps(cuco_engine *e, int32_t *ti_begstep, int32_t *ti_endstep);

/* ---------------------------------------------------------------- */
/* Snapshots and the output schedule (host side)                     */
/* ---------------------------------------------------------------- */

/* savepositions (io.c:9-40): writes HEADER then header->NumPart PARTICLE
 * records, native byte order. */
int cuco_write_snapshot(const char *path, const cuco_header *header,
                        const cuco_particle *P);
/* read_ic (read_ic.c:9-45): reads the HEADER and, when P is not NULL, the
 * NumPart records into P (capacity records).  A file whose size does not
 * match its HEADER is rejected. */
int cuco_read_snapshot(const char *path, cuco_header *header, cuco_particle *P,
                       int64_t capacity);
/* read_outputlist (begrun.c:80-102): up to max_times scale factors (read
 * as float, as the reference does) into times; *count receives how many. */
int cuco_read_output_list(const char *path, double *times, int32_t max_times,
                          int32_t *count);
/* find_next_outputtime (run.c:75-104): smallest integer time of a listed
 * scale factor in [time_begin, time_max] after ti_curr, else 2*TIMEBASE. */
int32_t cuco_find_next_output_time(const double *times, int32_t count,
                                   double time_begin, double time_max,
                                   double timebase_interval, int32_t ti_curr);
/* The handle's particles as PARTICLE records in original order (P[] of
 * savepositions): Pos, Mass, internal Vel, the last GravAccel (PP) and
 * GravPM, and the integer timeline's Ti_endstep/Ti_begstep (0 before
 * cuco_timeline_init).  After cuco_timeline_run, GravAccel is current for
 * the particles active on the last step only (the ones it kicked); an
 * inactive particle's value is from its own last active step or from the
 * last whole evaluation, whichever came later (the reference recomputes
 * every particle's GravAccel each step, accel.c:22, but kicks only the
 * active ones, so the dynamics are the same). */
int cuco_download_records(cuco_engine *e, cuco_particle *P);
/* Uploads Pos, Vel (internal velocities, as P[] holds them after init.c:
 * 27-36) and Mass of n PARTICLE records; as cuco_upload_particles. */
int cuco_upload_records(cuco_engine *e, const cuco_particle *P, int64_t n);
/* The particle part of init() (init.c:27-36), host only: Vel *= a^(3/2)
 * with a = time_begin (IC files hold peculiar velocities; the run holds
 * a^2 dx/dt), GravAccel, GravPM, Ti_endstep and Ti_begstep set to 0. */
int cuco_init_records(cuco_particle *P, int64_t n, double time_begin);

/* ---------------------------------------------------------------- */
/* Checkpoints (restart)                                             */
/* ---------------------------------------------------------------- */

/* The reference has no restart: savepositions (io.c:9-40) keeps neither
 * the run's internal velocities nor its integer timeline, and read_ic
 * treats a snapshot as new initial conditions (init.c:27-36 rescales Vel
 * and zeroes every Ti).  A checkpoint holds the whole state a handle needs
 * to continue: cuco_checkpoint_read into a handle created with the same
 * parameters (and, for bit-identical FP32 results, the same tuning), then
 * cuco_timeline_run continues exactly as the handle that wrote it would.
 *
 * File layout, native byte order: this header; nboxes cuco_params (the
 * boxes' parameters, soft_h the softening in force); nboxes int64 particle
 * counts; then per particle in original order n double4 (x, y, z, mass),
 * n double4 (internal velocity, 0); when timeline = 1, n int32 Ti_begstep
 * and n int32 Ti_endstep; when pp_precision >= 0, n accelerations of the
 * short-range force (double4 for CUCO_FP64, float4 for CUCO_FP32, the
 * fourth component unused); then likewise the PM force. */
typedef struct cuco_checkpoint_header {
    char magic[8];              /* "CUCOCKPT" */
    int32_t version;            /* 1 */
    int32_t nboxes;             /* 1, or the ensemble's box count */
    int64_t n;                  /* particles */
    int32_t timeline;           /* 1: params, state and the Ti arrays are held */
    int32_t pp_precision;       /* precision of the stored PP accelerations, -1: none */
    int32_t pm_precision;       /* likewise for the PM accelerations */
    int32_t reserved;           /* 0 */
    cuco_timeline params;       /* integrator parameters (timeline = 1) */
    cuco_timeline_state state;  /* integer-timeline state (timeline = 1) */
} cuco_checkpoint_header;

/* Size in bytes of the checkpoint file a header describes (-1 when its
 * fields are out of range). */
int64_t cuco_checkpoint_bytes(const cuco_checkpoint_header *header);
/* Reads and checks a checkpoint's header (magic, version, field ranges,
 * file size), host only. */
int cuco_checkpoint_info(const char *path, cuco_checkpoint_header *header);
/* Writes the handle's particles, velocities, last accelerations and, after
 * cuco_timeline_init, its integer timeline to path. */
int cuco_checkpoint_write(cuco_engine *e, const char *path);
/* Restores a checkpoint into a handle: the box count, per-box particle
 * counts and parameters (except the softening, which is restored) must
 * match the handle's; otherwise CUCO_ERR_INVALID_ARG and the handle is
 * unchanged.  Replaces the particles; the integer timeline is restored
 * when the file holds one, else the handle has none. */
int cuco_checkpoint_read(cuco_engine *e, const char *path);

/* ---------------------------------------------------------------- */
/* Device-pointer API (mirrors proto.cuh:26-58)                      */
/* ---------------------------------------------------------------- */

/* allocateArray / freeArray / copyArray{To,From}Device (gravlist.cu:33-54). */
int cuco_device_alloc(void **dptr, size_t bytes);
int cuco_device_free(void *dptr);
int cuco_copy_to_device(void *dst, const void *src, size_t bytes);
int cuco_copy_from_device(void *dst, const void *src, size_t bytes);
/* Device-to-device copy with the engine's own streaming kernel (16-byte
 * vector loads/stores; bytes must be a multiple of 16).  Its counters give
 * the HBM bytes of a copy of known size. */
int cuco_copy_device_to_device(void *dst, const void *src, size_t bytes);
/* RadixSort::sort(keys, values, n, keyBits) (radixsort.h:170-187): stable
 * key/value sort of n uint32 pairs on the device, ordering by bits
 * [0, key_bits) of the key.  values may be NULL (keys only).  Scratch is
 * allocated per call. */
int cuco_radix_sort(uint32_t *d_keys, uint32_t *d_values, uint32_t n,
                    uint32_t key_bits);
/* cuco_radix_sort without any allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and result; the scratch is a static device array of the
 * library (33 MiB per device, held until the process ends), so n is at most
 * CUCO_RADIX_SORT_NOALLOC_MAX (larger n: CUCO_ERR_INVALID_ARG).  All calls
 * on one device share that array: calls that overlap on one device race.
 * Synchronous (null stream). */
#define CUCO_RADIX_SORT_NOALLOC_MAX 4194304u
int cuco_radix_sort_noalloc(uint32_t *d_keys, uint32_t *d_values, uint32_t n,
                            uint32_t key_bits);

/* Unnormalized 3D FFTs on device arrays, replacing the cuFFT plans of
 * pm_periodic.c:34-35 and their executions cufftExecR2C (:140) and
 * cufftExecC2R (:201).  Real arrays are M^3 (x-major), complex arrays
 * M*M*(M/2+1) interleaved (re, im); precision selects double or float
 * elements.  c2r overwrites its input; c2r(r2c(x)) = M^3 x. */
int cuco_fft3d_r2c(int32_t M, int32_t precision, const void *d_in, void *d_out);
int cuco_fft3d_c2r(int32_t M, int32_t precision, void *d_in, void *d_out);

/* ---------------------------------------------------------------- */
/* Inspection (parity tests, diagnostics)                            */
/* ---------------------------------------------------------------- */

/* Sorted keys and the sorting permutation (sorted position -> original
 * index) of the last evaluation; n entries each. */
int cuco_get_sorted(cuco_engine *e, uint32_t *keys, uint32_t *index);
/* Cell table of the last evaluation in the reference's format
 * (gravlist_kernel.cu:136-155): ncells = PMGRID^3 entries each, 0xffffffff
 * for empty cells in both arrays.  Only for the FP64 (reference-order) key. */
int cuco_get_cell_table(cuco_engine *e, uint32_t *cell_start,
                        uint32_t *cell_end);
/* The 1000-entry short-range table uploaded to the device (gravlist.cpp:56-59). */
int cuco_get_shortrange_table(cuco_engine *e, float *table);
/* Mesh of the last PM evaluation: which = 0 density (after deposit),
 * 1 potential (after c2r).  PMGRID^3 doubles, x-major (an ensemble: every
 * box's mesh in box order). */
int cuco_get_pm_grid(cuco_engine *e, int32_t which, double *out);
/* Exact count of ordered pairs i != j with r^2 <= rcut^2 (FP64 test) and
 * of candidate pairs the FP32 kernel evaluates, for the current positions
 * (with pp_symmetric a pair evaluated for both particles counts once). */
int cuco_count_pairs(cuco_engine *e, uint64_t *pairs_in,
                     uint64_t *fp32_candidates);

/* ---------------------------------------------------------------- */
/* Timing (HIP events on the handle's stream)                        */
/* ---------------------------------------------------------------- */

enum {
    CUCO_STAGE_KEYS = 0,
    CUCO_STAGE_SORT,
    CUCO_STAGE_CELLS,   /* cell table + reorder */
    CUCO_STAGE_PP,
    CUCO_STAGE_DEPOSIT,
    CUCO_STAGE_FFT_FWD, /* z r2c + y pass */
    CUCO_STAGE_GREEN,   /* fused x forward, Green's function, x inverse */
    CUCO_STAGE_FFT_INV, /* y pass + z c2r */
    CUCO_STAGE_INTERP,  /* gradient + interpolation */
    CUCO_STAGE_COUNT
};
/* Enable (1) or disable (0) per-stage event timing. */
int cuco_set_timing(cuco_engine *e, int32_t enable);
/* Milliseconds per stage summed over evaluations since the last reset,
 * and the number of evaluations; reset != 0 clears the sums afterwards. */
int cuco_get_stage_times(cuco_engine *e, double *ms, int32_t nstages,
                         int32_t *evaluations, int32_t reset);
const char *cuco_stage_name(int32_t stage);
/* Blocks until all work queued on the handle's stream is done. */
int cuco_synchronize(cuco_engine *e);

#ifdef __cplusplus
}
#endif
#endif
