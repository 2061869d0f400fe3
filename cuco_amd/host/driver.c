/* A C host program in the shape of cuco's own (main.c, begrun.c, run.c,
 * accel.c), linked against libcuco_engine: the reference's call sites stay,
 * and the bodies of pmforce_periodic() and gravity_linklist_cuda() forward
 * to the engine.  It reads an IC file (HEADER + PARTICLE[], read_ic.c),
 * then either
 *
 *   driver forces <ic> <out> <pmgrid> <fp64|fp32> <softening>
 *       one compute_accelerations() on the IC with the comoving Plummer
 *       softening <softening> (h = 2.8 x that); writes P[] (with GravPM and
 *       GravAccel) to <out> in the snapshot format (io.c), or
 *
 *   driver run <ic> <outbase> <pmgrid> <fp64|fp32> <outputlist> <a_begin>
 *              <a_max> <softening> <softening_max_phys> <err_tol>
 *              <max_dt> [max_steps]
 *       the reference's run(): individual timesteps on the integer timeline,
 *       a snapshot <outbase>_<k>.cuco at every listed output time
 *       (find_next_sync_point_and_drift, run.c:47-56) and at the end, and
 *       beside each listed one a checkpoint <outbase>_<k>.ckpt, or
 *
 *   driver resume <ic> <outbase> <pmgrid> <fp64|fp32> <outputlist> <ckpt>
 *                 <k> [max_steps]
 *       continues a run from its checkpoint (which the reference cannot:
 *       it has no restart), numbering the next snapshot <k>; the snapshots
 *       are those the uninterrupted run writes.  <ic> gives the box
 *       (HEADER) and the particle count.
 *
 * Units and cosmology are the reference's (begrun.c:33-70): kpc/h,
 * 1e10 Msun/h, km/s, Omega0 = 0.3, OmegaLambda = 0.7, h = 0.7.  The IC
 * holds peculiar velocities: run mode applies init()'s Vel *= a^(3/2) at
 * a_begin (init.c:27-36) before the first step, as the reference does;
 * forces mode does not use velocities.  Exit status 0 on success; every
 * engine error is printed. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "cuco_engine.h"

#define ASMTH 1.25 /* allvars.h:22-23 */
#define RCUT 4.5

static cuco_particle *P;
static int NumPart;
static cuco_header header;
static cuco_engine *engine;
static int precision;

static struct {
    int pmgrid;
    double BoxSize, G, Hubble, Asmth, Rcut, Time;
    double SofteningHalo, SofteningHaloMaxPhys, ForceSoftening;
} All;

static int check(int rc, const char *what)
{
    if (rc != CUCO_OK) {
        fprintf(stderr, "%s failed: %s (%s)\n", what, cuco_status_string(rc),
                cuco_last_error());
        exit(2);
    }
    return rc;
}

static void set_units(void) /* begrun.c:58-70 */
{
    const double UnitLength_in_m = 3.085678e19, UnitMass_in_kg = 1.98892e40;
    const double UnitVelocity_in_m_per_s = 1e3;
    const double UnitTime_in_s = UnitLength_in_m / UnitVelocity_in_m_per_s;
    All.G = 6.673e-11 / pow(UnitLength_in_m, 3) * UnitMass_in_kg * pow(UnitTime_in_s, 2);
    All.Hubble = 3.2407789e-18 * UnitTime_in_s;
}

static void read_ic(const char *fname) /* read_ic.c:9-45 */
{
    check(cuco_read_snapshot(fname, &header, NULL, 0), "read_ic header");
    NumPart = header.NumPart;
    P = (cuco_particle *)malloc(sizeof(cuco_particle) * (size_t)(NumPart > 0 ? NumPart : 1));
    if (!P) {
        fprintf(stderr, "out of memory\n");
        exit(1);
    }
    check(cuco_read_snapshot(fname, &header, P, NumPart), "read_ic");
    All.BoxSize = header.BoxSize;
}

static void set_softenings(void) /* gravlist.cpp:28-45 */
{
    double soft = All.SofteningHalo * All.Time > All.SofteningHaloMaxPhys
                      ? All.SofteningHaloMaxPhys / All.Time
                      : All.SofteningHalo;
    All.ForceSoftening = 2.8 * soft;
}

void pm_init_periodic(void) /* pm_periodic.c:25-43 */
{
    All.Asmth = ASMTH * All.BoxSize / All.pmgrid;
    All.Rcut = RCUT * All.Asmth;
}

void linklist_init(void) /* gravlist.cpp:48-115 */
{
    set_softenings();
    cuco_params p = {All.pmgrid, 0, All.BoxSize, All.G, All.Asmth, All.Rcut,
                     All.ForceSoftening};
    check(cuco_create(&p, NumPart, 0, &engine), "linklist_init");
}

void pmforce_periodic(void) /* longrange.c:21 */
{
    check(cuco_pmforce_periodic(engine, P, NumPart, precision), "pmforce_periodic");
}

void gravity_linklist_cuda(void) /* accel.c:22 */
{
    set_softenings(); /* gravlist.cpp:151 */
    check(cuco_set_softening(engine, All.ForceSoftening), "set_softenings");
    check(cuco_gravity_linklist(engine, P, NumPart, precision), "gravity_linklist_cuda");
}

void compute_accelerations(void) /* accel.c:8-23 (PM step) */
{
    pmforce_periodic();
    gravity_linklist_cuda();
}

static void savepositions(const char *fname) /* io.c:9-40 */
{
    cuco_header h = header;
    h.NumPart = NumPart;
    h.Time = (float)All.Time;
    h.Redshift = (float)(1 / All.Time - 1);
    check(cuco_write_snapshot(fname, &h, P), "savepositions");
}

/* The output loop of run() (run.c:9-26) from the handle's current state:
 * snapshots <outbase>_<count>.cuco (and checkpoints .ckpt) at the listed
 * times, a last snapshot at the end. */
static int advance(const cuco_timeline *t, const double *times, int32_t ntimes,
                   int32_t ti_next, int count, int steps, int max_steps, const char *outbase)
{
    const double interval = (log(t->time_max) - log(t->time_begin)) / CUCO_TIMEBASE;
    char fname[512];
    for (;;) {
        cuco_timeline_state s;
        const int budget = max_steps - steps;
        if (budget <= 0) break;
        check(cuco_timeline_run(engine, precision, budget,
                                ti_next < 2 * CUCO_TIMEBASE ? ti_next : -1, &s),
              "run");
        steps = s.steps;
        All.Time = s.time;
        if (s.at_stop) {
            check(cuco_download_records(engine, P), "download");
            snprintf(fname, sizeof(fname), "%s_%d.cuco", outbase, count);
            savepositions(fname);
            printf("output %s at a = %g after %d steps\n", fname, s.time, s.steps);
            snprintf(fname, sizeof(fname), "%s_%d.ckpt", outbase, count++);
            check(cuco_checkpoint_write(engine, fname), "checkpoint");
            ti_next = cuco_find_next_output_time(times, ntimes, t->time_begin,
                                                 t->time_max, interval, ti_next + 1);
            continue;
        }
        if (s.finished) break;
    }
    check(cuco_download_records(engine, P), "download");
    snprintf(fname, sizeof(fname), "%s_%d.cuco", outbase, count++);
    savepositions(fname); /* run.c:27 */
    printf("final %s at a = %g after %d steps\n", fname, All.Time, steps);
    return 0;
}

static int run(int argc, char **argv)
{
    double times[500];
    int32_t ntimes = 0;
    check(cuco_read_output_list(argv[6], times, 500, &ntimes), "read_outputlist");
    cuco_timeline t;
    t.time_begin = atof(argv[7]);
    t.time_max = atof(argv[8]);
    t.omega0 = 0.3;
    t.omega_lambda = 0.7;
    t.hubble = All.Hubble;
    t.softening_halo = atof(argv[9]);
    t.softening_halo_max_phys = atof(argv[10]);
    t.err_tol_int_accuracy = atof(argv[11]);
    t.max_size_timestep = atof(argv[12]);
    t.min_size_timestep = 0.0;
    const int max_steps = argc > 13 ? atoi(argv[13]) : 1 << 30;
    All.SofteningHalo = t.softening_halo;
    All.SofteningHaloMaxPhys = t.softening_halo_max_phys;
    All.Time = t.time_begin;
    check(cuco_init_records(P, NumPart, All.Time), "init"); /* init.c:27-36 */
    linklist_init();
    check(cuco_upload_records(engine, P, NumPart), "upload");
    check(cuco_timeline_init(engine, &t), "init");
    const double interval = (log(t.time_max) - log(t.time_begin)) / CUCO_TIMEBASE;
    const int32_t ti_next =
        cuco_find_next_output_time(times, ntimes, t.time_begin, t.time_max, interval, 0);
    return advance(&t, times, ntimes, ti_next, 0, 0, max_steps, argv[3]);
}

static int resume(int argc, char **argv)
{
    double times[500];
    int32_t ntimes = 0;
    check(cuco_read_output_list(argv[6], times, 500, &ntimes), "read_outputlist");
    cuco_checkpoint_header ck;
    check(cuco_checkpoint_info(argv[7], &ck), "checkpoint_info");
    if (!ck.timeline || ck.n != NumPart) {
        fprintf(stderr, "%s: no integer timeline, or not %d particles\n", argv[7], NumPart);
        return 1;
    }
    const cuco_timeline t = ck.params;
    const int max_steps = argc > 9 ? atoi(argv[9]) : 1 << 30;
    All.SofteningHalo = t.softening_halo;
    All.SofteningHaloMaxPhys = t.softening_halo_max_phys;
    All.Time = ck.state.time;
    linklist_init();
    check(cuco_checkpoint_read(engine, argv[7]), "checkpoint_read");
    const double interval = (log(t.time_max) - log(t.time_begin)) / CUCO_TIMEBASE;
    const int32_t ti_next = cuco_find_next_output_time(times, ntimes, t.time_begin, t.time_max,
                                                       interval, ck.state.ti_current + 1);
    return advance(&t, times, ntimes, ti_next, atoi(argv[8]), ck.state.steps, max_steps,
                   argv[3]);
}

int main(int argc, char **argv)
{
    if (argc < 7 ||
        (strcmp(argv[1], "forces") && strcmp(argv[1], "run") && strcmp(argv[1], "resume")) ||
        (!strcmp(argv[1], "run") && argc < 13) || (!strcmp(argv[1], "resume") && argc < 9)) {
        fprintf(stderr, "usage: driver forces <ic> <out> <pmgrid> <fp64|fp32> <softening>\n"
                        "       driver run <ic> <outbase> <pmgrid> <fp64|fp32> <outputlist>"
                        " <a_begin> <a_max> <soft> <soft_max_phys> <err_tol> <max_dt>"
                        " [max_steps]\n"
                        "       driver resume <ic> <outbase> <pmgrid> <fp64|fp32> <outputlist>"
                        " <ckpt> <k> [max_steps]\n");
        return 1;
    }
    set_units();
    read_ic(argv[2]);
    All.pmgrid = atoi(argv[4]);
    precision = strcmp(argv[5], "fp32") ? CUCO_FP64 : CUCO_FP32;
    pm_init_periodic();
    if (!strcmp(argv[1], "run")) return run(argc, argv);
    if (!strcmp(argv[1], "resume")) return resume(argc, argv);
    All.Time = header.Time;
    All.SofteningHalo = atof(argv[6]);
    All.SofteningHaloMaxPhys = 1e30;
    linklist_init();
    compute_accelerations();
    savepositions(argv[3]);
    cuco_destroy(engine);
    free(P);
    return 0;
}
