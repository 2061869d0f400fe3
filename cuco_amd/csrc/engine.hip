// C ABI of the force engine: handle life cycle, device-resident particle
// state, force orchestration (sort once, PM and PP from the same order),
// integrator, hipGraph capture of a KDK step, timing and inspection.
//
// Reference call structure kept here: compute_accelerations (accel.c:8-23)
// runs PM on PM steps and PP always; gravity_linklist_cuda
// (gravlist.cpp:148-193) is key -> sort -> cell table -> force;
// pmforce_periodic (pm_periodic.c:67-314) is deposit -> r2c -> Green -> c2r
// -> gradient/interpolation.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <string>
#include <vector>

#include "internal.h"

namespace cuco {

static thread_local std::string g_last_error;

void set_error(const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

}  // namespace cuco

using namespace cuco;

namespace {

struct StageEvents {
    hipEvent_t start[CUCO_STAGE_COUNT];
    hipEvent_t stop[CUCO_STAGE_COUNT];
    bool used[CUCO_STAGE_COUNT];
};

int ceil_log2(uint64_t v)
{
    int b = 0;
    while ((1ull << b) < v) ++b;
    return b;
}

template <typename T>
void free_ptr(T *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

template <typename T>
int dev_alloc(T *&p, size_t bytes)
{
    void *q = nullptr;
    hipError_t err = hipMalloc(&q, bytes ? bytes : 16);
    if (err != hipSuccess) {
        set_error("hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(err));
        return CUCO_ERR_OUT_OF_MEMORY;
    }
    p = static_cast<T *>(q);
    return CUCO_OK;
}

}  // namespace

struct cuco_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    cuco_params prm{};
    int M = 0;
    uint32_t capacity = 0;
    uint32_t n = 0;
    bool have_particles = false;
    double uniform_mass = 0;  // > 0 when every particle has this mass

    double4 *pos = nullptr, *vel = nullptr;
    double4 *acc_pp64 = nullptr, *acc_pm64 = nullptr;
    float4 *acc_pp32 = nullptr, *acc_pm32 = nullptr;
    int pp_prec = -1, pm_prec = -1;

    uint32_t *keys = nullptr, *vals = nullptr;
    SortScratch sc;
    uint32_t *skeys = nullptr, *perm = nullptr;
    uint32_t *off = nullptr;
    int sorted_mode = -1;
    double4 *spos64 = nullptr;
    float4 *spos32 = nullptr;
    GridInfo gx{}, gm{};

    float *table = nullptr;    // 1000 entries, gravlist.cpp:56-59
    float *table32 = nullptr;  // 1024 entries for the FP32 kernel
    float htable[kNTab];
    PairParams pp{};
    TileInfo tiles;
    int pp32_workgroups = 0;
    int pp32_wg_per_cu = 8;
    int ncu = 256;
    int min_block_shift = 2;

    void *rho[2] = {nullptr, nullptr};
    void *phi[2] = {nullptr, nullptr};
    void *cgrid[2] = {nullptr, nullptr};
    void *dep_scratch[2] = {nullptr, nullptr};
    PmTables tb[2];
    double pm_fac = 0;

    unsigned long long *dcount = nullptr;
    double *dfactors = nullptr;
    uint32_t *dstep = nullptr;
    int factors_cap = 0;
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    int graph_prec = -1;
    bool capturing = false;

    bool timing = false;
    StageEvents *cur_events = nullptr;
    std::vector<StageEvents *> pending, spare;
    double stage_ms[CUCO_STAGE_COUNT] = {};
    int evals = 0;
};

namespace {

int set_device(cuco_engine *e)
{
    CUCO_HIP(hipSetDevice(e->device));
    return CUCO_OK;
}

#define CHECK_ENGINE(e)                                                      \
    do {                                                                     \
        if (!(e)) {                                                          \
            set_error("null engine handle");                                 \
            return CUCO_ERR_INVALID_ARG;                                     \
        }                                                                    \
        int rc_ = set_device(e);                                             \
        if (rc_) return rc_;                                                 \
    } while (0)

#define TRY(call)              \
    do {                       \
        int rc_ = (call);      \
        if (rc_) return rc_;   \
    } while (0)

// ------------------------------------------------------------ timing
int begin_eval(cuco_engine *e)
{
    if (!e->timing || e->capturing || e->cur_events) return CUCO_OK;
    StageEvents *ev;
    if (!e->spare.empty()) {
        ev = e->spare.back();
        e->spare.pop_back();
    } else {
        ev = new StageEvents();
        for (int s = 0; s < CUCO_STAGE_COUNT; ++s) {
            CUCO_HIP(hipEventCreate(&ev->start[s]));
            CUCO_HIP(hipEventCreate(&ev->stop[s]));
        }
    }
    for (int s = 0; s < CUCO_STAGE_COUNT; ++s) ev->used[s] = false;
    e->cur_events = ev;
    return CUCO_OK;
}

void end_eval(cuco_engine *e)
{
    if (e->cur_events) {
        e->pending.push_back(e->cur_events);
        e->cur_events = nullptr;
    }
}

int stage_start(cuco_engine *e, int s)
{
    if (e->cur_events) {
        CUCO_HIP(hipEventRecord(e->cur_events->start[s], e->stream));
        e->cur_events->used[s] = true;
    }
    return CUCO_OK;
}

int stage_stop(cuco_engine *e, int s)
{
    if (e->cur_events) CUCO_HIP(hipEventRecord(e->cur_events->stop[s], e->stream));
    return CUCO_OK;
}

int harvest(cuco_engine *e)
{
    for (StageEvents *ev : e->pending) {
        for (int s = 0; s < CUCO_STAGE_COUNT; ++s) {
            if (!ev->used[s]) continue;
            CUCO_HIP(hipEventSynchronize(ev->stop[s]));
            float ms = 0;
            CUCO_HIP(hipEventElapsedTime(&ms, ev->start[s], ev->stop[s]));
            e->stage_ms[s] += ms;
        }
        e->evals++;
        e->spare.push_back(ev);
    }
    e->pending.clear();
    return CUCO_OK;
}

// ------------------------------------------------------------ parameters
void make_table(float *t)
{
    // gravlist.cpp:56-59
    for (int i = 0; i < kNTab; ++i) {
        double u = 3.0 / kNTab * (i + 0.5);
        t[i] = (float)(erfc(u) + 2.0 * u / sqrt(M_PI) * exp(-u * u));
    }
}

int upload_tables(cuco_engine *e)
{
    make_table(e->htable);
    float t32[kTabPadded];
    const double edge = e->pp.asmthfac * e->pp.rcut;
    const bool aligned = fabs(edge - nearbyint(edge)) < 1e-9 * edge;
    const int kcut = aligned ? (int)nearbyint(edge) : kNTab;
    for (int k = 0; k < kTabPadded; ++k)
        t32[k] = (k < kNTab && k < kcut) ? e->htable[k] : 0.0f;
    CUCO_HIP(hipMemcpy(e->table, e->htable, sizeof(float) * kNTab, hipMemcpyHostToDevice));
    CUCO_HIP(hipMemcpy(e->table32, t32, sizeof(t32), hipMemcpyHostToDevice));
    return CUCO_OK;
}

void set_pair_params(cuco_engine *e)
{
    // gravlist.cpp:36-40, :61-70
    const cuco_params &p = e->prm;
    PairParams &q = e->pp;
    q.L = p.box_size;
    q.boxhalf = p.box_size / 2.0;
    q.rcut = p.rcut;
    q.rcut2 = p.rcut * p.rcut;
    q.asmthfac = 0.5 / p.asmth * (kNTab / 3.0);
    q.h = p.soft_h;
    q.h_inv = 1.0 / p.soft_h;
    q.h3_inv = q.h_inv * q.h_inv * q.h_inv;
    q.G = p.G;
    q.fac = p.pm_grid / p.box_size;
    q.M = p.pm_grid;
}

template <typename T>
int make_pm_tables(cuco_engine *e, PmTables &tb)
{
    const int M = e->M;
    std::vector<Cx<T>> tw(M);
    std::vector<T> ge(M), gs(M);
    const double asmth2 = std::pow(2 * M_PI * e->prm.asmth / e->prm.box_size, 2);
    for (int k = 0; k < M; ++k) {
        const double a = 2.0 * M_PI * k / M;
        tw[k] = {(T)cos(a), (T)-sin(a)};
        const double kk = k > M / 2 ? k - M : k;
        ge[k] = (T)exp(-kk * kk * asmth2);
        double f = 1.0;
        if (kk != 0) {
            f = M_PI * kk / M;
            f = sin(f) / f;
        }
        const double ff = 1.0 / f;
        gs[k] = (T)(ff * ff * ff * ff);
    }
    Cx<T> *dtw = nullptr;
    T *dge = nullptr, *dgs = nullptr;
    TRY(dev_alloc(dtw, sizeof(Cx<T>) * M));
    TRY(dev_alloc(dge, sizeof(T) * M));
    TRY(dev_alloc(dgs, sizeof(T) * M));
    CUCO_HIP(hipMemcpy(dtw, tw.data(), sizeof(Cx<T>) * M, hipMemcpyHostToDevice));
    CUCO_HIP(hipMemcpy(dge, ge.data(), sizeof(T) * M, hipMemcpyHostToDevice));
    CUCO_HIP(hipMemcpy(dgs, gs.data(), sizeof(T) * M, hipMemcpyHostToDevice));
    tb.tw = dtw;
    tb.gexp = dge;
    tb.gsinc = dgs;
    return CUCO_OK;
}

template <typename T>
int ensure_pm(cuco_engine *e)
{
    const int k = sizeof(T) == 8 ? 0 : 1;
    if (e->cgrid[k]) return CUCO_OK;
    const size_t M = e->M;
    TRY(dev_alloc(e->rho[k], sizeof(T) * M * M * M));
    TRY(dev_alloc(e->phi[k], sizeof(T) * M * M * M));
    TRY(dev_alloc(e->cgrid[k], sizeof(Cx<T>) * M * M * (M / 2 + 1)));
    TRY(dev_alloc(e->dep_scratch[k], pm_deposit_scratch_bytes(e->capacity, k == 0)));
    return make_pm_tables<T>(e, e->tb[k]);
}

// ------------------------------------------------------------ sort
int sort_particles(cuco_engine *e, int mode)
{
    if (e->sorted_mode == mode) return CUCO_OK;
    const GridInfo &g = mode == kKeyXMajor ? e->gx : e->gm;
    TRY(stage_start(e, CUCO_STAGE_KEYS));
    TRY(launch_calc_keys(e->pos, e->n, g, e->keys, e->stream));
    TRY(stage_stop(e, CUCO_STAGE_KEYS));
    TRY(stage_start(e, CUCO_STAGE_SORT));
    TRY(radix_sort_pairs(e->keys, e->vals, true, e->n, g.key_bits, e->sc,
                         e->stream, &e->skeys, &e->perm));
    TRY(stage_stop(e, CUCO_STAGE_SORT));
    TRY(stage_start(e, CUCO_STAGE_CELLS));
    TRY(launch_cell_offsets(e->skeys, e->n, g.nkeys, e->off, e->stream));
    if (mode == kKeyXMajor) {
        TRY(launch_reorder64(e->pos, e->perm, e->n, e->spos64, e->stream));
    } else {
        TRY(launch_reorder32(e->pos, e->perm, e->skeys, e->n, g, e->spos32, e->stream));
        TRY(launch_build_tiles(e->off, e->n, g, e->tiles, e->stream));
    }
    TRY(stage_stop(e, CUCO_STAGE_CELLS));
    e->sorted_mode = mode;
    return CUCO_OK;
}

int check_ready(cuco_engine *e)
{
    if (!e->have_particles) {
        set_error("no particles uploaded");
        return CUCO_ERR_STATE;
    }
    return CUCO_OK;
}

int check_prec(int32_t precision)
{
    if (precision != CUCO_FP64 && precision != CUCO_FP32) {
        set_error("unknown precision %d", precision);
        return CUCO_ERR_INVALID_ARG;
    }
    return CUCO_OK;
}

int mode_of(int precision) { return precision == CUCO_FP64 ? kKeyXMajor : kKeyMorton; }

int do_pp(cuco_engine *e, int precision, unsigned long long *counter)
{
    TRY(sort_particles(e, mode_of(precision)));
    TRY(stage_start(e, CUCO_STAGE_PP));
    if (precision == CUCO_FP64) {
        TRY(launch_pp64(e->spos64, e->perm, e->off, e->n, e->pp, e->table,
                        e->acc_pp64, counter, e->stream));
    } else {
        TRY(launch_pp32(e->spos32, e->skeys, e->perm, e->off, e->n, e->gm, e->pp,
                        e->table32, e->tiles, e->acc_pp32, counter,
                        e->pp32_workgroups, e->uniform_mass, e->stream));
    }
    TRY(stage_stop(e, CUCO_STAGE_PP));
    e->pp_prec = precision;
    return CUCO_OK;
}

template <typename T>
int do_pm_t(cuco_engine *e, int precision)
{
    TRY(ensure_pm<T>(e));
    TRY(sort_particles(e, mode_of(precision)));
    const int k = sizeof(T) == 8 ? 0 : 1;
    const GridInfo &g = precision == CUCO_FP64 ? e->gx : e->gm;
    const void *spos = precision == CUCO_FP64 ? (const void *)e->spos64 : (const void *)e->spos32;
    void *acc = precision == CUCO_FP64 ? (void *)e->acc_pm64 : (void *)e->acc_pm32;
    T *rho = static_cast<T *>(e->rho[k]);
    T *phi = static_cast<T *>(e->phi[k]);
    Cx<T> *c = static_cast<Cx<T> *>(e->cgrid[k]);
    TRY(stage_start(e, CUCO_STAGE_DEPOSIT));
    TRY(pm_deposit<T>(spos, e->skeys, e->off, e->n, g, rho, e->dep_scratch[k], e->stream));
    TRY(stage_stop(e, CUCO_STAGE_DEPOSIT));
    TRY(stage_start(e, CUCO_STAGE_FFT_FWD));
    TRY(pm_fft_forward<T>(rho, c, e->M, e->tb[k], e->stream));
    TRY(stage_stop(e, CUCO_STAGE_FFT_FWD));
    TRY(stage_start(e, CUCO_STAGE_GREEN));
    TRY(pm_green_x<T>(c, e->M, e->tb[k], e->stream));
    TRY(stage_stop(e, CUCO_STAGE_GREEN));
    TRY(stage_start(e, CUCO_STAGE_FFT_INV));
    TRY(pm_fft_inverse<T>(c, phi, e->M, e->tb[k], e->stream));
    TRY(stage_stop(e, CUCO_STAGE_FFT_INV));
    TRY(stage_start(e, CUCO_STAGE_INTERP));
    TRY(pm_interp<T>(phi, spos, e->skeys, e->perm, e->n, g, e->pm_fac, acc, e->stream));
    TRY(stage_stop(e, CUCO_STAGE_INTERP));
    e->pm_prec = precision;
    return CUCO_OK;
}

int do_pm(cuco_engine *e, int precision)
{
    return precision == CUCO_FP64 ? do_pm_t<double>(e, precision)
                                  : do_pm_t<float>(e, precision);
}

int do_forces(cuco_engine *e, int precision, int do_pm_flag)
{
    if (do_pm_flag) TRY(do_pm(e, precision));  // accel.c:14-16
    return do_pp(e, precision, nullptr);       // accel.c:22
}

int kick_current(cuco_engine *e, double kpp, double kpm)
{
    if (e->pp_prec < 0 && e->pm_prec < 0) {
        set_error("kick before any force evaluation");
        return CUCO_ERR_STATE;
    }
    const bool pp32 = e->pp_prec == CUCO_FP32, pm32 = e->pm_prec == CUCO_FP32;
    if (e->pp_prec >= 0 && e->pm_prec >= 0 && pp32 != pm32) {
        // Mixed precisions: two kicks, PP first (timestep.c:223-226 then :253-255).
        if (pp32) TRY(launch_kick32(e->vel, e->acc_pp32, nullptr, e->n, kpp, 0, e->stream));
        else TRY(launch_kick64(e->vel, e->acc_pp64, nullptr, e->n, kpp, 0, e->stream));
        if (pm32) TRY(launch_kick32(e->vel, nullptr, e->acc_pm32, e->n, 0, kpm, e->stream));
        else TRY(launch_kick64(e->vel, nullptr, e->acc_pm64, e->n, 0, kpm, e->stream));
        return CUCO_OK;
    }
    if (pp32 || pm32)
        return launch_kick32(e->vel, e->pp_prec >= 0 ? e->acc_pp32 : nullptr,
                             e->pm_prec >= 0 ? e->acc_pm32 : nullptr, e->n, kpp, kpm,
                             e->stream);
    return launch_kick64(e->vel, e->pp_prec >= 0 ? e->acc_pp64 : nullptr,
                         e->pm_prec >= 0 ? e->acc_pm64 : nullptr, e->n, kpp, kpm,
                         e->stream);
}

int destroy_graph(cuco_engine *e)
{
    if (e->graph_exec) (void)hipGraphExecDestroy(e->graph_exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->graph_exec = nullptr;
    e->graph = nullptr;
    e->graph_prec = -1;
    return CUCO_OK;
}

int wrap_host(double *x, double L)
{
    // predict.c:31-40
    int guard = 0;
    while (*x < 0 && guard++ < 64) *x += L;
    while (*x >= L && guard++ < 128) *x -= L;
    if (!(*x >= 0 && *x < L)) {
        set_error("position %g cannot be wrapped into [0, %g)", *x, L);
        return CUCO_ERR_INVALID_ARG;
    }
    return CUCO_OK;
}

}  // namespace

extern "C" {

const char *cuco_status_string(int status)
{
    switch (status) {
    case CUCO_OK: return "ok";
    case CUCO_ERR_INVALID_ARG: return "invalid argument";
    case CUCO_ERR_HIP: return "HIP runtime error";
    case CUCO_ERR_OUT_OF_MEMORY: return "out of device memory";
    case CUCO_ERR_STATE: return "invalid state";
    case CUCO_ERR_UNSUPPORTED: return "unsupported";
    default: return "unknown status";
    }
}

const char *cuco_last_error(void) { return g_last_error.c_str(); }

int cuco_create(const cuco_params *params, int64_t max_particles, int32_t device,
                cuco_engine **out)
{
    if (!params || !out) {
        set_error("null argument");
        return CUCO_ERR_INVALID_ARG;
    }
    *out = nullptr;
    const int M = params->pm_grid;
    if (M < 16 || M > 1024 || !fft_size_supported(M)) {
        set_error("pm_grid %d unsupported (need 16..1024, even, factors 2,3,5)", M);
        return CUCO_ERR_UNSUPPORTED;
    }
    if (!(params->box_size > 0) || !(params->asmth > 0) || !(params->rcut > 0) ||
        !(params->soft_h > 0) || !(params->G > 0) || params->reserved0 != 0) {
        set_error("box_size, asmth, rcut, soft_h and G must be positive");
        return CUCO_ERR_INVALID_ARG;
    }
    // The neighbour cell range must not wrap onto itself (each cell once).
    if (2.0 * params->rcut * M / params->box_size + 2.0 > M) {
        set_error("rcut %g too large for mesh %d in box %g", params->rcut, M,
                  params->box_size);
        return CUCO_ERR_INVALID_ARG;
    }
    if (max_particles <= 0 || max_particles > 0x7fffffffLL) {
        set_error("max_particles %lld out of range", (long long)max_particles);
        return CUCO_ERR_INVALID_ARG;
    }
    CUCO_HIP(hipSetDevice(device));
    cuco_engine *e = new cuco_engine();
    e->device = device;
    e->prm = *params;
    e->M = M;
    e->capacity = (uint32_t)max_particles;
    set_pair_params(e);

    GridInfo &gx = e->gx;
    gx.M = M;
    gx.key_mode = kKeyXMajor;
    gx.bits = ceil_log2((uint64_t)M);
    gx.key_bits = ceil_log2((uint64_t)M * M * M);
    gx.nkeys = (uint32_t)((uint64_t)M * M * M);
    gx.L = params->box_size;
    gx.fac = M / params->box_size;
    gx.cell = params->box_size / M;
    GridInfo &gm = e->gm;
    gm = gx;
    gm.key_mode = kKeyMorton;
    gm.bits = ceil_log2((uint64_t)M);
    gm.key_bits = 3 * gm.bits;
    gm.nkeys = (uint32_t)(1ull << gm.key_bits);

    e->pm_fac = params->G / (M_PI * params->box_size);  // pm_periodic.c:85-86
    e->pm_fac *= 1 / (2 * params->box_size / M);

    const size_t cap = e->capacity;
    int rc = CUCO_OK;
#define ALLOC(p, bytes) \
    if (rc == CUCO_OK) rc = dev_alloc(p, bytes)
    ALLOC(e->pos, sizeof(double4) * cap);
    ALLOC(e->vel, sizeof(double4) * cap);
    ALLOC(e->acc_pp64, sizeof(double4) * cap);
    ALLOC(e->acc_pm64, sizeof(double4) * cap);
    ALLOC(e->acc_pp32, sizeof(float4) * cap);
    ALLOC(e->acc_pm32, sizeof(float4) * cap);
    ALLOC(e->keys, sizeof(uint32_t) * cap);
    ALLOC(e->vals, sizeof(uint32_t) * cap);
    ALLOC(e->sc.keys_tmp, sizeof(uint32_t) * cap);
    ALLOC(e->sc.vals_tmp, sizeof(uint32_t) * cap);
    ALLOC(e->sc.counts, sizeof(uint32_t) * sort_count_words(e->capacity));
    ALLOC(e->sc.totals, sizeof(uint32_t) * 256);
    e->sc.capacity = e->capacity;
    const size_t offw = (size_t)std::max(gx.nkeys, gm.nkeys) + 1;
    ALLOC(e->off, sizeof(uint32_t) * offw);
    ALLOC(e->spos64, sizeof(double4) * cap);
    ALLOC(e->spos32, sizeof(float4) * cap);
    ALLOC(e->table, sizeof(float) * kNTab);
    ALLOC(e->table32, sizeof(float) * kTabPadded);
    ALLOC(e->dcount, sizeof(unsigned long long) * 2);
    // Tile arrays sized for the smallest tunable Morton block (most tiles).
    e->tiles.block_shift = gm.bits >= 4 ? 3 : gm.bits - 1;
    e->min_block_shift = std::min(2, e->tiles.block_shift);
    e->tiles.nblocks = gm.nkeys >> (3 * e->min_block_shift);
    e->tiles.max_tiles = (e->capacity + 63) / 64 + e->tiles.nblocks;
    ALLOC(e->tiles.tile_start, sizeof(uint32_t) * e->tiles.max_tiles);
    ALLOC(e->tiles.tile_count, sizeof(uint32_t) * e->tiles.max_tiles);
    ALLOC(e->tiles.block_tiles, sizeof(uint32_t) * e->tiles.nblocks);
    ALLOC(e->tiles.counters, sizeof(uint32_t) * 4);
    ALLOC(e->tiles.scan_scratch, sizeof(uint32_t) * scan_scratch_words(e->tiles.nblocks));
#undef ALLOC
    if (rc == CUCO_OK) {
        hipError_t err = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
        if (err != hipSuccess) {
            set_error("hipStreamCreate failed: %s", hipGetErrorString(err));
            rc = CUCO_ERR_HIP;
        }
    }
    if (rc == CUCO_OK) rc = upload_tables(e);
    if (rc == CUCO_OK) {
        int ncu = 0;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || ncu <= 0)
            ncu = 256;
        e->ncu = ncu;
        e->pp32_workgroups = ncu * e->pp32_wg_per_cu;
    }
    if (rc != CUCO_OK) {
        std::string msg = g_last_error;
        cuco_destroy(e);
        g_last_error = msg;
        return rc;
    }
    *out = e;
    return CUCO_OK;
}

int cuco_destroy(cuco_engine *e)
{
    if (!e) return CUCO_OK;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    destroy_graph(e);
    for (StageEvents *ev : e->pending) e->spare.push_back(ev);
    if (e->cur_events) e->spare.push_back(e->cur_events);
    for (StageEvents *ev : e->spare) {
        for (int s = 0; s < CUCO_STAGE_COUNT; ++s) {
            (void)hipEventDestroy(ev->start[s]);
            (void)hipEventDestroy(ev->stop[s]);
        }
        delete ev;
    }
    free_ptr(e->pos); free_ptr(e->vel);
    free_ptr(e->acc_pp64); free_ptr(e->acc_pm64);
    free_ptr(e->acc_pp32); free_ptr(e->acc_pm32);
    free_ptr(e->keys); free_ptr(e->vals);
    free_ptr(e->sc.keys_tmp); free_ptr(e->sc.vals_tmp);
    free_ptr(e->sc.counts); free_ptr(e->sc.totals);
    free_ptr(e->off); free_ptr(e->spos64); free_ptr(e->spos32);
    free_ptr(e->table); free_ptr(e->table32); free_ptr(e->dcount);
    free_ptr(e->tiles.tile_start); free_ptr(e->tiles.tile_count);
    free_ptr(e->tiles.block_tiles); free_ptr(e->tiles.counters);
    free_ptr(e->tiles.scan_scratch);
    for (int k = 0; k < 2; ++k) {
        free_ptr(e->rho[k]); free_ptr(e->phi[k]); free_ptr(e->cgrid[k]);
        free_ptr(e->dep_scratch[k]);
        free_ptr(e->tb[k].tw); free_ptr(e->tb[k].gexp); free_ptr(e->tb[k].gsinc);
    }
    free_ptr(e->dfactors); free_ptr(e->dstep);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return CUCO_OK;
}

int cuco_set_softening(cuco_engine *e, double soft_h)
{
    CHECK_ENGINE(e);
    if (!(soft_h > 0)) {
        set_error("softening must be positive");
        return CUCO_ERR_INVALID_ARG;
    }
    e->prm.soft_h = soft_h;
    set_pair_params(e);
    destroy_graph(e);
    return CUCO_OK;
}

int cuco_upload_particles(cuco_engine *e, int64_t n, const double *pos,
                          const double *vel, const double *mass,
                          double uniform_mass)
{
    CHECK_ENGINE(e);
    if (n <= 0 || n > (int64_t)e->capacity || !pos) {
        set_error("n=%lld outside [1, %u] or null positions", (long long)n, e->capacity);
        return CUCO_ERR_INVALID_ARG;
    }
    std::vector<double4> hp((size_t)n), hv((size_t)n);
    const double L = e->prm.box_size;
    double common = mass ? mass[0] : uniform_mass;
    for (int64_t i = 0; i < n; ++i) {
        double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
        if (!std::isfinite(x) || !std::isfinite(y) || !std::isfinite(z)) {
            set_error("non-finite position for particle %lld", (long long)i);
            return CUCO_ERR_INVALID_ARG;
        }
        TRY(wrap_host(&x, L));
        TRY(wrap_host(&y, L));
        TRY(wrap_host(&z, L));
        hp[i] = make_double4(x, y, z, mass ? mass[i] : uniform_mass);
        if (hp[i].w != common) common = 0;
        hv[i] = vel ? make_double4(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2], 0.0)
                    : make_double4(0, 0, 0, 0);
    }
    CUCO_HIP(hipMemcpyAsync(e->pos, hp.data(), sizeof(double4) * n, hipMemcpyHostToDevice, e->stream));
    CUCO_HIP(hipMemcpyAsync(e->vel, hv.data(), sizeof(double4) * n, hipMemcpyHostToDevice, e->stream));
    CUCO_HIP(hipStreamSynchronize(e->stream));
    if ((uint32_t)n != e->n) destroy_graph(e);
    e->n = (uint32_t)n;
    e->uniform_mass = common > 0 ? common : 0;
    e->have_particles = true;
    e->sorted_mode = -1;
    e->pp_prec = e->pm_prec = -1;
    return CUCO_OK;
}

int cuco_download_particles(cuco_engine *e, double *pos, double *vel)
{
    CHECK_ENGINE(e);
    TRY(check_ready(e));
    std::vector<double4> h(e->n);
    for (int which = 0; which < 2; ++which) {
        double *dst = which == 0 ? pos : vel;
        if (!dst) continue;
        CUCO_HIP(hipMemcpyAsync(h.data(), which == 0 ? e->pos : e->vel,
                                sizeof(double4) * e->n, hipMemcpyDeviceToHost, e->stream));
        CUCO_HIP(hipStreamSynchronize(e->stream));
        for (uint32_t i = 0; i < e->n; ++i) {
            dst[3 * i] = h[i].x;
            dst[3 * i + 1] = h[i].y;
            dst[3 * i + 2] = h[i].z;
        }
    }
    return CUCO_OK;
}

int cuco_download_accel(cuco_engine *e, double *acc_pp, double *acc_pm)
{
    CHECK_ENGINE(e);
    TRY(check_ready(e));
    for (int which = 0; which < 2; ++which) {
        double *dst = which == 0 ? acc_pp : acc_pm;
        const int prec = which == 0 ? e->pp_prec : e->pm_prec;
        if (!dst) continue;
        if (prec < 0) {
            set_error("%s force not computed", which == 0 ? "PP" : "PM");
            return CUCO_ERR_STATE;
        }
        if (prec == CUCO_FP64) {
            std::vector<double4> h(e->n);
            CUCO_HIP(hipMemcpyAsync(h.data(), which == 0 ? e->acc_pp64 : e->acc_pm64,
                                    sizeof(double4) * e->n, hipMemcpyDeviceToHost, e->stream));
            CUCO_HIP(hipStreamSynchronize(e->stream));
            for (uint32_t i = 0; i < e->n; ++i) {
                dst[3 * i] = h[i].x;
                dst[3 * i + 1] = h[i].y;
                dst[3 * i + 2] = h[i].z;
            }
        } else {
            std::vector<float4> h(e->n);
            CUCO_HIP(hipMemcpyAsync(h.data(), which == 0 ? e->acc_pp32 : e->acc_pm32,
                                    sizeof(float4) * e->n, hipMemcpyDeviceToHost, e->stream));
            CUCO_HIP(hipStreamSynchronize(e->stream));
            for (uint32_t i = 0; i < e->n; ++i) {
                dst[3 * i] = h[i].x;
                dst[3 * i + 1] = h[i].y;
                dst[3 * i + 2] = h[i].z;
            }
        }
    }
    return CUCO_OK;
}

int cuco_pp_force(cuco_engine *e, int32_t precision)
{
    CHECK_ENGINE(e);
    TRY(check_prec(precision));
    TRY(check_ready(e));
    TRY(begin_eval(e));
    int rc = do_pp(e, precision, nullptr);
    end_eval(e);
    return rc;
}

int cuco_pm_force(cuco_engine *e, int32_t precision)
{
    CHECK_ENGINE(e);
    TRY(check_prec(precision));
    TRY(check_ready(e));
    TRY(begin_eval(e));
    int rc = do_pm(e, precision);
    end_eval(e);
    return rc;
}

int cuco_compute_accelerations(cuco_engine *e, int32_t precision, int32_t do_pm_flag)
{
    CHECK_ENGINE(e);
    TRY(check_prec(precision));
    TRY(check_ready(e));
    // A complete evaluation, as gravity_linklist_cuda is: keys and sort are
    // recomputed even when the positions have not moved.
    e->sorted_mode = -1;
    TRY(begin_eval(e));
    int rc = do_forces(e, precision, do_pm_flag);
    end_eval(e);
    return rc;
}

int cuco_drift(cuco_engine *e, double drift_factor)
{
    CHECK_ENGINE(e);
    TRY(check_ready(e));
    TRY(launch_drift(e->pos, e->vel, e->n, e->prm.box_size, drift_factor, e->stream));
    e->sorted_mode = -1;
    return CUCO_OK;
}

int cuco_kick(cuco_engine *e, double kick_pp, double kick_pm)
{
    CHECK_ENGINE(e);
    TRY(check_ready(e));
    return kick_current(e, kick_pp, kick_pm);
}

int cuco_kdk_steps(cuco_engine *e, int32_t precision, int32_t nsteps,
                   const double *factors, int32_t use_graph)
{
    CHECK_ENGINE(e);
    TRY(check_prec(precision));
    TRY(check_ready(e));
    if (nsteps < 0 || (nsteps > 0 && !factors)) {
        set_error("bad step count or null factors");
        return CUCO_ERR_INVALID_ARG;
    }
    if (e->pp_prec != precision || e->pm_prec != precision) {
        set_error("forces of the requested precision must be current before KDK");
        return CUCO_ERR_STATE;
    }
    if (!use_graph) {
        for (int s = 0; s < nsteps; ++s) {
            TRY(kick_current(e, factors[3 * s], factors[3 * s]));
            TRY(launch_drift(e->pos, e->vel, e->n, e->prm.box_size, factors[3 * s + 1], e->stream));
            e->sorted_mode = -1;
            TRY(begin_eval(e));
            int rc = do_forces(e, precision, 1);
            end_eval(e);
            if (rc) return rc;
            TRY(kick_current(e, factors[3 * s + 2], factors[3 * s + 2]));
        }
        CUCO_HIP(hipStreamSynchronize(e->stream));
        return CUCO_OK;
    }
    if (nsteps > e->factors_cap) {
        free_ptr(e->dfactors);
        TRY(dev_alloc(e->dfactors, sizeof(double) * 3 * nsteps));
        e->factors_cap = nsteps;
        destroy_graph(e);
    }
    if (!e->dstep) TRY(dev_alloc(e->dstep, sizeof(uint32_t) * 4));
    CUCO_HIP(hipMemcpyAsync(e->dfactors, factors, sizeof(double) * 3 * nsteps,
                            hipMemcpyHostToDevice, e->stream));
    TRY(launch_fill_u32(e->dstep, 0u, 1, e->stream));
    if (!e->graph_exec || e->graph_prec != precision) {
        destroy_graph(e);
        // Forces (and so the PM buffers) exist already: capture allocates nothing.
        TRY(precision == CUCO_FP64 ? ensure_pm<double>(e) : ensure_pm<float>(e));
        CUCO_HIP(hipStreamSynchronize(e->stream));
        CUCO_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeRelaxed));
        e->capturing = true;
        const int fp32 = precision == CUCO_FP32;
        const void *app = fp32 ? (const void *)e->acc_pp32 : (const void *)e->acc_pp64;
        const void *apm = fp32 ? (const void *)e->acc_pm32 : (const void *)e->acc_pm64;
        int rc = launch_kick_table(e->vel, app, apm, fp32, e->n, e->dfactors, e->dstep, 0, e->stream);
        if (!rc) rc = launch_drift_table(e->pos, e->vel, e->n, e->prm.box_size, e->dfactors, e->dstep, e->stream);
        e->sorted_mode = -1;
        if (!rc) rc = do_forces(e, precision, 1);
        if (!rc) rc = launch_kick_table(e->vel, app, apm, fp32, e->n, e->dfactors, e->dstep, 2, e->stream);
        if (!rc) rc = launch_step_advance(e->dstep, e->stream);
        hipGraph_t g = nullptr;
        hipError_t err = hipStreamEndCapture(e->stream, &g);
        e->capturing = false;
        e->sorted_mode = -1;
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (err != hipSuccess) {
            set_error("hipStreamEndCapture failed: %s", hipGetErrorString(err));
            return CUCO_ERR_HIP;
        }
        e->graph = g;
        err = hipGraphInstantiate(&e->graph_exec, g, nullptr, nullptr, 0);
        if (err != hipSuccess) {
            set_error("hipGraphInstantiate failed: %s", hipGetErrorString(err));
            destroy_graph(e);
            return CUCO_ERR_HIP;
        }
        e->graph_prec = precision;
    }
    for (int s = 0; s < nsteps; ++s) CUCO_HIP(hipGraphLaunch(e->graph_exec, e->stream));
    CUCO_HIP(hipStreamSynchronize(e->stream));
    e->sorted_mode = -1;  // the sort buffers hold the last step's order
    return CUCO_OK;
}

int cuco_pmforce_periodic(cuco_engine *e, cuco_particle *P, int64_t n, int32_t precision)
{
    CHECK_ENGINE(e);
    if (!P) {
        set_error("null particle array");
        return CUCO_ERR_INVALID_ARG;
    }
    std::vector<double> pos((size_t)n * 3), mass((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        for (int d = 0; d < 3; ++d) pos[3 * i + d] = P[i].Pos[d];
        mass[i] = P[i].Mass;
    }
    TRY(cuco_upload_particles(e, n, pos.data(), nullptr, mass.data(), 0.0));
    TRY(cuco_pm_force(e, precision));
    std::vector<double> acc((size_t)n * 3);
    TRY(cuco_download_accel(e, nullptr, acc.data()));
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) P[i].GravPM[d] = (float)acc[3 * i + d];
    return CUCO_OK;
}

int cuco_gravity_linklist(cuco_engine *e, cuco_particle *P, int64_t n, int32_t precision)
{
    CHECK_ENGINE(e);
    if (!P) {
        set_error("null particle array");
        return CUCO_ERR_INVALID_ARG;
    }
    std::vector<double> pos((size_t)n * 3), mass((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        for (int d = 0; d < 3; ++d) pos[3 * i + d] = P[i].Pos[d];
        mass[i] = P[i].Mass;
    }
    TRY(cuco_upload_particles(e, n, pos.data(), nullptr, mass.data(), 0.0));
    TRY(cuco_pp_force(e, precision));
    std::vector<double> acc((size_t)n * 3);
    TRY(cuco_download_accel(e, acc.data(), nullptr));
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) P[i].GravAccel[d] = (float)acc[3 * i + d];
    return CUCO_OK;
}

int cuco_device_alloc(void **dptr, size_t bytes)
{
    if (!dptr) {
        set_error("null pointer");
        return CUCO_ERR_INVALID_ARG;
    }
    return dev_alloc(*dptr, bytes);
}

int cuco_device_free(void *dptr)
{
    if (dptr) CUCO_HIP(hipFree(dptr));
    return CUCO_OK;
}

int cuco_copy_to_device(void *dst, const void *src, size_t bytes)
{
    CUCO_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return CUCO_OK;
}

int cuco_copy_from_device(void *dst, const void *src, size_t bytes)
{
    CUCO_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return CUCO_OK;
}

int cuco_copy_device_to_device(void *dst, const void *src, size_t bytes)
{
    if (!dst || !src || bytes % 16) {
        set_error("null pointer or size %zu not a multiple of 16", bytes);
        return CUCO_ERR_INVALID_ARG;
    }
    TRY(launch_copy16(dst, src, bytes, nullptr));
    CUCO_HIP(hipDeviceSynchronize());
    return CUCO_OK;
}

int cuco_radix_sort(uint32_t *d_keys, uint32_t *d_values, uint32_t n, uint32_t key_bits)
{
    if (n == 0) return CUCO_OK;
    if (!d_keys || key_bits == 0 || key_bits > 32) {
        set_error("null keys or key_bits %u outside 1..32", key_bits);
        return CUCO_ERR_INVALID_ARG;
    }
    SortScratch sc;
    sc.capacity = n;
    int rc = dev_alloc(sc.keys_tmp, sizeof(uint32_t) * n);
    if (!rc && d_values) rc = dev_alloc(sc.vals_tmp, sizeof(uint32_t) * n);
    if (!rc) rc = dev_alloc(sc.counts, sizeof(uint32_t) * sort_count_words(n));
    if (!rc) rc = dev_alloc(sc.totals, sizeof(uint32_t) * 256);
    uint32_t *ko = nullptr, *vo = nullptr;
    if (!rc) rc = radix_sort_pairs(d_keys, d_values, false, n, (int)key_bits, sc, nullptr, &ko, &vo);
    if (!rc && ko != d_keys) {
        if (hipMemcpy(d_keys, ko, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice) != hipSuccess ||
            (d_values && hipMemcpy(d_values, vo, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice) != hipSuccess)) {
            set_error("hipMemcpy after sort failed");
            rc = CUCO_ERR_HIP;
        }
    }
    if (!rc && hipDeviceSynchronize() != hipSuccess) {
        set_error("sort failed on the device");
        rc = CUCO_ERR_HIP;
    }
    free_ptr(sc.keys_tmp); free_ptr(sc.vals_tmp); free_ptr(sc.counts); free_ptr(sc.totals);
    return rc;
}

int cuco_get_sorted(cuco_engine *e, uint32_t *keys, uint32_t *index)
{
    CHECK_ENGINE(e);
    TRY(check_ready(e));
    if (e->sorted_mode < 0) {
        set_error("no current sort (positions changed since the last force)");
        return CUCO_ERR_STATE;
    }
    if (keys) CUCO_HIP(hipMemcpyAsync(keys, e->skeys, sizeof(uint32_t) * e->n, hipMemcpyDeviceToHost, e->stream));
    if (index) CUCO_HIP(hipMemcpyAsync(index, e->perm, sizeof(uint32_t) * e->n, hipMemcpyDeviceToHost, e->stream));
    CUCO_HIP(hipStreamSynchronize(e->stream));
    return CUCO_OK;
}

int cuco_get_cell_table(cuco_engine *e, uint32_t *cell_start, uint32_t *cell_end)
{
    CHECK_ENGINE(e);
    TRY(check_ready(e));
    if (e->sorted_mode != kKeyXMajor) {
        set_error("cell table is available after an FP64 (reference order) evaluation");
        return CUCO_ERR_STATE;
    }
    const uint32_t nc = e->gx.nkeys;
    uint32_t *ds = nullptr, *de = nullptr;
    TRY(dev_alloc(ds, sizeof(uint32_t) * nc));
    int rc = dev_alloc(de, sizeof(uint32_t) * nc);
    if (!rc) rc = launch_cell_start_end(e->off, nc, ds, de, e->stream);
    if (!rc && (hipMemcpyAsync(cell_start, ds, sizeof(uint32_t) * nc, hipMemcpyDeviceToHost, e->stream) != hipSuccess ||
                hipMemcpyAsync(cell_end, de, sizeof(uint32_t) * nc, hipMemcpyDeviceToHost, e->stream) != hipSuccess ||
                hipStreamSynchronize(e->stream) != hipSuccess)) {
        set_error("cell table copy failed");
        rc = CUCO_ERR_HIP;
    }
    free_ptr(ds);
    free_ptr(de);
    return rc;
}

int cuco_get_shortrange_table(cuco_engine *e, float *table)
{
    CHECK_ENGINE(e);
    CUCO_HIP(hipMemcpy(table, e->table, sizeof(float) * kNTab, hipMemcpyDeviceToHost));
    return CUCO_OK;
}

int cuco_get_pm_grid(cuco_engine *e, int32_t which, double *out)
{
    CHECK_ENGINE(e);
    if (e->pm_prec < 0 || (which != 0 && which != 1) || !out) {
        set_error("no PM evaluation, or which not in {0,1}");
        return CUCO_ERR_STATE;
    }
    const int k = e->pm_prec == CUCO_FP64 ? 0 : 1;
    const size_t M3 = (size_t)e->M * e->M * e->M;
    void *src = which == 0 ? e->rho[k] : e->phi[k];
    CUCO_HIP(hipStreamSynchronize(e->stream));
    if (k == 0) {
        CUCO_HIP(hipMemcpy(out, src, sizeof(double) * M3, hipMemcpyDeviceToHost));
    } else {
        std::vector<float> h(M3);
        CUCO_HIP(hipMemcpy(h.data(), src, sizeof(float) * M3, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < M3; ++i) out[i] = h[i];
    }
    return CUCO_OK;
}

int cuco_count_pairs(cuco_engine *e, uint64_t *pairs_in, uint64_t *fp32_candidates)
{
    CHECK_ENGINE(e);
    TRY(check_ready(e));
    CUCO_HIP(hipMemsetAsync(e->dcount, 0, sizeof(unsigned long long) * 2, e->stream));
    if (pairs_in) TRY(do_pp(e, CUCO_FP64, e->dcount));
    if (fp32_candidates) TRY(do_pp(e, CUCO_FP32, e->dcount + 1));
    unsigned long long h[2] = {0, 0};
    CUCO_HIP(hipMemcpyAsync(h, e->dcount, sizeof(h), hipMemcpyDeviceToHost, e->stream));
    CUCO_HIP(hipStreamSynchronize(e->stream));
    if (pairs_in) *pairs_in = h[0];
    if (fp32_candidates) *fp32_candidates = h[1];
    return CUCO_OK;
}

int cuco_get_tuning(cuco_engine *e, cuco_tuning *t)
{
    CHECK_ENGINE(e);
    if (!t) {
        set_error("null tuning");
        return CUCO_ERR_INVALID_ARG;
    }
    t->pp_block_shift = e->tiles.block_shift;
    t->pp_workgroups_per_cu = e->pp32_wg_per_cu;
    t->reserved0 = t->reserved1 = 0;
    return CUCO_OK;
}

int cuco_set_tuning(cuco_engine *e, const cuco_tuning *t)
{
    CHECK_ENGINE(e);
    if (!t || t->pp_block_shift < e->min_block_shift || t->pp_block_shift > 4 ||
        t->pp_block_shift >= e->gm.bits || t->pp_workgroups_per_cu < 1 ||
        t->pp_workgroups_per_cu > 32) {
        set_error("tuning out of range");
        return CUCO_ERR_INVALID_ARG;
    }
    if (t->pp_block_shift != e->tiles.block_shift) {
        e->tiles.block_shift = t->pp_block_shift;
        if (e->sorted_mode == kKeyMorton) e->sorted_mode = -1;  // rebuild tiles
        destroy_graph(e);
    }
    if (t->pp_workgroups_per_cu != e->pp32_wg_per_cu) {
        e->pp32_wg_per_cu = t->pp_workgroups_per_cu;
        e->pp32_workgroups = e->ncu * e->pp32_wg_per_cu;
        destroy_graph(e);
    }
    return CUCO_OK;
}

int cuco_autotune(cuco_engine *e, cuco_tuning *chosen, double *best_ms)
{
    CHECK_ENGINE(e);
    TRY(check_ready(e));
    hipEvent_t a, b;
    CUCO_HIP(hipEventCreate(&a));
    CUCO_HIP(hipEventCreate(&b));
    const bool timing = e->timing;
    e->timing = false;
    cuco_tuning best{e->tiles.block_shift, e->pp32_wg_per_cu, 0, 0};
    float best_t = 1e30f;
    int rc = CUCO_OK;
    for (int shift = e->min_block_shift; shift <= std::min(4, e->gm.bits - 1) && !rc; ++shift) {
        for (int wg : {4, 8, 16}) {
            cuco_tuning t{shift, wg, 0, 0};
            rc = cuco_set_tuning(e, &t);
            if (!rc) rc = sort_particles(e, kKeyMorton);
            float ms = 1e30f;
            for (int rep = 0; rep < 2 && !rc; ++rep) {
                if (hipEventRecord(a, e->stream) != hipSuccess) rc = CUCO_ERR_HIP;
                if (!rc) rc = launch_pp32(e->spos32, e->skeys, e->perm, e->off, e->n, e->gm, e->pp,
                                          e->table32, e->tiles, e->acc_pp32, nullptr,
                                          e->pp32_workgroups, e->uniform_mass, e->stream);
                if (!rc && (hipEventRecord(b, e->stream) != hipSuccess ||
                            hipEventSynchronize(b) != hipSuccess)) rc = CUCO_ERR_HIP;
                float t_ms = 0;
                if (!rc && hipEventElapsedTime(&t_ms, a, b) == hipSuccess) ms = std::min(ms, t_ms);
            }
            if (!rc && ms < best_t) {
                best_t = ms;
                best = t;
            }
            if (rc) break;
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    e->timing = timing;
    if (rc) {
        if (rc == CUCO_ERR_HIP && g_last_error.empty()) set_error("autotune timing failed");
        return rc;
    }
    TRY(cuco_set_tuning(e, &best));
    e->pp_prec = CUCO_FP32;
    if (chosen) *chosen = best;
    if (best_ms) *best_ms = best_t;
    return CUCO_OK;
}

int cuco_set_timing(cuco_engine *e, int32_t enable)
{
    CHECK_ENGINE(e);
    e->timing = enable != 0;
    return CUCO_OK;
}

int cuco_get_stage_times(cuco_engine *e, double *ms, int32_t nstages,
                         int32_t *evaluations, int32_t reset)
{
    CHECK_ENGINE(e);
    TRY(harvest(e));
    for (int s = 0; s < nstages && s < CUCO_STAGE_COUNT; ++s) ms[s] = e->stage_ms[s];
    if (evaluations) *evaluations = e->evals;
    if (reset) {
        for (double &v : e->stage_ms) v = 0;
        e->evals = 0;
    }
    return CUCO_OK;
}

const char *cuco_stage_name(int32_t stage)
{
    static const char *names[CUCO_STAGE_COUNT] = {
        "keys", "sort", "cells", "pp", "deposit", "fft_fwd", "green_x", "fft_inv", "interp"};
    return (stage >= 0 && stage < CUCO_STAGE_COUNT) ? names[stage] : "unknown";
}

int cuco_synchronize(cuco_engine *e)
{
    CHECK_ENGINE(e);
    CUCO_HIP(hipStreamSynchronize(e->stream));
    return CUCO_OK;
}

int cuco_fft3d_r2c(int32_t M, int32_t precision, const void *d_in, void *d_out)
{
    TRY(check_prec(precision));
    if (!fft_size_supported(M) || !d_in || !d_out) {
        set_error("unsupported FFT size %d or null pointer", M);
        return CUCO_ERR_UNSUPPORTED;
    }
    cuco_engine tmp;
    tmp.M = M;
    tmp.prm.asmth = 1.0;
    tmp.prm.box_size = 1.0;
    PmTables tb;
    int rc = precision == CUCO_FP64 ? make_pm_tables<double>(&tmp, tb) : make_pm_tables<float>(&tmp, tb);
    if (!rc) {
        if (precision == CUCO_FP64) {
            rc = pm_fft_forward<double>(static_cast<const double *>(d_in), static_cast<Cx<double> *>(d_out), M, tb, nullptr);
            if (!rc) rc = pm_fft_x<double>(static_cast<Cx<double> *>(d_out), M, false, nullptr, tb);
        } else {
            rc = pm_fft_forward<float>(static_cast<const float *>(d_in), static_cast<Cx<float> *>(d_out), M, tb, nullptr);
            if (!rc) rc = pm_fft_x<float>(static_cast<Cx<float> *>(d_out), M, false, nullptr, tb);
        }
    }
    if (!rc && hipDeviceSynchronize() != hipSuccess) {
        set_error("FFT failed on the device");
        rc = CUCO_ERR_HIP;
    }
    free_ptr(tb.tw); free_ptr(tb.gexp); free_ptr(tb.gsinc);
    return rc;
}

int cuco_fft3d_c2r(int32_t M, int32_t precision, void *d_in, void *d_out)
{
    TRY(check_prec(precision));
    if (!fft_size_supported(M) || !d_in || !d_out) {
        set_error("unsupported FFT size %d or null pointer", M);
        return CUCO_ERR_UNSUPPORTED;
    }
    cuco_engine tmp;
    tmp.M = M;
    tmp.prm.asmth = 1.0;
    tmp.prm.box_size = 1.0;
    PmTables tb;
    int rc = precision == CUCO_FP64 ? make_pm_tables<double>(&tmp, tb) : make_pm_tables<float>(&tmp, tb);
    if (!rc) {
        if (precision == CUCO_FP64) {
            rc = pm_fft_x<double>(static_cast<Cx<double> *>(d_in), M, true, nullptr, tb);
            if (!rc) rc = pm_fft_inverse<double>(static_cast<Cx<double> *>(d_in), static_cast<double *>(d_out), M, tb, nullptr);
        } else {
            rc = pm_fft_x<float>(static_cast<Cx<float> *>(d_in), M, true, nullptr, tb);
            if (!rc) rc = pm_fft_inverse<float>(static_cast<Cx<float> *>(d_in), static_cast<float *>(d_out), M, tb, nullptr);
        }
    }
    if (!rc && hipDeviceSynchronize() != hipSuccess) {
        set_error("FFT failed on the device");
        rc = CUCO_ERR_HIP;
    }
    free_ptr(tb.tw); free_ptr(tb.gexp); free_ptr(tb.gsinc);
    return rc;
}

}  // extern "C"
