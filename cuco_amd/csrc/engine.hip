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

// Host view of one box: its parameters and particle slots.
struct HostBox {
    cuco_params prm{};
    uint32_t cap = 0;  // particle slots
    uint32_t n = 0;    // particles now
    PairParams pp{};
    GridInfo g[2]{};   // per KeyMode
    double pm_fac = 0;
    double mass_max = 0;    // largest particle mass of the box
};

struct cuco_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    // Second stream for the PM chain when it overlaps the pair force, and
    // the fork / join events between the two.
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    bool pm_overlap = false;
    bool ensemble = false;  // box sizes fixed at creation
    std::vector<HostBox> box;
    uint32_t capacity = 0;
    uint32_t n = 0;
    bool have_particles = false;
    double uniform_mass = 0;  // > 0 when every particle has this mass

    // Box descriptors and launch maps on the device, with host mirrors.
    BoxSet bs;
    std::vector<BoxDev> hdev[2];
    std::vector<uint32_t> hmaps;
    uint32_t nkeys[2] = {0, 0};
    int key_bits[2] = {0, 0};
    int min_bits = 0;  // smallest Morton bits per coordinate over boxes
    size_t grid_elems = 0, cgrid_elems = 0, tab_elems = 0;
    P32 *p32 = nullptr;  // per box
    bool table_cut = false;

    double4 *pos = nullptr, *vel = nullptr;
    double4 *acc_pp64 = nullptr, *acc_pm64 = nullptr;
    float4 *acc_pp32 = nullptr, *acc_pm32 = nullptr;
    int pp_prec = -1, pm_prec = -1;

    uint32_t *keys = nullptr, *vals = nullptr;
    SortScratch sc;
    uint32_t *skeys = nullptr, *perm = nullptr;
    uint32_t *off = nullptr;
    uint4 *runs = nullptr;     // long empty key ranges of the cell table
    uint32_t *nruns = nullptr;
    uint32_t run_cap = 0;
    int sorted_mode = -1;
    double4 *spos64 = nullptr;
    float4 *spos32 = nullptr;

    float *table = nullptr;    // 1000 entries, gravlist.cpp:56-59
    float *table32 = nullptr;  // 1024 entries for the FP32 kernel
    float htable[kNTab];
    TileInfo tiles;
    ActiveTiles atiles;  // FP32 active-set tiles of the integer timeline
    PairSym psym;  // each-pair-once variant of the FP32 pair force
    int pp32_workgroups = 0;
    int pp32_wg_per_cu = 8;
    int pp_cell_level = 0;
    int pp_wide_tiles = 0;  // each-pair-once tiles of 128 targets
    int ncu = 256;
    int min_block_shift = 2;

    void *rho[2] = {nullptr, nullptr};
    void *phi[2] = {nullptr, nullptr};
    void *cgrid[2] = {nullptr, nullptr};
    void *dep_scratch[2] = {nullptr, nullptr};
    PmTables tb[2];

    // Integer timeline (individual power-of-two timesteps).
    int32_t *ti_beg = nullptr, *ti_end = nullptr;
    int32_t *ti_scalars = nullptr;  // [0] min Ti_endstep, [1] error flags, [2] active count
    unsigned long long *tl_work = nullptr;  // pair-work estimate: all, active
    double *kick_table = nullptr;
    double *drift_table = nullptr;
    bool tl_init = false;
    cuco_timeline tl{};
    std::vector<double> drift_tab, kick_tab;
    double tl_log_begin = 0, tl_log_max = 0, tl_interval = 0;
    cuco_timeline_state tls{};

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

// Handler of the function-try-blocks around the C entry points that build
// host arrays: no C++ exception crosses the C boundary.
#define CUCO_CATCH                                                          \
    catch (const std::bad_alloc &) {                                        \
        ::cuco::set_error("host memory allocation failed");                 \
        return CUCO_ERR_OUT_OF_MEMORY;                                      \
    }                                                                       \
    catch (const std::exception &ex_) {                                     \
        ::cuco::set_error("host exception: %s", ex_.what());                \
        return CUCO_ERR_INVALID_ARG;                                        \
    }

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

int stage_start(cuco_engine *e, int s, hipStream_t st = nullptr)
{
    if (e->cur_events) {
        CUCO_HIP(hipEventRecord(e->cur_events->start[s], st ? st : e->stream));
        e->cur_events->used[s] = true;
    }
    return CUCO_OK;
}

int stage_stop(cuco_engine *e, int s, hipStream_t st = nullptr)
{
    if (e->cur_events) CUCO_HIP(hipEventRecord(e->cur_events->stop[s], st ? st : e->stream));
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

void set_pair_params(HostBox &b)
{
    // gravlist.cpp:36-40, :61-70
    const cuco_params &p = b.prm;
    PairParams &q = b.pp;
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
    b.pm_fac = p.G / (M_PI * p.box_size);  // pm_periodic.c:85-86
    b.pm_fac *= 1 / (2 * p.box_size / p.pm_grid);
}

// The FP32 table, the per-box FP32 pair constants and the choice between
// the table cut and the explicit r^2 test.  The table is shared by all
// boxes, so its zero tail starts at the cut bin only when every box puts
// rcut on the same bin edge.
int update_pair_constants(cuco_engine *e)
{
    make_table(e->htable);
    bool same_edge = true;
    double edge0 = e->box[0].pp.asmthfac * e->box[0].pp.rcut;
    bool cut = true;
    std::vector<P32> h;
    for (const HostBox &b : e->box) {
        const double edge = b.pp.asmthfac * b.pp.rcut;
        same_edge = same_edge && fabs(edge - edge0) <= 1e-9 * edge0;
        cut = cut && pp32_table_cut(b.pp);
        h.push_back(make_p32(b.pp, b.g[kKeyMorton], e->uniform_mass, b.mass_max, b.cap));
    }
    const bool aligned = fabs(edge0 - nearbyint(edge0)) < 1e-9 * edge0;
    e->table_cut = cut && same_edge;
    const int kcut = (same_edge && aligned) ? (int)nearbyint(edge0) : kNTab;
    float t32[kTabPadded];
    for (int k = 0; k < kTabPadded; ++k)
        t32[k] = (k < kNTab && k < kcut) ? e->htable[k] : 0.0f;
    CUCO_HIP(hipMemcpy(e->table, e->htable, sizeof(float) * kNTab, hipMemcpyHostToDevice));
    CUCO_HIP(hipMemcpy(e->table32, t32, sizeof(t32), hipMemcpyHostToDevice));
    CUCO_HIP(hipMemcpy(e->p32, h.data(), sizeof(P32) * h.size(), hipMemcpyHostToDevice));
    return CUCO_OK;
}

// Box descriptors and block maps for the current particle counts.
int update_boxes(cuco_engine *e)
{
    const int nb = (int)e->box.size();
    const size_t stride = (size_t)nb + 1;
    e->hmaps.assign(kMapCount * stride, 0u);
    for (int m = 0; m < 2; ++m) e->hdev[m].assign(nb, BoxDev{});
    uint32_t p0 = 0, wave0 = 0, kb[2] = {0, 0};
    uint64_t grid0 = 0, cgrid0 = 0, tab0 = 0;
    for (int b = 0; b < nb; ++b) {
        const HostBox &h = e->box[b];
        const int M = h.prm.pm_grid;
        for (int m = 0; m < 2; ++m) {
            BoxDev &d = e->hdev[m][b];
            d.g = h.g[m];
            d.key_base = kb[m];
            d.p0 = p0;
            d.n = h.n;
            d.wave0 = wave0;
            d.grid0 = grid0;
            d.cgrid0 = cgrid0;
            d.tab0 = (uint32_t)tab0;
            d.pm_fac = h.pm_fac;
            d.dep_mmax = h.mass_max;
            d.half = make_plan(M / 2);
            d.full = make_plan(M);
            kb[m] += h.g[m].nkeys;
        }
        uint32_t *mp = e->hmaps.data();
        mp[kMapPart * stride + b + 1] = (h.n + 255) / 256;
        for (int map = kMapGatherX; map < kMapCount; ++map)
            mp[map * stride + b + 1] = pm_blocks(M, map, e->bs.fft_split);
        p0 += h.n;
        wave0 += (h.cap + 63) / 64;
        grid0 += (uint64_t)M * M * M;
        cgrid0 += (uint64_t)M * M * (M / 2 + 1);
        tab0 += M;
    }
    for (int map = 0; map < kMapCount; ++map) {
        uint32_t *mp = e->hmaps.data() + map * stride;
        for (int b = 0; b < nb; ++b) mp[b + 1] += mp[b];
        e->bs.total[map] = mp[nb];
    }
    e->bs.fft32_mask = 0;
    for (const HostBox &h : e->box) e->bs.fft32_mask |= fft32_mask_bit(h.prm.pm_grid);
    for (int m = 0; m < 2; ++m)
        CUCO_HIP(hipMemcpy(e->bs.dev[m], e->hdev[m].data(), sizeof(BoxDev) * nb,
                           hipMemcpyHostToDevice));
    CUCO_HIP(hipMemcpy(e->bs.maps, e->hmaps.data(), sizeof(uint32_t) * e->hmaps.size(),
                       hipMemcpyHostToDevice));
    return CUCO_OK;
}

// Twiddles and Green's function factors of every box, concatenated in box
// order (box b at BoxDev::tab0).
template <typename T>
void fill_pm_tables(int M, double asmth, double L, Cx<T> *tw, T *ge, T *gs, T *gp)
{
    const double asmth2 = std::pow(2 * M_PI * asmth / L, 2);
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
        gp[k] = (T)(exp(-kk * kk * asmth2) * (ff * ff * ff * ff));
    }
}

template <typename T>
int upload_pm_tables(const std::vector<int> &Ms, const std::vector<double> &asmth,
                     const std::vector<double> &L, PmTables &tb)
{
    size_t total = 0;
    for (int M : Ms) total += M;
    std::vector<Cx<T>> tw(total);
    std::vector<T> ge(total), gs(total), gp(total);
    size_t o = 0;
    for (size_t b = 0; b < Ms.size(); ++b) {
        fill_pm_tables<T>(Ms[b], asmth[b], L[b], tw.data() + o, ge.data() + o, gs.data() + o,
                          gp.data() + o);
        o += Ms[b];
    }
    Cx<T> *dtw = nullptr;
    T *dge = nullptr, *dgs = nullptr, *dgp = nullptr;
    TRY(dev_alloc(dtw, sizeof(Cx<T>) * total));
    tb.tw = dtw;
    TRY(dev_alloc(dge, sizeof(T) * total));
    tb.gexp = dge;
    TRY(dev_alloc(dgs, sizeof(T) * total));
    tb.gsinc = dgs;
    TRY(dev_alloc(dgp, sizeof(T) * total));
    tb.gprod = dgp;
    CUCO_HIP(hipMemcpy(dtw, tw.data(), sizeof(Cx<T>) * total, hipMemcpyHostToDevice));
    CUCO_HIP(hipMemcpy(dge, ge.data(), sizeof(T) * total, hipMemcpyHostToDevice));
    CUCO_HIP(hipMemcpy(dgs, gs.data(), sizeof(T) * total, hipMemcpyHostToDevice));
    CUCO_HIP(hipMemcpy(dgp, gp.data(), sizeof(T) * total, hipMemcpyHostToDevice));
    return CUCO_OK;
}

void free_pm_tables(PmTables &tb)
{
    free_ptr(tb.tw);
    free_ptr(tb.gexp);
    free_ptr(tb.gsinc);
    free_ptr(tb.gprod);
}

template <typename T>
int make_pm_tables(cuco_engine *e, PmTables &tb)
{
    std::vector<int> Ms;
    std::vector<double> as, Ls;
    for (const HostBox &b : e->box) {
        Ms.push_back(b.prm.pm_grid);
        as.push_back(b.prm.asmth);
        Ls.push_back(b.prm.box_size);
    }
    return upload_pm_tables<T>(Ms, as, Ls, tb);
}

template <typename T>
int ensure_pm(cuco_engine *e)
{
    const int k = sizeof(T) == 8 ? 0 : 1;
    if (e->cgrid[k]) return CUCO_OK;
    TRY(dev_alloc(e->rho[k], sizeof(T) * e->grid_elems));
    TRY(dev_alloc(e->phi[k], sizeof(T) * e->grid_elems));
    TRY(dev_alloc(e->cgrid[k], sizeof(Cx<T>) * e->cgrid_elems));
    TRY(dev_alloc(e->dep_scratch[k], pm_deposit_scratch_bytes(e->bs, k == 0)));
    if (k == 1) TRY(pm_deposit_scratch_init(e->bs, e->dep_scratch[k], e->stream));
    return make_pm_tables<T>(e, e->tb[k]);
}

// ------------------------------------------------------------ sort
int sort_particles(cuco_engine *e, int mode)
{
    if (e->sorted_mode == mode) return CUCO_OK;
    TRY(stage_start(e, CUCO_STAGE_KEYS));
    TRY(launch_calc_keys(e->pos, e->bs, mode, e->keys, e->stream));
    TRY(stage_stop(e, CUCO_STAGE_KEYS));
    TRY(stage_start(e, CUCO_STAGE_SORT));
    TRY(radix_sort_pairs(e->keys, e->vals, true, e->n, e->key_bits[mode], e->sc,
                         e->stream, &e->skeys, &e->perm));
    TRY(stage_stop(e, CUCO_STAGE_SORT));
    TRY(stage_start(e, CUCO_STAGE_CELLS));
    TRY(launch_cell_offsets(e->skeys, e->n, e->nkeys[mode], e->off, e->runs,
                            e->run_cap, e->nruns, e->stream));
    if (mode == kKeyXMajor) {
        TRY(launch_reorder64(e->pos, e->perm, e->n, e->spos64, e->stream));
    } else {
        TRY(launch_reorder32(e->pos, e->perm, e->skeys, e->bs, mode, e->spos32, e->stream));
        TRY(launch_build_tiles(e->off, e->bs, e->nkeys[mode], e->tiles, e->stream));
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

int do_pp(cuco_engine *e, int precision, unsigned long long *counter,
          const int32_t *active_end = nullptr, int32_t ti_current = 0)
{
    if (precision != CUCO_FP64) {
        // The FP32 pair kernel packs a candidate cell's offset from its
        // tile's anchor cell into 10 bits per axis (|q| <= 511): rcut in
        // cells plus a Morton super-block (<= 32 cells) plus margin.
        for (const HostBox &h : e->box) {
            const double reach = h.pp.rcut * h.pp.fac + 36.0;
            if (reach > 511.0) {
                set_error("FP32 pair force: rcut of %.1f cells exceeds the FP32 kernel's "
                          "reach of 475 cells (use CUCO_FP64)", h.pp.rcut * h.pp.fac);
                return CUCO_ERR_UNSUPPORTED;
            }
        }
    }
    TRY(sort_particles(e, mode_of(precision)));
    TRY(stage_start(e, CUCO_STAGE_PP));
    if (precision == CUCO_FP64) {
        for (size_t b = 0; b < e->box.size(); ++b) {
            const BoxDev &d = e->hdev[kKeyXMajor][b];
            TRY(launch_pp64(e->spos64, e->perm, e->off + d.key_base, d.p0, d.n,
                            e->box[b].pp, e->table, e->acc_pp64, counter,
                            active_end, ti_current, e->stream));
        }
    } else {
        e->psym.n_sorted = e->n;
        e->atiles.t.block_shift = e->tiles.block_shift;
        TRY(launch_pp32(e->spos32, e->skeys, e->perm, e->off, e->bs, e->p32,
                        e->table_cut, e->uniform_mass, e->table32, e->tiles,
                        e->acc_pp32, counter, e->pp32_workgroups, active_end,
                        ti_current, e->psym, e->atiles.scan ? &e->atiles : nullptr,
                        e->n, e->stream));
    }
    TRY(stage_stop(e, CUCO_STAGE_PP));
    e->pp_prec = precision;
    return CUCO_OK;
}

template <typename T>
int do_pm_t(cuco_engine *e, int precision, hipStream_t st)
{
    TRY(ensure_pm<T>(e));
    const int mode = mode_of(precision);
    TRY(sort_particles(e, mode));
    const int k = sizeof(T) == 8 ? 0 : 1;
    const void *spos = precision == CUCO_FP64 ? (const void *)e->spos64 : (const void *)e->spos32;
    void *acc = precision == CUCO_FP64 ? (void *)e->acc_pm64 : (void *)e->acc_pm32;
    T *rho = static_cast<T *>(e->rho[k]);
    T *phi = static_cast<T *>(e->phi[k]);
    Cx<T> *c = static_cast<Cx<T> *>(e->cgrid[k]);
    TRY(stage_start(e, CUCO_STAGE_DEPOSIT, st));
    TRY(pm_deposit<T>(spos, e->skeys, e->off, e->bs, mode, rho, e->dep_scratch[k], st));
    TRY(stage_stop(e, CUCO_STAGE_DEPOSIT, st));
    TRY(stage_start(e, CUCO_STAGE_FFT_FWD, st));
    TRY(pm_fft_forward<T>(rho, c, e->bs, e->tb[k], st));
    TRY(stage_stop(e, CUCO_STAGE_FFT_FWD, st));
    TRY(stage_start(e, CUCO_STAGE_GREEN, st));
    TRY(pm_green_x<T>(c, e->bs, e->tb[k], st));
    TRY(stage_stop(e, CUCO_STAGE_GREEN, st));
    TRY(stage_start(e, CUCO_STAGE_FFT_INV, st));
    TRY(pm_fft_inverse<T>(c, phi, e->bs, e->tb[k], st));
    TRY(stage_stop(e, CUCO_STAGE_FFT_INV, st));
    TRY(stage_start(e, CUCO_STAGE_INTERP, st));
    TRY(pm_interp<T>(phi, spos, e->skeys, e->perm, e->off, e->bs, mode, e->dep_scratch[k],
                     acc, st));
    TRY(stage_stop(e, CUCO_STAGE_INTERP, st));
    e->pm_prec = precision;
    return CUCO_OK;
}

int do_pm(cuco_engine *e, int precision, hipStream_t st = nullptr)
{
    if (!st) st = e->stream;
    return precision == CUCO_FP64 ? do_pm_t<double>(e, precision, st)
                                  : do_pm_t<float>(e, precision, st);
}

// accel.c:14-22: PM on PM steps, then PP.  With pm_overlap the PM chain is
// issued first on the second stream (it waits for the sort) and the pair
// force on the main stream; the main stream then waits for the PM chain.
// The two touch disjoint outputs and only read the sorted arrays.
int do_forces(cuco_engine *e, int precision, int do_pm_flag,
              const int32_t *active_end = nullptr, int32_t ti_current = 0)
{
    if (!do_pm_flag) return do_pp(e, precision, nullptr, active_end, ti_current);
    if (!e->pm_overlap) {
        TRY(do_pm(e, precision));
        return do_pp(e, precision, nullptr, active_end, ti_current);
    }
    TRY(precision == CUCO_FP64 ? ensure_pm<double>(e) : ensure_pm<float>(e));
    TRY(sort_particles(e, mode_of(precision)));
    CUCO_HIP(hipEventRecord(e->ev_fork, e->stream));
    CUCO_HIP(hipStreamWaitEvent(e->stream2, e->ev_fork, 0));
    int rc = do_pm(e, precision, e->stream2);
    if (!rc) rc = do_pp(e, precision, nullptr, active_end, ti_current);
    // Join even after an error so a capture never ends with a dangling fork.
    hipError_t a = hipEventRecord(e->ev_join, e->stream2);
    hipError_t b = hipStreamWaitEvent(e->stream, e->ev_join, 0);
    if (rc) return rc;
    if (a != hipSuccess || b != hipSuccess) {
        set_error("stream join failed: %s", hipGetErrorString(a != hipSuccess ? a : b));
        return CUCO_ERR_HIP;
    }
    return CUCO_OK;
}

int kick_current(cuco_engine *e, double kpp, double kpm)
{
    if (e->pp_prec < 0 && e->pm_prec < 0) {
        set_error("kick before any force evaluation");
        return CUCO_ERR_STATE;
    }
    const bool pp32 = e->pp_prec == CUCO_FP32, pm32 = e->pm_prec == CUCO_FP32;
    if (e->pp_prec >= 0 && e->pm_prec >= 0 && pp32 != pm32) {
        // Mixed precisions: two kicks, PP first (timestep.c:56-59 then :86-92).
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

// A one-box set for the standalone 3D FFTs (unit box, no particles).
struct FftBox {
    BoxSet bs;
    PmTables tb;
    ~FftBox()
    {
        free_ptr(bs.dev[0]);
        free_ptr(bs.maps);
        free_pm_tables(tb);
    }
};

template <typename T>
int make_fft_box(int M, FftBox &f)
{
    BoxDev d{};
    d.g.M = M;
    d.g.L = 1.0;
    d.g.fac = M;
    d.g.cell = 1.0 / M;
    d.half = make_plan(M / 2);
    d.full = make_plan(M);
    f.bs.nboxes = 1;
    std::vector<uint32_t> maps(kMapCount * 2, 0u);
    for (int map = kMapGatherX; map < kMapCount; ++map) maps[map * 2 + 1] = pm_blocks(M, map);
    for (int map = 0; map < kMapCount; ++map) f.bs.total[map] = maps[map * 2 + 1];
    f.bs.fft32_mask = fft32_mask_bit(M);
    TRY(dev_alloc(f.bs.dev[0], sizeof(BoxDev)));
    f.bs.dev[1] = f.bs.dev[0];
    TRY(dev_alloc(f.bs.maps, sizeof(uint32_t) * maps.size()));
    CUCO_HIP(hipMemcpy(f.bs.dev[0], &d, sizeof(BoxDev), hipMemcpyHostToDevice));
    CUCO_HIP(hipMemcpy(f.bs.maps, maps.data(), sizeof(uint32_t) * maps.size(), hipMemcpyHostToDevice));
    return upload_pm_tables<T>({M}, {1.0}, {1.0}, f.tb);
}

template <typename T>
int fft3d_single_t(int M, bool inverse, void *d_in, void *d_out)
{
    FftBox f;
    TRY(make_fft_box<T>(M, f));
    if (!inverse) {
        TRY(pm_fft_forward<T>(static_cast<const T *>(d_in), static_cast<Cx<T> *>(d_out), f.bs, f.tb, nullptr));
        TRY(pm_fft_x<T>(static_cast<Cx<T> *>(d_out), f.bs, false, f.tb, nullptr));
    } else {
        TRY(pm_fft_x<T>(static_cast<Cx<T> *>(d_in), f.bs, true, f.tb, nullptr));
        TRY(pm_fft_inverse<T>(static_cast<Cx<T> *>(d_in), static_cast<T *>(d_out), f.bs, f.tb, nullptr));
    }
    if (hipDeviceSynchronize() != hipSuccess) {
        set_error("FFT failed on the device");
        return CUCO_ERR_HIP;
    }
    return CUCO_OK;
}

int fft3d_single(int M, int precision, bool inverse, void *d_in, void *d_out)
{
    TRY(check_prec(precision));
    if (!fft_size_supported(M) || !d_in || !d_out) {
        set_error("unsupported FFT size %d or null pointer", M);
        return CUCO_ERR_UNSUPPORTED;
    }
    return precision == CUCO_FP64 ? fft3d_single_t<double>(M, inverse, d_in, d_out)
                                  : fft3d_single_t<float>(M, inverse, d_in, d_out);
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

}  // extern "C"

namespace {

int check_params(const cuco_params *params)
{
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
    return CUCO_OK;
}

void set_grids(HostBox &b)
{
    const int M = b.prm.pm_grid;
    GridInfo &gx = b.g[kKeyXMajor];
    gx.M = M;
    gx.key_mode = kKeyXMajor;
    gx.bits = ceil_log2((uint64_t)M);
    gx.key_bits = ceil_log2((uint64_t)M * M * M);
    gx.nkeys = (uint32_t)((uint64_t)M * M * M);
    gx.L = b.prm.box_size;
    gx.fac = M / b.prm.box_size;
    gx.cell = b.prm.box_size / M;
    GridInfo &gm = b.g[kKeyMorton];
    gm = gx;
    gm.key_mode = kKeyMorton;
    gm.bits = ceil_log2((uint64_t)M);
    gm.key_bits = 3 * gm.bits;
    gm.nkeys = (uint32_t)(1ull << gm.key_bits);
}

int create_engine(const cuco_params *params, const uint32_t *caps, int nboxes,
                  bool ensemble, int32_t device, cuco_engine **out)
{
    *out = nullptr;
    uint64_t cap_total = 0, waves = 0, keys[2] = {0, 0};
    for (int b = 0; b < nboxes; ++b) {
        TRY(check_params(&params[b]));
        cap_total += caps[b];
        waves += (caps[b] + 63) / 64;
    }
    if (cap_total == 0 || cap_total > 0x7fffffffULL) {
        set_error("total particle count %llu out of range", (unsigned long long)cap_total);
        return CUCO_ERR_INVALID_ARG;
    }
    CUCO_HIP(hipSetDevice(device));
    cuco_engine *e = new cuco_engine();
    e->device = device;
    e->ensemble = ensemble;
    e->box.resize(nboxes);
    e->min_bits = 64;
    for (int b = 0; b < nboxes; ++b) {
        HostBox &h = e->box[b];
        h.prm = params[b];
        h.cap = caps[b];
        h.n = ensemble ? caps[b] : 0;
        set_pair_params(h);
        set_grids(h);
        for (int m = 0; m < 2; ++m) keys[m] += h.g[m].nkeys;
        e->min_bits = std::min(e->min_bits, h.g[kKeyMorton].bits);
        const uint64_t M = h.prm.pm_grid;
        e->grid_elems += M * M * M;
        e->cgrid_elems += M * M * (M / 2 + 1);
        e->tab_elems += M;
    }
    if (keys[0] >= 0xffffffffULL || keys[1] >= 0xffffffffULL) {
        delete e;
        set_error("the boxes' cell keys exceed 32 bits");
        return CUCO_ERR_UNSUPPORTED;
    }
    for (int m = 0; m < 2; ++m) {
        e->nkeys[m] = (uint32_t)keys[m];
        e->key_bits[m] = ceil_log2(keys[m]);
    }
    e->capacity = (uint32_t)cap_total;
    e->bs.nboxes = nboxes;
    e->bs.capacity = e->capacity;
    e->bs.wave_slots = (uint32_t)waves;

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
    const size_t offw = (size_t)std::max(e->nkeys[0], e->nkeys[1]) + 1;
    ALLOC(e->off, sizeof(uint32_t) * offw);
    e->run_cap = cell_run_capacity((uint32_t)(offw - 1));
    ALLOC(e->runs, sizeof(uint4) * e->run_cap);
    ALLOC(e->nruns, sizeof(uint32_t));
    ALLOC(e->spos64, sizeof(double4) * cap);
    ALLOC(e->spos32, sizeof(float4) * cap);
    ALLOC(e->table, sizeof(float) * kNTab);
    ALLOC(e->table32, sizeof(float) * kTabPadded);
    ALLOC(e->dcount, sizeof(unsigned long long) * 2);
    ALLOC(e->p32, sizeof(P32) * nboxes);
    ALLOC(e->bs.dev[0], sizeof(BoxDev) * nboxes);
    ALLOC(e->bs.dev[1], sizeof(BoxDev) * nboxes);
    ALLOC(e->bs.maps, sizeof(uint32_t) * kMapCount * (nboxes + 1));
    ALLOC(e->bs.tile_map, sizeof(uint32_t) * (nboxes + 1));
    // FP32 pair force evaluates each pair once by default (40 bytes per
    // particle of own and fixed-point j sums); cuco_set_tuning can switch to
    // the kernel that visits every ordered pair.
    ALLOC(e->psym.iacc, sizeof(float4) * cap);
    ALLOC(e->psym.jacc, sizeof(unsigned long long) * 3 * cap);
    e->psym.on = true;
    // Tile arrays sized for the smallest tunable Morton block (most tiles).
    // A Morton block never straddles two boxes: every box's key range is a
    // multiple of 8^bits with bits > shift.
    e->tiles.block_shift = e->min_bits >= 4 ? 3 : e->min_bits - 1;
    e->min_block_shift = std::min(2, e->tiles.block_shift);
    e->tiles.nblocks = e->nkeys[kKeyMorton] >> (3 * e->min_block_shift);
    e->tiles.max_tiles = (e->capacity + 63) / 64 + e->tiles.nblocks;
    ALLOC(e->tiles.tile_start, sizeof(uint32_t) * e->tiles.max_tiles);
    ALLOC(e->tiles.tile_count, sizeof(uint32_t) * e->tiles.max_tiles);
    ALLOC(e->tiles.block_tiles, sizeof(uint32_t) * e->tiles.nblocks);
    ALLOC(e->tiles.counters, sizeof(uint32_t) * 4);
    ALLOC(e->tiles.scan_scratch, sizeof(uint32_t) * scan_scratch_words(e->tiles.nblocks));
#undef ALLOC
    if (rc == CUCO_OK) {
        // Both streams keep the default priority.
        hipError_t err = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
        if (err == hipSuccess) err = hipStreamCreateWithFlags(&e->stream2, hipStreamNonBlocking);
        if (err == hipSuccess) err = hipEventCreateWithFlags(&e->ev_fork, hipEventDisableTiming);
        if (err == hipSuccess) err = hipEventCreateWithFlags(&e->ev_join, hipEventDisableTiming);
        if (err != hipSuccess) {
            set_error("stream / event creation failed: %s", hipGetErrorString(err));
            rc = CUCO_ERR_HIP;
        }
    }
    if (rc == CUCO_OK) rc = update_pair_constants(e);
    if (rc == CUCO_OK) rc = update_boxes(e);
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

}  // namespace

extern "C" {

int cuco_create(const cuco_params *params, int64_t max_particles, int32_t device,
                cuco_engine **out)
try {
    if (!params || !out) {
        set_error("null argument");
        return CUCO_ERR_INVALID_ARG;
    }
    *out = nullptr;
    if (max_particles <= 0 || max_particles > 0x7fffffffLL) {
        set_error("max_particles %lld out of range", (long long)max_particles);
        return CUCO_ERR_INVALID_ARG;
    }
    const uint32_t cap = (uint32_t)max_particles;
    return create_engine(params, &cap, 1, false, device, out);
} CUCO_CATCH

int cuco_ensemble_create(const cuco_params *params, const int64_t *n_per_box,
                         int32_t nboxes, int32_t device, cuco_engine **out)
try {
    if (!params || !n_per_box || !out) {
        set_error("null argument");
        return CUCO_ERR_INVALID_ARG;
    }
    *out = nullptr;
    if (nboxes < 1 || nboxes > 65536) {
        set_error("nboxes %d outside 1..65536", nboxes);
        return CUCO_ERR_INVALID_ARG;
    }
    std::vector<uint32_t> caps(nboxes);
    for (int b = 0; b < nboxes; ++b) {
        if (n_per_box[b] <= 0 || n_per_box[b] > 0x7fffffffLL) {
            set_error("box %d: particle count %lld out of range", b, (long long)n_per_box[b]);
            return CUCO_ERR_INVALID_ARG;
        }
        caps[b] = (uint32_t)n_per_box[b];
    }
    return create_engine(params, caps.data(), nboxes, true, device, out);
} CUCO_CATCH

int cuco_ensemble_size(cuco_engine *e, int32_t *nboxes, int64_t *n_per_box)
{
    CHECK_ENGINE(e);
    if (nboxes) *nboxes = (int32_t)e->box.size();
    if (n_per_box)
        for (size_t b = 0; b < e->box.size(); ++b)
            n_per_box[b] = e->ensemble ? e->box[b].cap : e->box[b].n;
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
    free_ptr(e->off); free_ptr(e->runs); free_ptr(e->nruns);
    free_ptr(e->spos64); free_ptr(e->spos32);
    free_ptr(e->table); free_ptr(e->table32); free_ptr(e->dcount);
    free_ptr(e->p32);
    free_ptr(e->psym.iacc); free_ptr(e->psym.jacc);
    free_ptr(e->bs.dev[0]); free_ptr(e->bs.dev[1]);
    free_ptr(e->bs.maps); free_ptr(e->bs.tile_map);
    free_ptr(e->tiles.tile_start); free_ptr(e->tiles.tile_count);
    free_ptr(e->tiles.block_tiles); free_ptr(e->tiles.counters);
    free_ptr(e->tiles.scan_scratch);
    free_ptr(e->atiles.scan); free_ptr(e->atiles.nscratch); free_ptr(e->atiles.list);
    free_ptr(e->atiles.tile_map); free_ptr(e->atiles.t.tile_start);
    free_ptr(e->atiles.t.tile_count); free_ptr(e->atiles.t.block_tiles);
    free_ptr(e->atiles.t.counters); free_ptr(e->atiles.t.scan_scratch);
    for (int k = 0; k < 2; ++k) {
        free_ptr(e->rho[k]); free_ptr(e->phi[k]); free_ptr(e->cgrid[k]);
        free_ptr(e->dep_scratch[k]);
        free_pm_tables(e->tb[k]);
    }
    free_ptr(e->dfactors); free_ptr(e->dstep);
    if (e->stream2) (void)hipStreamSynchronize(e->stream2);
    free_ptr(e->ti_beg); free_ptr(e->ti_end);
    free_ptr(e->ti_scalars); free_ptr(e->tl_work); free_ptr(e->kick_table);
    free_ptr(e->drift_table);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    if (e->stream2) (void)hipStreamDestroy(e->stream2);
    if (e->ev_fork) (void)hipEventDestroy(e->ev_fork);
    if (e->ev_join) (void)hipEventDestroy(e->ev_join);
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
    for (HostBox &b : e->box) {
        b.prm.soft_h = soft_h;
        set_pair_params(b);
    }
    destroy_graph(e);
    return update_pair_constants(e);
}

int cuco_set_params(cuco_engine *e, int32_t box, const cuco_params *params)
{
    CHECK_ENGINE(e);
    if (!params || box < -1 || box >= (int)e->box.size()) {
        set_error("null params or box %d outside -1..%d", box, (int)e->box.size() - 1);
        return CUCO_ERR_INVALID_ARG;
    }
    TRY(check_params(params));
    const int b0 = box < 0 ? 0 : box, b1 = box < 0 ? (int)e->box.size() : box + 1;
    for (int b = b0; b < b1; ++b) {
        const cuco_params &cur = e->box[b].prm;
        if (params->pm_grid != cur.pm_grid || params->box_size != cur.box_size) {
            set_error("box %d: pm_grid and box_size are fixed when the handle is created", b);
            return CUCO_ERR_INVALID_ARG;
        }
    }
    const bool asmth_changed = [&] {
        for (int b = b0; b < b1; ++b)
            if (e->box[b].prm.asmth != params->asmth) return true;
        return false;
    }();
    for (int b = b0; b < b1; ++b) {
        e->box[b].prm = *params;
        set_pair_params(e->box[b]);
    }
    if (asmth_changed) {
        // The Green's function tables depend on asmth: rebuild them.
        for (int k = 0; k < 2; ++k) {
            if (!e->cgrid[k]) continue;
            free_pm_tables(e->tb[k]);
            TRY(k == 0 ? make_pm_tables<double>(e, e->tb[0]) : make_pm_tables<float>(e, e->tb[1]));
        }
    }
    TRY(update_boxes(e));
    destroy_graph(e);
    e->pp_prec = e->pm_prec = -1;
    return update_pair_constants(e);
}

}  // extern "C"

namespace {

// Checks a particle count against the handle: [1, capacity], and exactly the
// capacity for an ensemble.
int check_count(const cuco_engine *e, int64_t n)
{
    if (e->ensemble && n != (int64_t)e->capacity) {
        set_error("ensemble needs exactly %u particles (got %lld)", e->capacity, (long long)n);
        return CUCO_ERR_INVALID_ARG;
    }
    if (n <= 0 || n > (int64_t)e->capacity) {
        set_error("n=%lld outside [1, %u]", (long long)n, e->capacity);
        return CUCO_ERR_INVALID_ARG;
    }
    return CUCO_OK;
}

// keep_vel: positions (and masses) only.  When the handle already holds
// these n particles their device velocities and integer-timeline state stay;
// otherwise the velocities start at zero.
int upload_impl(cuco_engine *e, int64_t n, const double *pos, const double *vel,
                const double *mass, double uniform_mass, bool keep_vel)
{
    if (e->ensemble) {
        if (n != (int64_t)e->capacity || !pos) {
            set_error("ensemble upload needs exactly %u particles (got %lld) and positions",
                      e->capacity, (long long)n);
            return CUCO_ERR_INVALID_ARG;
        }
    } else if (n <= 0 || n > (int64_t)e->capacity || !pos) {
        set_error("n=%lld outside [1, %u] or null positions", (long long)n, e->capacity);
        return CUCO_ERR_INVALID_ARG;
    }
    std::vector<double4> hp((size_t)n), hv((size_t)n);
    double common = mass ? mass[0] : uniform_mass;
    bool mass_max_changed = false;
    int64_t i = 0;
    for (size_t b = 0; b < e->box.size(); ++b) {
        const double L = e->box[b].prm.box_size;
        const int64_t nb = e->ensemble ? (int64_t)e->box[b].cap : n;
        double mmax = 0;
        for (int64_t k = 0; k < nb; ++k, ++i) {
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
            mmax = std::max(mmax, hp[i].w);
            hv[i] = vel ? make_double4(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2], 0.0)
                        : make_double4(0, 0, 0, 0);
        }
        if (mmax != e->box[b].mass_max) mass_max_changed = true;
        e->box[b].mass_max = mmax;
    }
    const bool keep = keep_vel && e->have_particles && (uint32_t)n == e->n;
    CUCO_HIP(hipMemcpyAsync(e->pos, hp.data(), sizeof(double4) * n, hipMemcpyHostToDevice, e->stream));
    if (!keep)
        CUCO_HIP(hipMemcpyAsync(e->vel, hv.data(), sizeof(double4) * n, hipMemcpyHostToDevice, e->stream));
    CUCO_HIP(hipStreamSynchronize(e->stream));
    if ((uint32_t)n != e->n) destroy_graph(e);
    e->n = (uint32_t)n;
    if (!e->ensemble) e->box[0].n = (uint32_t)n;
    const double um = common > 0 ? common : 0;
    const bool mass_changed = um != e->uniform_mass || mass_max_changed;
    e->uniform_mass = um;
    TRY(update_boxes(e));
    if (mass_changed) {
        destroy_graph(e);
        TRY(update_pair_constants(e));
    }
    e->have_particles = true;
    e->sorted_mode = -1;
    e->pp_prec = e->pm_prec = -1;
    if (!keep) e->tl_init = false;
    return CUCO_OK;
}

}  // namespace

extern "C" {

int cuco_upload_particles(cuco_engine *e, int64_t n, const double *pos,
                          const double *vel, const double *mass,
                          double uniform_mass)
try {
    CHECK_ENGINE(e);
    return upload_impl(e, n, pos, vel, mass, uniform_mass, false);
} CUCO_CATCH

int cuco_download_particles(cuco_engine *e, double *pos, double *vel)
try {
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
} CUCO_CATCH

int cuco_download_accel(cuco_engine *e, double *acc_pp, double *acc_pm)
try {
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
} CUCO_CATCH

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
    TRY(launch_drift(e->pos, e->vel, e->bs, drift_factor, e->stream));
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
            TRY(launch_drift(e->pos, e->vel, e->bs, factors[3 * s + 1], e->stream));
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
        if (!rc) rc = launch_drift_table(e->pos, e->vel, e->bs, e->dfactors, e->dstep, e->stream);
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

}  // extern "C"

namespace {

// get_drift_factor / get_gravkick_factor (driftfac.c:66-126) at two integer
// times; the device kick uses the same operations (integrate.hip).
double timeline_factor(const cuco_engine *e, const std::vector<double> &table,
                       int time0, int time1)
{
    auto at = [&](double u) {
        int i = (int)u;
        if (i >= kDriftTable) i = kDriftTable - 1;
        if (i < 1) return u * table[0];
        return table[i - 1] + (table[i] - table[i - 1]) * (u - i);
    };
    const double lb = e->tl_log_begin, lm = e->tl_log_max;
    const double a1 = lb + time0 * e->tl_interval;
    const double a2 = lb + time1 * e->tl_interval;
    const double u1 = (a1 - lb) / (lm - lb) * kDriftTable;
    const double u2 = (a2 - lb) / (lm - lb) * kDriftTable;
    return at(u2) - at(u1);
}

double timeline_time(const cuco_engine *e, int ti)
{
    return e->tl.time_begin * exp(ti * e->tl_interval);  // run.c:57
}

// Softening of gravlist.cpp:28-45 at scale factor a: comoving SofteningHalo
// until it exceeds SofteningHaloMaxPhys physically.
double softening_table(const cuco_timeline &t, double a)
{
    return t.softening_halo * a > t.softening_halo_max_phys
               ? t.softening_halo_max_phys / a
               : t.softening_halo;
}

int timeline_step(cuco_engine *e, int precision, int32_t ti_stop, bool &stopped)
{
    cuco_timeline_state &s = e->tls;
    stopped = false;
    // find_next_sync_point_and_drift (run.c:29-64)
    TRY(launch_ti_min(e->ti_end, e->n, e->ti_scalars, e->stream));
    int32_t h[2];
    CUCO_HIP(hipMemcpyAsync(h, e->ti_scalars, sizeof(h), hipMemcpyDeviceToHost, e->stream));
    CUCO_HIP(hipStreamSynchronize(e->stream));
    if (h[1]) {
        set_error("a timestep left the integer timeline (timestep.c:126-133)");
        return CUCO_ERR_STATE;
    }
    int32_t min_glob = h[0];
    if (min_glob >= s.pm_ti_endstep) min_glob = s.pm_ti_endstep;
    if (ti_stop > s.ti_current && min_glob >= ti_stop) {
        TRY(launch_drift(e->pos, e->vel, e->bs,
                         timeline_factor(e, e->drift_tab, s.ti_current, ti_stop), e->stream));
        e->sorted_mode = -1;
        s.ti_current = ti_stop;
        s.time = timeline_time(e, s.ti_current);
        stopped = true;
        return CUCO_OK;
    }
    TRY(launch_drift(e->pos, e->vel, e->bs,
                     timeline_factor(e, e->drift_tab, s.ti_current, min_glob), e->stream));
    e->sorted_mode = -1;
    s.ti_current = min_glob;
    s.time = timeline_time(e, s.ti_current);

    // Active particles of this step.
    TRY(launch_count_active(e->ti_end, e->n, s.ti_current, e->ti_scalars + 2, e->stream));

    // set_softenings (gravlist.cpp:28-45) at the new time.
    const double soft = softening_table(e->tl, s.time);
    for (HostBox &b : e->box) {
        b.prm.soft_h = 2.8 * soft;
        set_pair_params(b);
    }
    TRY(update_pair_constants(e));
    destroy_graph(e);

    // compute_accelerations (accel.c:8-23)
    const bool pm_step = s.pm_ti_endstep == s.ti_current;
    TRY(begin_eval(e));
    // FP32: when the active particles carry at least 4/5 of the estimated
    // pair work, the short-range force is evaluated for every particle, each
    // pair once, and only the active ones are kicked below (the reference
    // computes every particle's force on every step, accel.c:8-23); the
    // each-pair-once kernel costs about 4/5 of the one that visits every
    // ordered pair, whose active-set tiles take the other steps.
    bool whole = false;
    int rc = CUCO_OK;
    if (precision == CUCO_FP32 && e->psym.on) {
        rc = sort_particles(e, kKeyMorton);
        if (!rc) rc = launch_active_work(e->skeys, e->perm, e->off, e->n,
                                         3 * std::min(4, e->min_bits), e->ti_end,
                                         s.ti_current, e->tl_work, e->stream);
        unsigned long long w[2] = {0, 0};
        if (!rc && (hipMemcpyAsync(w, e->tl_work, sizeof(w), hipMemcpyDeviceToHost,
                                   e->stream) != hipSuccess ||
                    hipStreamSynchronize(e->stream) != hipSuccess)) {
            set_error("reading the pair-work estimate failed");
            rc = CUCO_ERR_HIP;
        }
        whole = (double)w[1] >= 0.8 * (double)w[0];
    }
    if (!rc) rc = do_forces(e, precision, pm_step, whole ? nullptr : e->ti_end, s.ti_current);
    end_eval(e);
    if (rc) return rc;

    // advance_and_find_timesteps (timestep.c:11-94)
    const double a = s.time;
    TimelineDev t{};
    t.ti_current = s.ti_current;
    t.log_begin = e->tl_log_begin;
    t.log_max = e->tl_log_max;
    t.interval = e->tl_interval;
    t.fac1 = 1 / (a * a);
    t.atime = a;
    double ha = e->tl.omega0 / (a * a * a) + (1 - e->tl.omega0 - e->tl.omega_lambda) / (a * a) +
                e->tl.omega_lambda;
    t.hubble_a = e->tl.hubble * sqrt(ha);
    t.err_tol = e->tl.err_tol_int_accuracy;
    t.soft_table = soft;
    t.dt_displacement = e->tl.max_size_timestep;
    t.min_dt = e->tl.min_size_timestep;
    const int fp32 = precision == CUCO_FP32;
    const void *app = fp32 ? (const void *)e->acc_pp32 : (const void *)e->acc_pp64;
    const void *apm = fp32 ? (const void *)e->acc_pm32 : (const void *)e->acc_pm64;
    TRY(launch_advance(e->vel, app, apm, fp32, e->ti_beg, e->ti_end, e->n, t,
                       e->kick_table, e->ti_scalars + 1, e->stream));
    if (pm_step) {
        int ti_step = kTimebase;
        while (ti_step > (e->tl.max_size_timestep / e->tl_interval)) ti_step >>= 1;
        if (ti_step > (s.pm_ti_endstep - s.pm_ti_begstep)) {
            if (((kTimebase - s.pm_ti_endstep) % ti_step) > 0)
                ti_step = s.pm_ti_endstep - s.pm_ti_begstep;
        }
        if (s.ti_current == kTimebase) ti_step = 0;
        const int tstart = (s.pm_ti_begstep + s.pm_ti_endstep) / 2;
        const int tend = s.pm_ti_endstep + ti_step / 2;
        const double k = timeline_factor(e, e->kick_tab, tstart, tend);
        s.pm_ti_begstep = s.pm_ti_endstep;
        s.pm_ti_endstep = s.pm_ti_begstep + ti_step;
        if (fp32) TRY(launch_kick32(e->vel, nullptr, e->acc_pm32, e->n, 0, k, e->stream));
        else TRY(launch_kick64(e->vel, nullptr, e->acc_pm64, e->n, 0, k, e->stream));
    }
    int32_t cnt = 0;
    CUCO_HIP(hipMemcpyAsync(&cnt, e->ti_scalars + 2, sizeof(cnt), hipMemcpyDeviceToHost, e->stream));
    CUCO_HIP(hipStreamSynchronize(e->stream));
    s.active = cnt;
    s.steps++;
    return CUCO_OK;
}

bool timeline_params_ok(const cuco_timeline *t)
{
    return t && t->time_begin > 0 && t->time_max > t->time_begin && t->hubble > 0 &&
           t->err_tol_int_accuracy > 0 && t->max_size_timestep > 0 &&
           t->min_size_timestep >= 0 && t->softening_halo > 0 &&
           t->softening_halo_max_phys > 0;
}

// The parts of init that do not touch the particles' timesteps: parameter
// checks, buffers, the drift and kick tables and the timebase.
int timeline_setup(cuco_engine *e, const cuco_timeline *t)
{
    if (!timeline_params_ok(t)) {
        set_error("timeline parameters out of range");
        return CUCO_ERR_INVALID_ARG;
    }
    if (!e->ti_beg) {
        TRY(dev_alloc(e->ti_beg, sizeof(int32_t) * e->capacity));
        TRY(dev_alloc(e->ti_end, sizeof(int32_t) * e->capacity));
        TRY(dev_alloc(e->ti_scalars, sizeof(int32_t) * 4));
        TRY(dev_alloc(e->tl_work, sizeof(unsigned long long) * 2));
        TRY(dev_alloc(e->kick_table, sizeof(double) * kDriftTable));
        TRY(dev_alloc(e->drift_table, sizeof(double) * kDriftTable));
    }
    if (!e->atiles.scan) {
        // Tiles of the active targets (FP32 active-set pair force).
        ActiveTiles &a = e->atiles;
        a.capacity = e->capacity;
        a.nkeys = e->nkeys[kKeyMorton];
        a.t.nblocks = e->tiles.nblocks;
        a.t.max_tiles = e->tiles.max_tiles;
        TRY(dev_alloc(a.scan, sizeof(uint32_t) * ((size_t)e->capacity + 1)));
        TRY(dev_alloc(a.nscratch, sizeof(uint32_t) * scan_scratch_words(e->capacity)));
        TRY(dev_alloc(a.list, sizeof(uint32_t) * (size_t)e->capacity));
        TRY(dev_alloc(a.tile_map, sizeof(uint32_t) * (e->box.size() + 1)));
        TRY(dev_alloc(a.t.tile_start, sizeof(uint32_t) * a.t.max_tiles));
        TRY(dev_alloc(a.t.tile_count, sizeof(uint32_t) * a.t.max_tiles));
        TRY(dev_alloc(a.t.block_tiles, sizeof(uint32_t) * a.t.nblocks));
        TRY(dev_alloc(a.t.counters, sizeof(uint32_t) * 4));
        TRY(dev_alloc(a.t.scan_scratch, sizeof(uint32_t) * scan_scratch_words(a.t.nblocks)));
    }
    e->tl = *t;
    // The tables are built on the device; the kick table stays there for
    // the per-particle kicks, and host copies serve the drift factors.
    e->drift_tab.assign(kDriftTable, 0.0);
    e->kick_tab.assign(kDriftTable, 0.0);
    TRY(launch_drift_tables(t->time_begin, t->time_max, t->omega0, t->omega_lambda, t->hubble,
                            e->drift_table, e->kick_table, e->stream));
    CUCO_HIP(hipMemcpyAsync(e->drift_tab.data(), e->drift_table, sizeof(double) * kDriftTable,
                            hipMemcpyDeviceToHost, e->stream));
    CUCO_HIP(hipMemcpyAsync(e->kick_tab.data(), e->kick_table, sizeof(double) * kDriftTable,
                            hipMemcpyDeviceToHost, e->stream));
    e->tl_log_begin = log(t->time_begin);
    e->tl_log_max = log(t->time_max);
    e->tl_interval = (log(t->time_max) - log(t->time_begin)) / kTimebase;  // init.c:23
    return CUCO_OK;
}

}  // namespace

extern "C" {

int cuco_timeline_init(cuco_engine *e, const cuco_timeline *t)
try {
    CHECK_ENGINE(e);
    TRY(check_ready(e));
    TRY(timeline_setup(e, t));
    TRY(launch_fill_u32(reinterpret_cast<uint32_t *>(e->ti_beg), 0u, e->n, e->stream));
    TRY(launch_fill_u32(reinterpret_cast<uint32_t *>(e->ti_end), 0u, e->n, e->stream));
    TRY(launch_fill_u32(reinterpret_cast<uint32_t *>(e->ti_scalars), 0u, 4, e->stream));
    CUCO_HIP(hipStreamSynchronize(e->stream));
    cuco_timeline_state &s = e->tls;
    s = cuco_timeline_state{};
    s.time = t->time_begin;
    s.timebase_interval = e->tl_interval;
    e->tl_init = true;
    return CUCO_OK;
} CUCO_CATCH

int cuco_timeline_run(cuco_engine *e, int32_t precision, int32_t max_steps,
                      int32_t ti_stop, cuco_timeline_state *state)
{
    CHECK_ENGINE(e);
    TRY(check_prec(precision));
    TRY(check_ready(e));
    if (!e->tl_init) {
        set_error("cuco_timeline_init has not been called");
        return CUCO_ERR_STATE;
    }
    cuco_timeline_state &s = e->tls;
    s.at_stop = 0;
    for (int k = 0; k < max_steps && !s.finished; ++k) {
        bool stopped = false;
        TRY(timeline_step(e, precision, ti_stop, stopped));
        if (stopped) {
            s.at_stop = 1;
            break;
        }
        // run.c:24
        if (!(s.ti_current < kTimebase && s.time <= e->tl.time_max)) s.finished = 1;
    }
    int32_t flag = 0;
    CUCO_HIP(hipMemcpyAsync(&flag, e->ti_scalars + 1, sizeof(flag), hipMemcpyDeviceToHost, e->stream));
    CUCO_HIP(hipStreamSynchronize(e->stream));
    if (state) *state = s;
    if (flag) {
        set_error("a timestep left the integer timeline (timestep.c:126-133)");
        return CUCO_ERR_STATE;
    }
    return CUCO_OK;
}

int cuco_get_timesteps(cuco_engine *e, int32_t *ti_begstep, int32_t *ti_endstep)
{
    CHECK_ENGINE(e);
    if (!e->tl_init) {
        set_error("cuco_timeline_init has not been called");
        return CUCO_ERR_STATE;
    }
    if (ti_begstep)
        CUCO_HIP(hipMemcpyAsync(ti_begstep, e->ti_beg, sizeof(int32_t) * e->n,
                                hipMemcpyDeviceToHost, e->stream));
    if (ti_endstep)
        CUCO_HIP(hipMemcpyAsync(ti_endstep, e->ti_end, sizeof(int32_t) * e->n,
                                hipMemcpyDeviceToHost, e->stream));
    CUCO_HIP(hipStreamSynchronize(e->stream));
    return CUCO_OK;
}

namespace {

struct CkFile {
    FILE *f = nullptr;
    ~CkFile()
    {
        if (f) fclose(f);
    }
};

size_t accel_bytes(int32_t precision)
{
    return precision < 0 ? 0 : (precision == CUCO_FP64 ? sizeof(double4) : sizeof(float4));
}

void *accel_buffer(cuco_engine *e, bool pp, int32_t precision)
{
    if (precision == CUCO_FP64) return pp ? (void *)e->acc_pp64 : (void *)e->acc_pm64;
    return pp ? (void *)e->acc_pp32 : (void *)e->acc_pm32;
}

}  // namespace

int cuco_checkpoint_write(cuco_engine *e, const char *path)
try {
    CHECK_ENGINE(e);
    TRY(check_ready(e));
    if (!path) {
        set_error("null path");
        return CUCO_ERR_INVALID_ARG;
    }
    const size_t n = e->n;
    cuco_checkpoint_header h{};
    memcpy(h.magic, "CUCOCKPT", 8);
    h.version = 1;
    h.nboxes = (int32_t)e->box.size();
    h.n = (int64_t)n;
    h.timeline = e->tl_init ? 1 : 0;
    h.pp_precision = e->pp_prec;
    h.pm_precision = e->pm_prec;
    if (e->tl_init) {
        h.params = e->tl;
        h.state = e->tls;
    }
    std::vector<cuco_params> prm;
    std::vector<int64_t> counts;
    for (const HostBox &b : e->box) {
        prm.push_back(b.prm);
        counts.push_back(e->ensemble ? (int64_t)b.cap : (int64_t)n);
    }
    std::vector<double4> hp(n), hv(n);
    std::vector<int32_t> tb(e->tl_init ? n : 0), te(e->tl_init ? n : 0);
    std::vector<char> app(accel_bytes(e->pp_prec) * n), apm(accel_bytes(e->pm_prec) * n);
    CUCO_HIP(hipMemcpyAsync(hp.data(), e->pos, sizeof(double4) * n, hipMemcpyDeviceToHost,
                            e->stream));
    CUCO_HIP(hipMemcpyAsync(hv.data(), e->vel, sizeof(double4) * n, hipMemcpyDeviceToHost,
                            e->stream));
    if (e->tl_init) {
        CUCO_HIP(hipMemcpyAsync(tb.data(), e->ti_beg, sizeof(int32_t) * n,
                                hipMemcpyDeviceToHost, e->stream));
        CUCO_HIP(hipMemcpyAsync(te.data(), e->ti_end, sizeof(int32_t) * n,
                                hipMemcpyDeviceToHost, e->stream));
    }
    if (!app.empty())
        CUCO_HIP(hipMemcpyAsync(app.data(), accel_buffer(e, true, e->pp_prec), app.size(),
                                hipMemcpyDeviceToHost, e->stream));
    if (!apm.empty())
        CUCO_HIP(hipMemcpyAsync(apm.data(), accel_buffer(e, false, e->pm_prec), apm.size(),
                                hipMemcpyDeviceToHost, e->stream));
    CUCO_HIP(hipStreamSynchronize(e->stream));

    CkFile fp;
    fp.f = fopen(path, "wb");
    if (!fp.f) {
        set_error("cannot open checkpoint file %s for writing", path);
        return CUCO_ERR_INVALID_ARG;
    }
    auto put = [&](const void *p, size_t bytes) {
        return bytes == 0 || fwrite(p, 1, bytes, fp.f) == bytes;
    };
    if (!put(&h, sizeof(h)) || !put(prm.data(), sizeof(cuco_params) * prm.size()) ||
        !put(counts.data(), sizeof(int64_t) * counts.size()) ||
        !put(hp.data(), sizeof(double4) * n) || !put(hv.data(), sizeof(double4) * n) ||
        !put(tb.data(), sizeof(int32_t) * tb.size()) ||
        !put(te.data(), sizeof(int32_t) * te.size()) || !put(app.data(), app.size()) ||
        !put(apm.data(), apm.size())) {
        set_error("short write to %s", path);
        return CUCO_ERR_INVALID_ARG;
    }
    const int closed = fclose(fp.f);
    fp.f = nullptr;
    if (closed != 0) {
        set_error("closing %s failed", path);
        return CUCO_ERR_INVALID_ARG;
    }
    return CUCO_OK;
} CUCO_CATCH

int cuco_checkpoint_read(cuco_engine *e, const char *path)
try {
    CHECK_ENGINE(e);
    cuco_checkpoint_header h;
    TRY(cuco_checkpoint_info(path, &h));
    if (h.nboxes != (int32_t)e->box.size()) {
        set_error("%s holds %d boxes, the handle %zu", path, h.nboxes, e->box.size());
        return CUCO_ERR_INVALID_ARG;
    }
    TRY(check_count(e, h.n));
    const size_t n = (size_t)h.n, nb = (size_t)h.nboxes;
    CkFile fp;
    fp.f = fopen(path, "rb");
    if (!fp.f) {
        set_error("cannot open checkpoint file %s", path);
        return CUCO_ERR_INVALID_ARG;
    }
    auto get = [&](void *p, size_t bytes) {
        return bytes == 0 || fread(p, 1, bytes, fp.f) == bytes;
    };
    std::vector<cuco_params> prm(nb);
    std::vector<int64_t> counts(nb);
    std::vector<double4> hp(n), hv(n);
    std::vector<int32_t> tb(h.timeline ? n : 0), te(h.timeline ? n : 0);
    std::vector<char> app(accel_bytes(h.pp_precision) * n), apm(accel_bytes(h.pm_precision) * n);
    if (fseek(fp.f, (long)sizeof(h), SEEK_SET) != 0 ||
        !get(prm.data(), sizeof(cuco_params) * nb) || !get(counts.data(), sizeof(int64_t) * nb) ||
        !get(hp.data(), sizeof(double4) * n) || !get(hv.data(), sizeof(double4) * n) ||
        !get(tb.data(), sizeof(int32_t) * tb.size()) ||
        !get(te.data(), sizeof(int32_t) * te.size()) || !get(app.data(), app.size()) ||
        !get(apm.data(), apm.size())) {
        set_error("%s: short read", path);
        return CUCO_ERR_INVALID_ARG;
    }
    // Everything is checked before the handle changes.
    int64_t i = 0;
    for (size_t b = 0; b < nb; ++b) {
        const cuco_params &a = prm[b], &c = e->box[b].prm;
        const int64_t want = e->ensemble ? (int64_t)e->box[b].cap : h.n;
        if (a.pm_grid != c.pm_grid || a.box_size != c.box_size || a.G != c.G ||
            a.asmth != c.asmth || a.rcut != c.rcut || counts[b] != want || !(a.soft_h > 0)) {
            set_error("%s: box %zu was written with other parameters or particle count", path, b);
            return CUCO_ERR_INVALID_ARG;
        }
        for (int64_t k = 0; k < counts[b]; ++k, ++i) {
            const double4 q = hp[i];
            if (!(q.x >= 0 && q.x < c.box_size && q.y >= 0 && q.y < c.box_size && q.z >= 0 &&
                  q.z < c.box_size)) {
                set_error("%s: particle %lld lies outside box %zu", path, (long long)i, b);
                return CUCO_ERR_INVALID_ARG;
            }
        }
    }
    if (h.timeline) {
        const cuco_timeline_state &st = h.state;
        bool ok = timeline_params_ok(&h.params) && st.ti_current >= 0 &&
                  st.ti_current <= kTimebase && st.pm_ti_begstep >= 0 &&
                  st.pm_ti_begstep <= st.pm_ti_endstep && st.pm_ti_endstep <= kTimebase;
        for (size_t k = 0; ok && k < n; ++k)
            ok = tb[k] >= 0 && tb[k] <= te[k] && te[k] <= kTimebase;
        if (!ok) {
            set_error("%s: integer-timeline parameters or state out of range", path);
            return CUCO_ERR_INVALID_ARG;
        }
    }

    std::vector<double> pos3(3 * n), vel3(3 * n), mass(n);
    for (size_t k = 0; k < n; ++k) {
        pos3[3 * k] = hp[k].x;
        pos3[3 * k + 1] = hp[k].y;
        pos3[3 * k + 2] = hp[k].z;
        mass[k] = hp[k].w;
        vel3[3 * k] = hv[k].x;
        vel3[3 * k + 1] = hv[k].y;
        vel3[3 * k + 2] = hv[k].z;
    }
    for (size_t b = 0; b < nb; ++b) {
        e->box[b].prm.soft_h = prm[b].soft_h;
        set_pair_params(e->box[b]);
    }
    TRY(upload_impl(e, h.n, pos3.data(), vel3.data(), mass.data(), 0.0, false));
    destroy_graph(e);
    TRY(update_pair_constants(e));
    if (h.timeline) {
        TRY(timeline_setup(e, &h.params));
        CUCO_HIP(hipMemcpyAsync(e->ti_beg, tb.data(), sizeof(int32_t) * n,
                                hipMemcpyHostToDevice, e->stream));
        CUCO_HIP(hipMemcpyAsync(e->ti_end, te.data(), sizeof(int32_t) * n,
                                hipMemcpyHostToDevice, e->stream));
        TRY(launch_fill_u32(reinterpret_cast<uint32_t *>(e->ti_scalars), 0u, 4, e->stream));
        e->tls = h.state;
        e->tl_init = true;
    }
    if (!app.empty())
        CUCO_HIP(hipMemcpyAsync(accel_buffer(e, true, h.pp_precision), app.data(), app.size(),
                                hipMemcpyHostToDevice, e->stream));
    if (!apm.empty())
        CUCO_HIP(hipMemcpyAsync(accel_buffer(e, false, h.pm_precision), apm.data(), apm.size(),
                                hipMemcpyHostToDevice, e->stream));
    CUCO_HIP(hipStreamSynchronize(e->stream));
    e->pp_prec = h.pp_precision;
    e->pm_prec = h.pm_precision;
    return CUCO_OK;
} CUCO_CATCH

int cuco_download_records(cuco_engine *e, cuco_particle *P)
try {
    CHECK_ENGINE(e);
    TRY(check_ready(e));
    if (!P) {
        set_error("null particle array");
        return CUCO_ERR_INVALID_ARG;
    }
    const size_t n = e->n;
    std::vector<double> pos(3 * n), vel(3 * n), app(3 * n, 0.0), apm(3 * n, 0.0);
    std::vector<double4> hp(n);
    CUCO_HIP(hipMemcpyAsync(hp.data(), e->pos, sizeof(double4) * n, hipMemcpyDeviceToHost, e->stream));
    CUCO_HIP(hipStreamSynchronize(e->stream));
    TRY(cuco_download_particles(e, pos.data(), vel.data()));
    if (e->pp_prec >= 0) TRY(cuco_download_accel(e, app.data(), nullptr));
    if (e->pm_prec >= 0) TRY(cuco_download_accel(e, nullptr, apm.data()));
    std::vector<int32_t> tb(n, 0), te(n, 0);
    if (e->tl_init) {
        CUCO_HIP(hipMemcpyAsync(tb.data(), e->ti_beg, sizeof(int32_t) * n, hipMemcpyDeviceToHost, e->stream));
        CUCO_HIP(hipMemcpyAsync(te.data(), e->ti_end, sizeof(int32_t) * n, hipMemcpyDeviceToHost, e->stream));
        CUCO_HIP(hipStreamSynchronize(e->stream));
    }
    for (size_t i = 0; i < n; ++i) {
        for (int d = 0; d < 3; ++d) {
            P[i].Pos[d] = (float)pos[3 * i + d];
            P[i].Vel[d] = (float)vel[3 * i + d];
            P[i].GravAccel[d] = (float)app[3 * i + d];
            P[i].GravPM[d] = (float)apm[3 * i + d];
        }
        P[i].Mass = (float)hp[i].w;
        P[i].Ti_endstep = te[i];
        P[i].Ti_begstep = tb[i];
    }
    return CUCO_OK;
} CUCO_CATCH

int cuco_upload_records(cuco_engine *e, const cuco_particle *P, int64_t n)
try {
    CHECK_ENGINE(e);
    if (!P || n <= 0) {
        set_error("null particle array or n <= 0");
        return CUCO_ERR_INVALID_ARG;
    }
    std::vector<double> pos((size_t)n * 3), vel((size_t)n * 3), mass((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        for (int d = 0; d < 3; ++d) {
            pos[3 * i + d] = P[i].Pos[d];
            vel[3 * i + d] = P[i].Vel[d];
        }
        mass[i] = P[i].Mass;
    }
    return cuco_upload_particles(e, n, pos.data(), vel.data(), mass.data(), 0.0);
} CUCO_CATCH

int cuco_pmforce_periodic(cuco_engine *e, cuco_particle *P, int64_t n, int32_t precision)
try {
    CHECK_ENGINE(e);
    if (!P) {
        set_error("null particle array");
        return CUCO_ERR_INVALID_ARG;
    }
    TRY(check_count(e, n));
    std::vector<double> pos((size_t)n * 3), mass((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        for (int d = 0; d < 3; ++d) pos[3 * i + d] = P[i].Pos[d];
        mass[i] = P[i].Mass;
    }
    TRY(upload_impl(e, n, pos.data(), nullptr, mass.data(), 0.0, true));
    TRY(cuco_pm_force(e, precision));
    std::vector<double> acc((size_t)n * 3);
    TRY(cuco_download_accel(e, nullptr, acc.data()));
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) P[i].GravPM[d] = (float)acc[3 * i + d];
    return CUCO_OK;
} CUCO_CATCH

int cuco_gravity_linklist(cuco_engine *e, cuco_particle *P, int64_t n, int32_t precision)
try {
    CHECK_ENGINE(e);
    if (!P) {
        set_error("null particle array");
        return CUCO_ERR_INVALID_ARG;
    }
    TRY(check_count(e, n));
    std::vector<double> pos((size_t)n * 3), mass((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        for (int d = 0; d < 3; ++d) pos[3 * i + d] = P[i].Pos[d];
        mass[i] = P[i].Mass;
    }
    TRY(upload_impl(e, n, pos.data(), nullptr, mass.data(), 0.0, true));
    TRY(cuco_pp_force(e, precision));
    std::vector<double> acc((size_t)n * 3);
    TRY(cuco_download_accel(e, acc.data(), nullptr));
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) P[i].GravAccel[d] = (float)acc[3 * i + d];
    return CUCO_OK;
} CUCO_CATCH

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
try {
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
} CUCO_CATCH

int cuco_device_drift_table(int32_t device, double a_begin, double a_max, double omega0,
                            double omega_lambda, double hubble, double *drift, double *kick)
{
    if (!(a_begin > 0) || !(a_max > a_begin) || !(hubble > 0) || !drift || !kick) {
        set_error("a_begin, a_max, hubble out of range or null tables");
        return CUCO_ERR_INVALID_ARG;
    }
    CUCO_HIP(hipSetDevice(device));
    double *d = nullptr;
    TRY(dev_alloc(d, sizeof(double) * 2 * kDriftTable));
    int rc = launch_drift_tables(a_begin, a_max, omega0, omega_lambda, hubble, d,
                                 d + kDriftTable, nullptr);
    if (!rc && (hipMemcpy(drift, d, sizeof(double) * kDriftTable, hipMemcpyDeviceToHost) !=
                    hipSuccess ||
                hipMemcpy(kick, d + kDriftTable, sizeof(double) * kDriftTable,
                          hipMemcpyDeviceToHost) != hipSuccess)) {
        set_error("copy of the drift tables failed");
        rc = CUCO_ERR_HIP;
    }
    free_ptr(d);
    return rc;
}

int cuco_radix_sort_noalloc(uint32_t *d_keys, uint32_t *d_values, uint32_t n,
                            uint32_t key_bits)
{
    if (n == 0) return CUCO_OK;
    if (!d_keys || key_bits == 0 || key_bits > 32) {
        set_error("null keys or key_bits %u outside 1..32", key_bits);
        return CUCO_ERR_INVALID_ARG;
    }
    TRY(radix_sort_static(d_keys, d_values, n, (int)key_bits, nullptr));
    CUCO_HIP(hipStreamSynchronize(nullptr));
    return CUCO_OK;
}

int cuco_get_sorted(cuco_engine *e, uint32_t *keys, uint32_t *index)
try {
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
} CUCO_CATCH

int cuco_get_cell_table(cuco_engine *e, uint32_t *cell_start, uint32_t *cell_end)
try {
    CHECK_ENGINE(e);
    TRY(check_ready(e));
    if (e->sorted_mode != kKeyXMajor) {
        set_error("cell table is available after an FP64 (reference order) evaluation");
        return CUCO_ERR_STATE;
    }
    const uint32_t nc = e->nkeys[kKeyXMajor];
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
} CUCO_CATCH

int cuco_get_shortrange_table(cuco_engine *e, float *table)
{
    CHECK_ENGINE(e);
    CUCO_HIP(hipMemcpy(table, e->table, sizeof(float) * kNTab, hipMemcpyDeviceToHost));
    return CUCO_OK;
}

int cuco_get_pm_grid(cuco_engine *e, int32_t which, double *out)
try {
    CHECK_ENGINE(e);
    if (e->pm_prec < 0 || (which != 0 && which != 1) || !out) {
        set_error("no PM evaluation, or which not in {0,1}");
        return CUCO_ERR_STATE;
    }
    const int k = e->pm_prec == CUCO_FP64 ? 0 : 1;
    const size_t M3 = e->grid_elems;  // every box's mesh, in box order
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
} CUCO_CATCH

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
    t->pm_overlap = e->pm_overlap ? 1 : 0;
    t->pp_symmetric = e->psym.on ? 1 : 0;
    t->pp_cell_level = e->pp_cell_level;
    t->fft_split = e->bs.fft_split;
    t->pp_wide_tiles = e->pp_wide_tiles;
    return CUCO_OK;
}

int cuco_set_tuning(cuco_engine *e, const cuco_tuning *t)
{
    CHECK_ENGINE(e);
    if (!t || t->pp_block_shift < e->min_block_shift || t->pp_block_shift > 4 ||
        t->pp_block_shift >= e->min_bits || t->pp_workgroups_per_cu < 1 ||
        t->pp_workgroups_per_cu > 32 || t->pm_overlap < 0 || t->pm_overlap > 1 ||
        t->pp_symmetric < 0 || t->pp_symmetric > 1 || t->pp_cell_level < 0 ||
        t->pp_cell_level > 2 || t->fft_split < 0 || t->fft_split > 1 ||
        t->pp_wide_tiles < 0 || t->pp_wide_tiles > 1) {
        set_error("tuning out of range");
        return CUCO_ERR_INVALID_ARG;
    }
    if (t->pp_cell_level != e->pp_cell_level) {
        e->pp_cell_level = t->pp_cell_level;
        e->tiles.coarse_cells = t->pp_cell_level == 0 ? 4.0f : t->pp_cell_level == 1 ? 1e30f : 0.0f;
        destroy_graph(e);
    }
    if (t->fft_split != e->bs.fft_split) {
        e->bs.fft_split = t->fft_split;
        TRY(update_boxes(e));  // the register-FFT block maps
        destroy_graph(e);
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
    if ((t->pm_overlap != 0) != e->pm_overlap) {
        e->pm_overlap = t->pm_overlap != 0;
        destroy_graph(e);
    }
    if ((t->pp_symmetric != 0) != e->psym.on) {
        e->psym.on = t->pp_symmetric != 0;
        destroy_graph(e);
    }
    // Tiles of 128 serve the each-pair-once kernel only.
    e->pp_wide_tiles = t->pp_wide_tiles;
    const uint32_t run = e->psym.on && e->pp_wide_tiles ? 128u : 64u;
    if (run != e->tiles.run) {
        e->tiles.run = run;
        if (e->sorted_mode == kKeyMorton) e->sorted_mode = -1;  // rebuild tiles
        destroy_graph(e);
    }
    return CUCO_OK;
}

int cuco_autotune(cuco_engine *e, cuco_tuning *chosen, double *best_ms)
try {
    CHECK_ENGINE(e);
    TRY(check_ready(e));
    hipEvent_t a, b;
    CUCO_HIP(hipEventCreate(&a));
    CUCO_HIP(hipEventCreate(&b));
    const bool timing = e->timing;
    e->timing = false;
    cuco_tuning best{e->tiles.block_shift, e->pp32_wg_per_cu, 0, 0, 0, 0, 0};
    float best_t = 1e30f;
    int rc = CUCO_OK;
    for (int shift = e->min_block_shift; shift <= std::min(4, e->min_bits - 1) && !rc; ++shift) {
        for (int cfg = 0; cfg < 6; ++cfg) {
            const int wg = cfg % 3 == 0 ? 4 : cfg % 3 == 1 ? 8 : 16;
            cuco_tuning t{shift, wg, 0, cfg / 3, 0, 0, 0};
            rc = cuco_set_tuning(e, &t);
            if (!rc) rc = sort_particles(e, kKeyMorton);
            float ms = 1e30f;
            for (int rep = 0; rep < 2 && !rc; ++rep) {
                if (hipEventRecord(a, e->stream) != hipSuccess) rc = CUCO_ERR_HIP;
                e->psym.n_sorted = e->n;
                if (!rc) rc = launch_pp32(e->spos32, e->skeys, e->perm, e->off, e->bs, e->p32,
                                          e->table_cut, e->uniform_mass, e->table32, e->tiles,
                                          e->acc_pp32, nullptr, e->pp32_workgroups,
                                          nullptr, 0, e->psym, nullptr, e->n, e->stream);
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
    // Candidate enumeration level of the each-pair-once kernel (mesh cells
    // or 2x2x2-cell blocks for wide tiles, or either for every tile), with
    // tiles of 64 or 128 targets.
    if (!rc && best.pp_symmetric) {
        const cuco_tuning base = best;
        for (int cfg = 0; cfg < 6 && !rc; ++cfg) {
            cuco_tuning t = base;
            t.pp_cell_level = cfg % 3;
            t.pp_wide_tiles = cfg / 3;
            rc = cuco_set_tuning(e, &t);
            if (!rc) rc = sort_particles(e, kKeyMorton);
            float ms = 1e30f;
            for (int rep = 0; rep < 2 && !rc; ++rep) {
                if (hipEventRecord(a, e->stream) != hipSuccess) rc = CUCO_ERR_HIP;
                e->psym.n_sorted = e->n;
                if (!rc) rc = launch_pp32(e->spos32, e->skeys, e->perm, e->off, e->bs, e->p32,
                                          e->table_cut, e->uniform_mass, e->table32, e->tiles,
                                          e->acc_pp32, nullptr, e->pp32_workgroups,
                                          nullptr, 0, e->psym, nullptr, e->n, e->stream);
                if (!rc && (hipEventRecord(b, e->stream) != hipSuccess ||
                            hipEventSynchronize(b) != hipSuccess)) rc = CUCO_ERR_HIP;
                float t_ms = 0;
                if (!rc && hipEventElapsedTime(&t_ms, a, b) == hipSuccess) ms = std::min(ms, t_ms);
            }
            if (!rc && ms < best_t) {
                best_t = ms;
                best = t;
            }
        }
    }
    // Register-FFT line split: the FP32 PM chain alone (particles sorted).
    if (!rc && e->bs.fft32_mask) {
        const cuco_tuning base = best;
        float best_pm = 1e30f;
        for (int split = 0; split < 2 && !rc; ++split) {
            cuco_tuning t = base;
            t.fft_split = split;
            rc = cuco_set_tuning(e, &t);
            if (!rc) rc = ensure_pm<float>(e);
            if (!rc) rc = sort_particles(e, kKeyMorton);
            float ms = 1e30f;
            for (int rep = 0; rep < 3 && !rc; ++rep) {
                if (hipEventRecord(a, e->stream) != hipSuccess) rc = CUCO_ERR_HIP;
                if (!rc) rc = do_pm(e, CUCO_FP32);
                if (!rc && (hipEventRecord(b, e->stream) != hipSuccess ||
                            hipEventSynchronize(b) != hipSuccess)) rc = CUCO_ERR_HIP;
                float t_ms = 0;
                if (!rc && hipEventElapsedTime(&t_ms, a, b) == hipSuccess) ms = std::min(ms, t_ms);
            }
            if (!rc && ms < best_pm * 0.98f) {
                best_pm = ms;
                best = t;
            }
        }
    }
    // Whole FP32 evaluations (sort, PM, PP) with and without the PM chain
    // on its own stream; for each overlap setting, the pair kernel is also
    // retimed at 4 and 8 workgroups per CU (fewer resident pair workgroups
    // leave room for the PM kernels).
    if (!rc) {
        float best_eval = 1e30f;
        cuco_tuning pick = best;
        for (int ov = 0; ov < 2 && !rc; ++ov) {
            for (int wg : {best.pp_workgroups_per_cu, 4}) {
                if (ov == 0 && wg != best.pp_workgroups_per_cu) continue;
                cuco_tuning t = best;
                t.pp_workgroups_per_cu = wg;
                t.pm_overlap = ov;
                rc = cuco_set_tuning(e, &t);
                float ms = 1e30f;
                for (int rep = 0; rep < 3 && !rc; ++rep) {
                    e->sorted_mode = -1;
                    if (hipEventRecord(a, e->stream) != hipSuccess) rc = CUCO_ERR_HIP;
                    if (!rc) rc = do_forces(e, CUCO_FP32, 1);
                    if (!rc && (hipEventRecord(b, e->stream) != hipSuccess ||
                                hipEventSynchronize(b) != hipSuccess)) rc = CUCO_ERR_HIP;
                    float t_ms = 0;
                    if (!rc && hipEventElapsedTime(&t_ms, a, b) == hipSuccess) ms = std::min(ms, t_ms);
                }
                if (!rc && ms < best_eval * 0.98f) {
                    best_eval = ms;
                    pick = t;
                }
            }
        }
        best = pick;
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
} CUCO_CATCH

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
    return fft3d_single(M, precision, false, const_cast<void *>(d_in), d_out);
}

int cuco_fft3d_c2r(int32_t M, int32_t precision, void *d_in, void *d_out)
{
    return fft3d_single(M, precision, true, d_in, d_out);
}

}  // extern "C"
