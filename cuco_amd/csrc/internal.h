// Launchers shared between the engine's translation units.
#pragma once

#include "common.h"
#include "cuco_engine.h"

namespace cuco {

// ------------------------------------------------------------ boxes
// Mesh sizes with register FFTs (fft32_mask_bit).
constexpr int kRegFftSizes = 9;

// Work units counted by the block maps (one prefix over boxes each).
enum BoxMap : int {
    kMapPart = 0,   // 256 particles per block
    kMapGatherX,    // FP64 deposit gather bricks (x-major keys)
    kMapRowsD,      // FFT z-row blocks, FP64
    kMapRowsF,      // FFT z-row blocks, FP32
    kMapLinesD,     // FFT y/x-line blocks, FP64
    kMapLinesF,     // FFT y/x-line blocks, FP32
    // Register-FFT z-row and y/x-line blocks, one map per mesh size
    // (fft32_mask_bit order): a launch for one size covers only the boxes
    // of that size.
    kMapRowsR,
    kMapLinesR = kMapRowsR + kRegFftSizes,
    kMapBrick = kMapLinesR + kRegFftSizes,  // 16^3-cell Morton bricks (FP32 deposit, interpolation)
    kMapCount
};

// Device-side description of the boxes of an engine plus the host copies
// of every map's total (the launch sizes).
struct BoxSet {
    int nboxes = 0;
    uint32_t capacity = 0;           // particle slots (deposit scratch)
    uint32_t wave_slots = 0;         // deposit wave slots
    BoxDev *dev[2] = {nullptr, nullptr};  // descriptors per KeyMode
    uint32_t *maps = nullptr;        // kMapCount x (nboxes+1)
    uint32_t total[kMapCount] = {};  // blocks per launch
    uint32_t *tile_map = nullptr;    // nboxes+1: first pair-kernel tile per box
    uint32_t fft32_mask = 0;         // register-FFT mesh sizes present (fft32_mask_bit)
    int fft_split = 0;               // register-FFT line split (fft32_blocks)
    BoxView view(int mode, int map) const
    {
        return BoxView{dev[mode], maps + (size_t)map * (nboxes + 1), nboxes};
    }
};
// Blocks of one box of mesh size M for a PM map (kMapGatherX..kMapBrick);
// fft_split selects the register-FFT line split.
uint32_t pm_blocks(int M, int map, int fft_split = 0);

// ------------------------------------------------------------ sort
struct SortScratch {
    uint32_t *keys_tmp = nullptr;
    uint32_t *vals_tmp = nullptr;
    uint32_t *counts = nullptr;  // 256 * ceil(n/4096) digit-major counters
    uint32_t *totals = nullptr;  // 256 per-digit totals
    uint32_t capacity = 0;       // elements the buffers hold
};

constexpr int kSortTile = 4096;
inline size_t sort_count_words(uint32_t n)
{
    return 256 * (size_t)((n + kSortTile - 1) / kSortTile + 1);
}

int launch_calc_keys(const double4 *pos, const BoxSet &bs, int mode,
                     uint32_t *keys, hipStream_t s);
// Stable LSD sort by bits [0, key_bits).  With vals_identity the input
// values are taken to be 0..n-1 and vals is only written.  On return
// *keys_out / *vals_out point at whichever buffers hold the result.
int radix_sort_pairs(uint32_t *keys, uint32_t *vals, bool vals_identity,
                     uint32_t n, int key_bits, const SortScratch &sc,
                     hipStream_t s, uint32_t **keys_out, uint32_t **vals_out);
// The same sort with its scratch in a static device array (no allocation):
// at most static_sort_capacity() pairs; the result is left in keys / vals.
// Calls that overlap on one device share the array.
uint32_t static_sort_capacity();
int radix_sort_static(uint32_t *keys, uint32_t *vals, uint32_t n, int key_bits,
                      hipStream_t s);
// off[c] = first sorted index with key >= c, c in [0, nkeys]; off[nkeys] = n.
// runs (cell_run_capacity(nkeys) entries) and nruns (one word) are scratch.
int launch_cell_offsets(const uint32_t *skeys, uint32_t n, uint32_t nkeys,
                        uint32_t *off, uint4 *runs, uint32_t run_cap, uint32_t *nruns,
                        hipStream_t s);
uint32_t cell_run_capacity(uint32_t nkeys);
int launch_reorder64(const double4 *pos, const uint32_t *perm, uint32_t n,
                     double4 *spos, hipStream_t s);
int launch_reorder32(const double4 *pos, const uint32_t *perm,
                     const uint32_t *skeys, const BoxSet &bs, int mode,
                     float4 *spos, hipStream_t s);
// Reference-format cellStart/cellEnd from the CSR offsets (x-major keys).
int launch_cell_start_end(const uint32_t *off, uint32_t ncells,
                          uint32_t *start, uint32_t *end, hipStream_t s);
// Generic exclusive scan of n uint32 (scratch: 2*ceil(n/1024)+2 words);
// out may alias in.  total (device, optional) receives the sum.
size_t scan_scratch_words(uint32_t n);
int exclusive_scan_u32(const uint32_t *in, uint32_t *out, uint32_t n,
                       uint32_t *scratch, uint32_t *total, hipStream_t s);
int launch_fill_u32(uint32_t *p, uint32_t value, size_t n, hipStream_t s);
int launch_copy16(void *dst, const void *src, size_t bytes, hipStream_t s);

// ------------------------------------------------------------ PP
// One box per launch: sorted particles [i0, i0+n), off at the box's first key.
// With active_end, only particles whose active_end[original index] equals
// ti_current get a new force (the active set of an integer-timeline step).
int launch_pp64(const double4 *spos, const uint32_t *perm, const uint32_t *off,
                uint32_t i0, uint32_t n, const PairParams &pp, const float *table,
                double4 *acc, unsigned long long *pair_counter,
                const int32_t *active_end, int32_t ti_current, hipStream_t s);

struct TileInfo {
    uint32_t *tile_start = nullptr;   // per tile: first sorted index
    uint32_t *tile_count = nullptr;   // per tile: particles (<= run)
    uint32_t *block_tiles = nullptr;  // per Morton super-block: tiles, then offsets
    uint32_t *counters = nullptr;     // [0] total tiles, [1] work counter
    uint32_t *scan_scratch = nullptr;
    uint32_t max_tiles = 0;
    uint32_t nblocks = 0;
    int block_shift = 3;              // Morton block = 2^shift cells per dim
    // Each-pair-once kernel: tiles at least this many cells wide enumerate
    // their candidates by 2x2x2-cell blocks instead of cells.
    float coarse_cells = 4.0f;
    // Tile length: 64 targets (one per lane), or 128 (two per lane, the
    // each-pair-once kernel only).
    uint32_t run = 64;
};

// Constants of the FP32 pair kernel for one box.  All lengths are scaled by
// s = 4 asmthfac (quarter table bins per length unit), so a pair's table
// bin is floor(r / 4), whose byte offset in the LDS table a far pair gets
// from one multiply-add and a mask.  A pair term in scaled units is
// dx_s * m * rinv_s^3 * T; the physical acceleration is G * s^2 times the
// scaled sum.
struct P32 {
    float s;                      // 4 asmthfac
    float rcut, rcut2, filter2;   // scaled
    float h, h_inv, h3_inv;       // scaled spline softening
    float cellf, inv_cellf;       // scaled cell size
    float eps;                    // added to r^2: keeps rsq finite at r = 0
    float out_scale;              // G * s^2
    float mass;                   // uniform particle mass of the box
    int M;
    float jscale, jinv;           // fixed-point scale of the each-pair-once j sums
    float hnear;                  // near-chunk test distance: max(h, 1)
};
// mass_max: the box's largest particle mass (bounds the fixed-point j sums).
P32 make_p32(const PairParams &pp, const GridInfo &g, double uniform_mass,
             double mass_max, uint32_t n_box);
// True when the zero tail of the LDS table can make the rcut cut alone.
bool pp32_table_cut(const PairParams &pp);

// Builds the tile list of every box and bs.tile_map (Morton keys).
int launch_build_tiles(const uint32_t *off, const BoxSet &bs, uint32_t nkeys,
                       TileInfo &t, hipStream_t s);
// Buffers of the each-pair-once variant: own sums (sorted order) and the
// j sums in 64-bit fixed point (3 planes of n_sorted).
struct PairSym {
    bool on = false;
    float4 *iacc = nullptr;
    unsigned long long *jacc = nullptr;
    uint32_t n_sorted = 0;
};
// Tiles of the active targets of an integer-timeline step (FP32 default
// variant): the active sorted indices in sorted order (list), cut into runs
// of at most 64 inside each Morton super-block; t.tile_start indexes list.
struct ActiveTiles {
    TileInfo t;
    uint32_t *scan = nullptr;      // n+1: active particles before each sorted index
    uint32_t *nscratch = nullptr;  // scan scratch for n entries
    uint32_t *list = nullptr;      // n: active sorted indices
    uint32_t *tile_map = nullptr;  // nboxes+1: first active tile per box
    uint32_t capacity = 0;         // particles the buffers hold
    uint32_t nkeys = 0;            // Morton keys of all boxes
};
// With active_end and at, only the active targets are evaluated, packed into
// their own tiles (at); without at, tiles holding an active target are
// evaluated whole and only active lanes store.
int launch_pp32(const float4 *spos, const uint32_t *skeys,
                const uint32_t *perm, const uint32_t *off, const BoxSet &bs,
                const P32 *p32s, bool table_cut, double uniform_mass,
                const float *table32, const TileInfo &t, float4 *acc,
                unsigned long long *cand_counter, int num_workgroups,
                const int32_t *active_end, int32_t ti_current,
                const PairSym &sym, const ActiveTiles *at, uint32_t n_sorted,
                hipStream_t s);

// ------------------------------------------------------------ PM
struct PmTables {
    void *tw = nullptr;      // M complex twiddles exp(-2 pi i k/M), precision T
    void *gexp = nullptr;    // M: exp(-k^2 asmth2) with signed k
    void *gsinc = nullptr;   // M: 1/sinc(pi k/M)^4
    void *gprod = nullptr;   // M: gexp * gsinc (one load per element, FP32 kernels)
};
// scratch: pm_deposit_scratch_bytes(bs, fp64) bytes (FP64: per-cell corner
// sums and per-wave partials; FP32: one 17^3 partial tile per brick).
template <typename T>
int pm_deposit(const void *spos, const uint32_t *skeys, const uint32_t *off,
               const BoxSet &bs, int mode, T *rho, void *scratch, hipStream_t s);
size_t pm_deposit_scratch_bytes(const BoxSet &bs, int fp64);
// Clears the parts of a new FP32 deposit scratch that must start at zero.
int pm_deposit_scratch_init(const BoxSet &bs, void *scratch, hipStream_t s);
// Fixed-point scale of the FP32 deposit tiles for a box of total mass m_total.
template <typename T>
int pm_fft_forward(const T *rho, Cx<T> *c, const BoxSet &bs, const PmTables &tb,
                   hipStream_t s);
template <typename T>
int pm_green_x(Cx<T> *c, const BoxSet &bs, const PmTables &tb, hipStream_t s);
template <typename T>
int pm_fft_inverse(Cx<T> *c, T *phi, const BoxSet &bs, const PmTables &tb,
                   hipStream_t s);
// off: the cell table of the sort (FP32: Morton keys, read per brick).
// dep_scratch: the FP32 deposit scratch of the same evaluation (its work
// units); unused for FP64.
template <typename T>
int pm_interp(const T *phi, const void *spos, const uint32_t *skeys,
              const uint32_t *perm, const uint32_t *off, const BoxSet &bs, int mode,
              const void *dep_scratch, void *acc, hipStream_t s);
template <typename T>
int pm_fft_x(Cx<T> *c, const BoxSet &bs, bool inverse, const PmTables &tb,
             hipStream_t s);
bool fft_size_supported(int M);
// FP32 register-resident FFTs of the boxes whose mesh is a power of two in
// 32..1024, or 96, 160 or 192 (fft32_supported); the generic FP32 maps count
// no blocks for
// those boxes, so every box takes the same kernels whatever set it is in.
// One launch per mesh size present (bs.fft32_mask).
bool fft32_supported(int M);
uint32_t fft32_mask_bit(int M);
// split 0: N1 = 4/8/16/32 threads per line of length N <= 32/128/512/1024;
// split 1: half as many threads per line, each holding twice the values.
// Lines of 3 or 5 times a power of two: 8 threads where N/8 is a multiple
// of 8, else 4, whatever the split.
uint32_t fft32_blocks(int M, bool rows, int split = 0);
int fft32_forward(const float *rho, Cx<float> *c, const BoxSet &bs, const PmTables &tb,
                  hipStream_t s);
int fft32_green_x(Cx<float> *c, const BoxSet &bs, const PmTables &tb, hipStream_t s);
int fft32_inverse(Cx<float> *c, float *phi, const BoxSet &bs, const PmTables &tb,
                  hipStream_t s);
int fft32_x(Cx<float> *c, const BoxSet &bs, bool inverse, const PmTables &tb, hipStream_t s);

// ------------------------------------------------------------ integrator
// init_drift_table (driftfac.c:14-42) on the device: kDriftTable entries
// of each table into device arrays.
int launch_drift_tables(double a_begin, double a_max, double omega0, double omega_lambda,
                        double hubble, double *drift, double *kick, hipStream_t s);
int launch_drift(double4 *pos, const double4 *vel, const BoxSet &bs,
                 double drift, hipStream_t s);
int launch_kick64(double4 *vel, const double4 *app, const double4 *apm,
                  uint32_t n, double kpp, double kpm, hipStream_t s);
int launch_kick32(double4 *vel, const float4 *app, const float4 *apm,
                  uint32_t n, double kpp, double kpm, hipStream_t s);
int launch_kick_table(double4 *vel, const void *app, const void *apm,
                      int fp32, uint32_t n, const double *factors,
                      const uint32_t *step, int which, hipStream_t s);
int launch_drift_table(double4 *pos, const double4 *vel, const BoxSet &bs,
                       const double *factors, const uint32_t *step,
                       hipStream_t s);
int launch_step_advance(uint32_t *step, hipStream_t s);

// Integer timeline (timestep.c, run.c): constants of one system step.
constexpr int kTimebase = 1 << 28;   // allvars.h:11
constexpr int kDriftTable = 1000;    // allvars.h:25
struct TimelineDev {
    int32_t ti_current;
    int32_t pad0;
    double log_begin, log_max, interval;  // ln TimeBegin, ln TimeMax, Timebase_interval
    double fac1, atime, hubble_a;         // 1/a^2, a, H(a) (timestep.c:18-24)
    double err_tol, soft_table;           // ErrTolIntAccuracy, SofteningTable
    double dt_displacement, min_dt;       // MaxSizeTimestep, MinSizeTimestep
};
int launch_advance(double4 *vel, const void *app, const void *apm, int fp32,
                   int32_t *ti_beg, int32_t *ti_end, uint32_t n,
                   const TimelineDev &t, const double *kick_table,
                   int32_t *flags, hipStream_t s);
// out[0] = min(ti_end[0..n)) (out is cleared first).  ti_end must be
// 16-byte aligned.
int launch_ti_min(const int32_t *ti_end, uint32_t n, int32_t *out, hipStream_t s);
// Pair-work estimate of the active set (sorted order, CSR cell offsets):
// out[0] = sum over particles of the occupancy of their aligned block of
// 2^shift keys, out[1] = the same over the active ones (out cleared first;
// every box's key range must be a multiple of 2^shift).
int launch_active_work(const uint32_t *skeys, const uint32_t *perm, const uint32_t *off,
                       uint32_t n, int shift, const int32_t *ti_end, int32_t ti_current,
                       unsigned long long *out, hipStream_t s);
// out[0] = number of i with ti_end[i] == ti_current (out is cleared first).
int launch_count_active(const int32_t *ti_end, uint32_t n, int32_t ti_current,
                        int32_t *out, hipStream_t s);

}  // namespace cuco
