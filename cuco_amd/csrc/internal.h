// Launchers shared between the engine's translation units.
#pragma once

#include "common.h"
#include "cuco_engine.h"

namespace cuco {

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

int launch_calc_keys(const double4 *pos, uint32_t n, const GridInfo &g,
                     uint32_t *keys, hipStream_t s);
// Stable LSD sort by bits [0, key_bits).  With vals_identity the input
// values are taken to be 0..n-1 and vals is only written.  On return
// *keys_out / *vals_out point at whichever buffers hold the result.
int radix_sort_pairs(uint32_t *keys, uint32_t *vals, bool vals_identity,
                     uint32_t n, int key_bits, const SortScratch &sc,
                     hipStream_t s, uint32_t **keys_out, uint32_t **vals_out);
// off[c] = first sorted index with key >= c, c in [0, nkeys]; off[nkeys] = n.
int launch_cell_offsets(const uint32_t *skeys, uint32_t n, uint32_t nkeys,
                        uint32_t *off, hipStream_t s);
int launch_reorder64(const double4 *pos, const uint32_t *perm, uint32_t n,
                     double4 *spos, hipStream_t s);
int launch_reorder32(const double4 *pos, const uint32_t *perm,
                     const uint32_t *skeys, uint32_t n, const GridInfo &g,
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
int launch_pp64(const double4 *spos, const uint32_t *perm, const uint32_t *off,
                uint32_t n, const PairParams &pp, const float *table,
                double4 *acc, unsigned long long *pair_counter, hipStream_t s);

struct TileInfo {
    uint32_t *tile_start = nullptr;   // per tile: first sorted index
    uint32_t *tile_count = nullptr;   // per tile: particles (<= 64)
    uint32_t *block_tiles = nullptr;  // per Morton block: tiles, then offsets
    uint32_t *counters = nullptr;     // [0] total tiles, [1] work counter
    uint32_t *scan_scratch = nullptr;
    uint32_t max_tiles = 0;
    uint32_t nblocks = 0;
    int block_shift = 3;              // Morton block = 2^shift cells per dim
};
int launch_build_tiles(const uint32_t *off, uint32_t n, const GridInfo &g,
                       TileInfo &t, hipStream_t s);
int launch_pp32(const float4 *spos, const uint32_t *skeys,
                const uint32_t *perm, const uint32_t *off, uint32_t n,
                const GridInfo &g, const PairParams &pp, const float *table32,
                const TileInfo &t, float4 *acc, unsigned long long *cand_counter,
                int num_workgroups, double uniform_mass, hipStream_t s);

// ------------------------------------------------------------ PM
struct PmTables {
    void *tw = nullptr;      // M complex twiddles exp(-2 pi i k/M), precision T
    void *gexp = nullptr;    // M: exp(-k^2 asmth2) with signed k
    void *gsinc = nullptr;   // M: 1/sinc(pi k/M)^4
};
// scratch: pm_deposit_scratch_bytes(n, fp64) bytes (per-cell corner sums
// and per-wave partials).
template <typename T>
int pm_deposit(const void *spos, const uint32_t *skeys, const uint32_t *off,
               uint32_t n, const GridInfo &g, T *rho, void *scratch,
               hipStream_t s);
size_t pm_deposit_scratch_bytes(uint32_t n, int fp64);
template <typename T>
int pm_fft_forward(const T *rho, Cx<T> *c, int M, const PmTables &tb,
                   hipStream_t s);
template <typename T>
int pm_green_x(Cx<T> *c, int M, const PmTables &tb, hipStream_t s);
template <typename T>
int pm_fft_inverse(Cx<T> *c, T *phi, int M, const PmTables &tb,
                   hipStream_t s);
template <typename T>
int pm_interp(const T *phi, const void *spos, const uint32_t *skeys,
              const uint32_t *perm, uint32_t n, const GridInfo &g, double fac,
              void *acc, hipStream_t s);
template <typename T>
int pm_fft_x(Cx<T> *c, int M, bool inverse, hipStream_t s, const PmTables &tb);
bool fft_size_supported(int M);

// ------------------------------------------------------------ integrator
int launch_drift(double4 *pos, const double4 *vel, uint32_t n, double L,
                 double drift, hipStream_t s);
int launch_kick64(double4 *vel, const double4 *app, const double4 *apm,
                  uint32_t n, double kpp, double kpm, hipStream_t s);
int launch_kick32(double4 *vel, const float4 *app, const float4 *apm,
                  uint32_t n, double kpp, double kpm, hipStream_t s);
int launch_kick_table(double4 *vel, const void *app, const void *apm,
                      int fp32, uint32_t n, const double *factors,
                      const uint32_t *step, int which, hipStream_t s);
int launch_drift_table(double4 *pos, const double4 *vel, uint32_t n, double L,
                       const double *factors, const uint32_t *step,
                       hipStream_t s);
int launch_step_advance(uint32_t *step, hipStream_t s);

}  // namespace cuco
