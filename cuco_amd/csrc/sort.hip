// Cell keys, stable LSD radix sort, CSR cell table and particle reorder.
//
// Reference behaviour (liulei/cuco):
//   key        gravlist_kernel.cu:92-118  (here clamped to M-1, any N)
//   sort       radixsort.h:170-187, radixsort.cu:1048-1145 -- a stable LSD
//              key/value sort; the permutation is the unique stable order.
//   cell table gravlist_kernel.cu:121-160 (cellStart/cellEnd); this engine
//              keeps CSR offsets (one array, no memset) and derives the
//              reference's start/end arrays on request.
//
// Sort design for gfx950: 8-bit digits, reduce-then-scan (no inter-block
// spin waits).  Per pass: an upsweep histogram per 4096-key tile, a per-digit
// scan over tiles, and a downsweep that ranks every key inside its tile with
// a wave-level multisplit (eight 64-lane ballots), reorders the tile through
// LDS by digit, and writes each digit run contiguously.
#include <algorithm>
#include <mutex>

#include "internal.h"

namespace cuco {

namespace {

constexpr int kThreads = 256;
constexpr int kItems = kSortTile / kThreads;  // 16
constexpr int kBins = 256;

__device__ inline uint32_t wave_incl_scan(uint32_t v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// Exclusive scan over a 256-thread block; sh must hold 4 words.
__device__ inline uint32_t block_excl_scan256(uint32_t v, uint32_t *sh,
                                              uint32_t &total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t inc = wave_incl_scan(v);
    if (lane == 63) sh[wave] = inc;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        uint32_t t = sh[w];
        off += (w < wave) ? t : 0u;
        tot += t;
    }
    __syncthreads();
    total = tot;
    return off + inc - v;
}

__global__ void k_calc_keys(const double4 *__restrict__ pos, BoxView v,
                            uint32_t *__restrict__ keys)
{
    int bi;
    uint32_t lb;
    if (!box_block(v, bi, lb)) return;
    const BoxDev &bx = v.box[bi];
    const uint32_t il = lb * blockDim.x + threadIdx.x;
    if (il >= bx.n) return;
    const uint32_t i = bx.p0 + il;
    const GridInfo &g = bx.g;
    double4 p = pos[i];
    int ix = cell_coord(g.fac, p.x, g.M);
    int iy = cell_coord(g.fac, p.y, g.M);
    int iz = cell_coord(g.fac, p.z, g.M);
    keys[i] = bx.key_base + cell_key(g, ix, iy, iz);
}

__global__ __launch_bounds__(kThreads) void k_rs_upsweep(
    const uint32_t *__restrict__ keys, uint32_t n, int shift, uint32_t nblocks,
    uint32_t *__restrict__ counts)
{
    __shared__ uint32_t hist[4][kBins];
    const int tid = threadIdx.x, wave = tid >> 6;
#pragma unroll
    for (int w = 0; w < 4; ++w) hist[w][tid] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * (uint32_t)kSortTile;
#pragma unroll 4
    for (int r = 0; r < kItems; ++r) {
        uint32_t e = base + r * kThreads + tid;
        if (e < n) atomicAdd(&hist[wave][(keys[e] >> shift) & 255u], 1u);
    }
    __syncthreads();
    counts[(size_t)tid * nblocks + blockIdx.x] =
        hist[0][tid] + hist[1][tid] + hist[2][tid] + hist[3][tid];
}

// One block per digit: exclusive scan of that digit's counters over tiles.
__global__ __launch_bounds__(kThreads) void k_rs_scan_digits(
    uint32_t *__restrict__ counts, uint32_t nblocks, uint32_t *__restrict__ totals)
{
    __shared__ uint32_t sh[4];
    uint32_t *c = counts + (size_t)blockIdx.x * nblocks;
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < nblocks; b0 += kThreads) {
        uint32_t b = b0 + threadIdx.x;
        uint32_t v = b < nblocks ? c[b] : 0u;
        uint32_t tot;
        uint32_t ex = block_excl_scan256(v, sh, tot);
        if (b < nblocks) c[b] = ex + carry;
        carry += tot;
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

__global__ __launch_bounds__(kThreads) void k_rs_downsweep(
    const uint32_t *__restrict__ keys_in, const uint32_t *__restrict__ vals_in,
    uint32_t *__restrict__ keys_out, uint32_t *__restrict__ vals_out,
    uint32_t n, int shift, uint32_t nblocks, const uint32_t *__restrict__ counts,
    const uint32_t *__restrict__ totals)
{
    __shared__ uint32_t s_keys[kSortTile];
    __shared__ uint32_t s_vals[kSortTile];
    __shared__ uint32_t s_wcnt[4][kBins];  // per-wave digit counts, then offsets
    __shared__ uint32_t s_tileoff[kBins];
    __shared__ uint32_t s_gbase[kBins];
    __shared__ uint32_t s_scan[4];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t base = blockIdx.x * (uint32_t)kSortTile;

    uint32_t tot;
    uint32_t dbase = block_excl_scan256(totals[tid], s_scan, tot);
    s_gbase[tid] = dbase + counts[(size_t)tid * nblocks + blockIdx.x];
#pragma unroll
    for (int w = 0; w < 4; ++w) s_wcnt[w][tid] = 0;
    __syncthreads();

    // Each wave ranks its own contiguous quarter of the tile, 64 keys at a
    // time: a key's rank among equal digits of the quarter is the wave's
    // running count of that digit plus its rank inside the 64 (eight-ballot
    // match).  No block barrier until the quarter is done.
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint32_t *wc = s_wcnt[wave];
    const uint32_t wbase = base + (uint32_t)wave * (kSortTile / 4);
    uint32_t my_key[kItems], my_val[kItems], my_rank[kItems];
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
        const uint32_t e = wbase + (uint32_t)r * 64 + lane;
        const bool valid = e < n;
        const uint32_t k = valid ? keys_in[e] : 0u;
        const uint32_t v = valid ? (vals_in ? vals_in[e] : e) : 0u;
        const uint32_t d = (k >> shift) & 255u;
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = __ballot(bit);
            m &= bit ? bb : ~bb;
        }
        const uint32_t rank_w = (uint32_t)__popcll(m & lt);
        const uint32_t before = valid ? wc[d] : 0u;
        wave_sync();  // every lane has read the count before its leader updates it
        if (valid && rank_w == 0) wc[d] = before + (uint32_t)__popcll(m);
        wave_sync();
        my_rank[r] = before + rank_w;
        my_key[r] = k;
        my_val[r] = v;
    }
    __syncthreads();

    // Per digit: tile offset, then each wave's start inside the digit run.
    const uint32_t c0 = s_wcnt[0][tid], c1 = s_wcnt[1][tid], c2 = s_wcnt[2][tid],
                   c3 = s_wcnt[3][tid];
    uint32_t tt;
    const uint32_t toff = block_excl_scan256(c0 + c1 + c2 + c3, s_scan, tt);
    s_tileoff[tid] = toff;
    s_wcnt[0][tid] = toff;
    s_wcnt[1][tid] = toff + c0;
    s_wcnt[2][tid] = toff + c0 + c1;
    s_wcnt[3][tid] = toff + c0 + c1 + c2;
    __syncthreads();

#pragma unroll
    for (int r = 0; r < kItems; ++r) {
        const uint32_t e = wbase + (uint32_t)r * 64 + lane;
        if (e < n) {
            const uint32_t d = (my_key[r] >> shift) & 255u;
            const uint32_t lp = wc[d] + my_rank[r];
            s_keys[lp] = my_key[r];
            s_vals[lp] = my_val[r];
        }
    }
    __syncthreads();

    const uint32_t tile_n = min((uint32_t)kSortTile, n - base);
    for (uint32_t lp = tid; lp < tile_n; lp += kThreads) {
        uint32_t k = s_keys[lp];
        uint32_t d = (k >> shift) & 255u;
        uint32_t gp = s_gbase[d] + (lp - s_tileoff[d]);
        keys_out[gp] = k;
        if (vals_out) vals_out[gp] = s_vals[lp];
    }
}

// Each i in [0, n] owns the keys (key[i-1], key[i]] and writes off[c] = i
// for them (key[-1] = -1, key[n] = nkeys): every entry is written exactly
// once.  A wave's ranges follow one another in key order and are written
// 64 entries at a time (coalesced); an entry's owner is the last non-empty
// range starting at or before it: each window marks the starts that fall
// inside it in LDS, and a lane takes the highest marked slot at or below
// it (a ballot and a leading-zero count), or the owner carried from the
// window before.  A range longer than kLongRun (a void, or the padding of
// a Morton key space around a mesh that is not a power of two: runs of
// millions of empty keys) would keep its wave filling alone, so it goes
// to a list of pieces of at most kRunPiece entries that k_fill_runs spreads
// over the whole chip.
constexpr uint32_t kLongRun = 2048;
constexpr uint32_t kRunPiece = 65536;

__global__ __launch_bounds__(256) void k_cell_offsets(
    const uint32_t *__restrict__ skeys, uint32_t n, uint32_t nkeys,
    uint32_t *__restrict__ off, uint4 *__restrict__ runs, uint32_t run_cap,
    uint32_t *__restrict__ nruns)
{
    __shared__ uint32_t s_own[4][64];
    __shared__ uint32_t s_lo[4][64];
    __shared__ uint32_t s_start[4][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    uint32_t lo = 0, len = 0;
    if (i <= n) {
        const int64_t prev = (i == 0) ? -1 : (int64_t)skeys[i - 1];
        const int64_t cur = (i == n) ? (int64_t)nkeys : (int64_t)skeys[i];
        lo = (uint32_t)(prev + 1);
        len = cur > prev ? (uint32_t)(cur - prev) : 0u;
    }
    if (len > kLongRun) {
        const uint32_t pieces = (len + kRunPiece - 1) / kRunPiece;
        const uint32_t slot = atomicAdd(nruns, pieces);
        for (uint32_t k = 0; k < pieces && slot + k < run_cap; ++k)
            runs[slot + k] = make_uint4(lo + k * kRunPiece, min(kRunPiece, len - k * kRunPiece), i, 0u);
        len = 0;
    }
    const uint32_t inc = wave_incl_scan(len);
    const uint32_t start = inc - len;  // my range's place in the wave's entries
    const uint32_t total = __builtin_amdgcn_readlane(inc, 63);
    const uint32_t first = blockIdx.x * 256u + (uint32_t)(w * 64);
    uint32_t *own = s_own[w];
    s_lo[w][lane] = lo;
    s_start[w][lane] = start;
    uint32_t carry = 0;  // owner of the entry before the window
    for (uint32_t f0 = 0; f0 < total; f0 += 64) {
        own[lane] = 0xffffffffu;
        wave_sync();
        if (len > 0 && start >= f0 && start - f0 < 64u) own[start - f0] = (uint32_t)lane;
        wave_sync();
        const uint64_t marks = __ballot(own[lane] != 0xffffffffu);
        const uint64_t upto = marks & ((2ull << lane) - 1ull);  // slots 0..lane
        const int q = upto ? 63 - (int)__clzll((long long)upto) : 0;
        const uint32_t o = upto ? own[q] : carry;
        const uint32_t f = f0 + (uint32_t)lane;
        if (f < total) off[s_lo[w][o] + (f - s_start[w][o])] = first + o;
        carry = __builtin_amdgcn_readlane(o, 63);
        wave_sync();
    }
}

// The long ranges k_cell_offsets left: one workgroup per piece.
__global__ __launch_bounds__(256) void k_fill_runs(const uint4 *__restrict__ runs,
                                                   const uint32_t *__restrict__ nruns,
                                                   uint32_t run_cap, uint32_t *__restrict__ off)
{
    const uint32_t nr = min(nruns[0], run_cap);
    for (uint32_t r = blockIdx.x; r < nr; r += gridDim.x) {
        const uint4 run = runs[r];
        for (uint32_t c = threadIdx.x; c < run.y; c += 256) off[run.x + c] = run.z;
    }
}

__global__ void k_cell_start_end(const uint32_t *__restrict__ off,
                                 uint32_t ncells, uint32_t *__restrict__ start,
                                 uint32_t *__restrict__ end)
{
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    uint32_t a = off[c], b = off[c + 1];
    start[c] = b > a ? a : 0xffffffffu;
    end[c] = b > a ? b : 0xffffffffu;
}

__global__ void k_reorder64(const double4 *__restrict__ pos,
                            const uint32_t *__restrict__ perm, uint32_t n,
                            double4 *__restrict__ spos)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    spos[i] = pos[perm[i]];
}

// Cell-relative FP32 coordinates: x - ix*cell computed in FP64, then rounded.
__global__ void k_reorder32(const double4 *__restrict__ pos,
                            const uint32_t *__restrict__ perm,
                            const uint32_t *__restrict__ skeys, BoxView v,
                            float4 *__restrict__ spos)
{
    int bi;
    uint32_t lb;
    if (!box_block(v, bi, lb)) return;
    const BoxDev &bx = v.box[bi];
    const uint32_t il = lb * blockDim.x + threadIdx.x;
    if (il >= bx.n) return;
    const uint32_t i = bx.p0 + il;
    const GridInfo &g = bx.g;
    double4 p = pos[perm[i]];
    int ix, iy, iz;
    key_cell(g, skeys[i] - bx.key_base, ix, iy, iz);
    spos[i] = make_float4((float)(p.x - ix * g.cell), (float)(p.y - iy * g.cell),
                          (float)(p.z - iz * g.cell), (float)p.w);
}

// ---- generic exclusive scan (reduce / scan block sums / add)
constexpr int kScanTile = 1024;  // 256 threads x 4

__global__ __launch_bounds__(kThreads) void k_scan_reduce(
    const uint32_t *__restrict__ in, uint32_t n, uint32_t *__restrict__ sums)
{
    __shared__ uint32_t sh[4];
    uint32_t base = blockIdx.x * kScanTile + threadIdx.x * 4;
    uint32_t v = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (base + k < n) v += in[base + k];
    uint32_t tot;
    block_excl_scan256(v, sh, tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kThreads) void k_scan_sums(uint32_t *__restrict__ sums,
                                                        uint32_t nb,
                                                        uint32_t *__restrict__ total)
{
    __shared__ uint32_t sh[4];
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < nb; b0 += kThreads) {
        uint32_t b = b0 + threadIdx.x;
        uint32_t v = b < nb ? sums[b] : 0u;
        uint32_t tot;
        uint32_t ex = block_excl_scan256(v, sh, tot);
        if (b < nb) sums[b] = ex + carry;
        carry += tot;
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

__global__ __launch_bounds__(kThreads) void k_scan_apply(
    const uint32_t *__restrict__ in, uint32_t *__restrict__ out, uint32_t n,
    const uint32_t *__restrict__ sums)
{
    __shared__ uint32_t sh[4];
    uint32_t base = blockIdx.x * kScanTile + threadIdx.x * 4;
    uint32_t v[4];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v[k] = (base + k < n) ? in[base + k] : 0u;
        s += v[k];
    }
    uint32_t tot;
    uint32_t ex = block_excl_scan256(s, sh, tot) + sums[blockIdx.x];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (base + k < n) out[base + k] = ex;
        ex += v[k];
    }
}

__global__ void k_iota(uint32_t *p, uint32_t n)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

__global__ void k_copy16(uint4 *__restrict__ dst, const uint4 *__restrict__ src,
                         size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void k_fill_u32(uint32_t *p, uint32_t value, size_t n)
{
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

__global__ void k_copy_u32(uint32_t *__restrict__ dst, const uint32_t *__restrict__ src,
                           uint32_t n)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// ---- sort scratch without allocation
// A file-scope device array holds the temporary keys and values, the digit
// counters and the digit totals of one sort of up to kStaticSortKeys pairs
// (33 MiB).  Every device the library runs on holds its own copy until the
// process ends; its address is looked up once per device.
constexpr uint32_t kStaticSortKeys = 1u << 22;
constexpr size_t kStaticCountWords =
    256 * ((size_t)(kStaticSortKeys + kSortTile - 1) / kSortTile + 1);
constexpr size_t kStaticSortWords = 2 * (size_t)kStaticSortKeys + kStaticCountWords + 256;

static __device__ __align__(256) uint32_t scratch[kStaticSortWords];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
uint32_t *g_scratch_ptr[kMaxDevices];

// Device address of `scratch` on the current device (the host-side name is
// only a stand-in; kernels must get this address).
int static_sort_scratch(uint32_t **out)
{
    int dev = -1;
    CUCO_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("device %d outside 0..%d", dev, kMaxDevices - 1);
        return CUCO_ERR_INVALID_ARG;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        CUCO_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = static_cast<uint32_t *>(p);
    }
    *out = g_scratch_ptr[dev];
    return CUCO_OK;
}

}  // namespace

uint32_t static_sort_capacity() { return kStaticSortKeys; }

int radix_sort_static(uint32_t *keys, uint32_t *vals, uint32_t n, int key_bits,
                      hipStream_t s)
{
    if (n > kStaticSortKeys) {
        set_error("allocation-free sort: %u keys exceed its maximum of %u", n,
                  kStaticSortKeys);
        return CUCO_ERR_INVALID_ARG;
    }
    if (n == 0) return CUCO_OK;
    uint32_t *base = nullptr;
    int rc = static_sort_scratch(&base);
    if (rc) return rc;
    SortScratch sc;
    sc.keys_tmp = base;
    sc.vals_tmp = base + kStaticSortKeys;
    sc.counts = base + 2 * (size_t)kStaticSortKeys;
    sc.totals = sc.counts + kStaticCountWords;
    sc.capacity = kStaticSortKeys;
    uint32_t *ko = nullptr, *vo = nullptr;
    rc = radix_sort_pairs(keys, vals, false, n, key_bits, sc, s, &ko, &vo);
    if (rc) return rc;
    if (ko != keys) {
        // an odd number of passes leaves the result in the scratch copies
        const unsigned grid = std::min(grid_for(n, 256), 4096u);
        k_copy_u32<<<grid, 256, 0, s>>>(keys, ko, n);
        CUCO_LAUNCH_CHECK();
        if (vals) {
            k_copy_u32<<<grid, 256, 0, s>>>(vals, vo, n);
            CUCO_LAUNCH_CHECK();
        }
    }
    return CUCO_OK;
}

int launch_calc_keys(const double4 *pos, const BoxSet &bs, int mode,
                     uint32_t *keys, hipStream_t s)
{
    if (bs.total[kMapPart] == 0) return CUCO_OK;
    k_calc_keys<<<xcd_grid(bs.total[kMapPart]), 256, 0, s>>>(
        pos, bs.view(mode, kMapPart), keys);
    CUCO_LAUNCH_CHECK();
    return CUCO_OK;
}

int radix_sort_pairs(uint32_t *keys, uint32_t *vals, bool vals_identity,
                     uint32_t n, int key_bits, const SortScratch &sc,
                     hipStream_t s, uint32_t **keys_out, uint32_t **vals_out)
{
    *keys_out = keys;
    *vals_out = vals;
    if (n == 0) return CUCO_OK;
    if (n > sc.capacity) {
        set_error("radix sort: %u elements exceed scratch capacity %u", n,
                  sc.capacity);
        return CUCO_ERR_INVALID_ARG;
    }
    const int passes = key_bits <= 0 ? 0 : (key_bits + 7) / 8;
    const uint32_t nblocks = (n + kSortTile - 1) / kSortTile;
    uint32_t *bk[2] = {keys, sc.keys_tmp};
    uint32_t *bv[2] = {vals, sc.vals_tmp};
    int cur = 0;
    if (passes == 0 && vals_identity && vals) {
        k_iota<<<grid_for(n, 256), 256, 0, s>>>(vals, n);
        CUCO_LAUNCH_CHECK();
    }
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p;
        const uint32_t *vin =
            (vals && !(p == 0 && vals_identity)) ? bv[cur] : nullptr;
        uint32_t *vout = vals ? bv[cur ^ 1] : nullptr;
        k_rs_upsweep<<<nblocks, kThreads, 0, s>>>(bk[cur], n, shift, nblocks,
                                                  sc.counts);
        CUCO_LAUNCH_CHECK();
        k_rs_scan_digits<<<kBins, kThreads, 0, s>>>(sc.counts, nblocks,
                                                    sc.totals);
        CUCO_LAUNCH_CHECK();
        k_rs_downsweep<<<nblocks, kThreads, 0, s>>>(bk[cur], vin, bk[cur ^ 1],
                                                    vout, n, shift, nblocks,
                                                    sc.counts, sc.totals);
        CUCO_LAUNCH_CHECK();
        cur ^= 1;
    }
    *keys_out = bk[cur];
    *vals_out = vals ? bv[cur] : nullptr;
    return CUCO_OK;
}

int launch_cell_offsets(const uint32_t *skeys, uint32_t n, uint32_t nkeys,
                        uint32_t *off, uint4 *runs, uint32_t run_cap, uint32_t *nruns,
                        hipStream_t s)
{
    const int rc = launch_fill_u32(nruns, 0u, 1, s);
    if (rc != CUCO_OK) return rc;
    k_cell_offsets<<<grid_for((int64_t)n + 1, 256), 256, 0, s>>>(skeys, n, nkeys, off, runs,
                                                                 run_cap, nruns);
    CUCO_LAUNCH_CHECK();
    k_fill_runs<<<2048, 256, 0, s>>>(runs, nruns, run_cap, off);
    CUCO_LAUNCH_CHECK();
    return CUCO_OK;
}

// Every range cut into pieces is longer than kLongRun and gives at most
// len / kLongRun pieces; the ranges cover the nkeys + 1 entries once.
uint32_t cell_run_capacity(uint32_t nkeys) { return (nkeys + 1u) / kLongRun + 2u; }

int launch_cell_start_end(const uint32_t *off, uint32_t ncells,
                          uint32_t *start, uint32_t *end, hipStream_t s)
{
    if (ncells == 0) return CUCO_OK;
    k_cell_start_end<<<grid_for(ncells, 256), 256, 0, s>>>(off, ncells, start,
                                                           end);
    CUCO_LAUNCH_CHECK();
    return CUCO_OK;
}

int launch_reorder64(const double4 *pos, const uint32_t *perm, uint32_t n,
                     double4 *spos, hipStream_t s)
{
    if (n == 0) return CUCO_OK;
    k_reorder64<<<grid_for(n, 256), 256, 0, s>>>(pos, perm, n, spos);
    CUCO_LAUNCH_CHECK();
    return CUCO_OK;
}

int launch_reorder32(const double4 *pos, const uint32_t *perm,
                     const uint32_t *skeys, const BoxSet &bs, int mode,
                     float4 *spos, hipStream_t s)
{
    if (bs.total[kMapPart] == 0) return CUCO_OK;
    k_reorder32<<<xcd_grid(bs.total[kMapPart]), 256, 0, s>>>(
        pos, perm, skeys, bs.view(mode, kMapPart), spos);
    CUCO_LAUNCH_CHECK();
    return CUCO_OK;
}

size_t scan_scratch_words(uint32_t n)
{
    return (size_t)(n + kScanTile - 1) / kScanTile + 2;
}

int exclusive_scan_u32(const uint32_t *in, uint32_t *out, uint32_t n,
                       uint32_t *scratch, uint32_t *total, hipStream_t s)
{
    if (n == 0) return CUCO_OK;
    const uint32_t nb = (n + kScanTile - 1) / kScanTile;
    k_scan_reduce<<<nb, kThreads, 0, s>>>(in, n, scratch);
    CUCO_LAUNCH_CHECK();
    k_scan_sums<<<1, kThreads, 0, s>>>(scratch, nb, total);
    CUCO_LAUNCH_CHECK();
    k_scan_apply<<<nb, kThreads, 0, s>>>(in, out, n, scratch);
    CUCO_LAUNCH_CHECK();
    return CUCO_OK;
}

int launch_copy16(void *dst, const void *src, size_t bytes, hipStream_t s)
{
    const size_t n = bytes / 16;
    if (n == 0) return CUCO_OK;
    const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
    k_copy16<<<grid, 256, 0, s>>>(static_cast<uint4 *>(dst),
                                  static_cast<const uint4 *>(src), n);
    CUCO_LAUNCH_CHECK();
    return CUCO_OK;
}

int launch_fill_u32(uint32_t *p, uint32_t value, size_t n, hipStream_t s)
{
    if (n == 0) return CUCO_OK;
    k_fill_u32<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, value, n);
    CUCO_LAUNCH_CHECK();
    return CUCO_OK;
}

}  // namespace cuco
