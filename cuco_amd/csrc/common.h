// Internal definitions shared by the engine's HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

namespace cuco {

constexpr int kNTab = 1000;       // allvars.h:27
constexpr int kTabPadded = 1024;  // FP32 table padded with zeros past rcut

// Key layouts of the cell sort.
enum KeyMode : int {
    kKeyXMajor = 0,  // (ix*M+iy)*M+iz, the reference order (gravlist_kernel.cu:114)
    kKeyMorton = 1,  // bit-interleaved (x,y,z): compact tiles for the FP32 PP kernel
};

// Scalars every kernel of one evaluation needs; passed by value.
struct GridInfo {
    int32_t M;          // mesh cells per dimension
    int32_t key_mode;   // KeyMode
    int32_t bits;       // bits per coordinate of the Morton key
    int32_t key_bits;   // total key bits
    uint32_t nkeys;     // size of the key space (table has nkeys+1 entries)
    double L;           // box size
    double fac;         // M/L (to_grid_fac, gravlist.cpp:61)
    double cell;        // L/M
};

struct PairParams {
    double L, boxhalf;
    double rcut, rcut2;
    double asmthfac;
    double h, h_inv, h3_inv;
    double G;
    double fac;  // M/L
    int32_t M;
};

// ---------------------------------------------------------------- keys
__host__ __device__ inline uint32_t spread3(uint32_t v)
{
    // Inserts two zero bits between the low 10 bits of v.
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__host__ __device__ inline uint32_t compact3(uint32_t v)
{
    v &= 0x09249249u;
    v = (v | (v >> 2)) & 0x030c30c3u;
    v = (v | (v >> 4)) & 0x0300f00fu;
    v = (v | (v >> 8)) & 0x030000ffu;
    v = (v | (v >> 16)) & 0x000003ffu;
    return v;
}

__host__ __device__ inline uint32_t cell_key(const GridInfo &g, int ix, int iy,
                                             int iz)
{
    if (g.key_mode == kKeyXMajor)
        return (uint32_t)(((int64_t)ix * g.M + iy) * g.M + iz);
    return (spread3((uint32_t)ix) << 2) | (spread3((uint32_t)iy) << 1) |
           spread3((uint32_t)iz);
}

__host__ __device__ inline void key_cell(const GridInfo &g, uint32_t key,
                                         int &ix, int &iy, int &iz)
{
    if (g.key_mode == kKeyXMajor) {
        iz = (int)(key % (uint32_t)g.M);
        uint32_t r = key / (uint32_t)g.M;
        iy = (int)(r % (uint32_t)g.M);
        ix = (int)(r / (uint32_t)g.M);
    } else {
        ix = (int)compact3(key >> 2);
        iy = (int)compact3(key >> 1);
        iz = (int)compact3(key);
    }
}

// (int)(fac*x) clamped to [0, M-1]: gravlist_kernel.cu:111-113 with the
// clamp of pm_periodic.c:100-101.
//
// The clamp happens in floating point, which gives the same index for every
// position in [0, L) and keeps a non-finite value from becoming an
// out-of-range integer (fmax(NaN, 0) is 0).
__host__ __device__ inline int cell_coord(double fac, double x, int M)
{
    return (int)fmin(fmax(fac * x, 0.0), (double)(M - 1));
}

// Orders one wave's LDS accesses before and after the call.
__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// off[k] and off[k+1] of a cell table in one 8-byte load (the entries are
// only 4-byte aligned, which global loads accept).
typedef uint32_t u32x2_a4 __attribute__((ext_vector_type(2), aligned(4)));
__device__ inline void load_range(const uint32_t *__restrict__ off, uint32_t k, uint32_t &s,
                                  uint32_t &e)
{
    const u32x2_a4 v = *reinterpret_cast<const u32x2_a4 *>(off + k);
    s = v.x;
    e = v.y;
}

__host__ __device__ inline int wrap_index(int a, int M)
{
    a = a >= M ? a - M : a;
    return a < 0 ? a + M : a;
}

// ---------------------------------------------------------------- FFT plans
// Stockham stages of a length-n transform: radix 4 first, then 2, 3, 5.
struct FftPlan {
    int nstages;
    int radix[16];
};

inline FftPlan make_plan(int n)
{
    FftPlan p{};
    int m = n;
    while (m % 4 == 0 && p.nstages < 16) { p.radix[p.nstages++] = 4; m /= 4; }
    while (m % 2 == 0 && p.nstages < 16) { p.radix[p.nstages++] = 2; m /= 2; }
    while (m % 3 == 0 && p.nstages < 16) { p.radix[p.nstages++] = 3; m /= 3; }
    while (m % 5 == 0 && p.nstages < 16) { p.radix[p.nstages++] = 5; m /= 5; }
    if (m != 1) p.nstages = -1;
    return p;
}

// ---------------------------------------------------------------- boxes
// Every launch works on a set of independent periodic boxes (an ensemble; a
// single simulation is a set of one).  Box b owns the particles [p0, p0+n)
// in both the original and the sorted order (its keys form one contiguous
// range starting at key_base), one real and one complex mesh, and a slice of
// the twiddle and Green's function tables.
struct BoxDev {
    GridInfo g;         // mesh and key layout of the box
    uint32_t key_base;  // first global key of the box
    uint32_t p0, n;     // particle range
    uint32_t wave0;     // first 64-particle deposit slot of the box
    uint64_t grid0;     // first element of the real mesh (M^3)
    uint64_t cgrid0;    // first element of the complex mesh (M*M*(M/2+1))
    uint32_t tab0;      // first entry of the twiddle / Green tables (M each)
    int32_t pad0;
    double pm_fac;      // G/(pi L) * M/(2L) (pm_periodic.c:85-86)
    double dep_mmax;    // FP32 deposit: the box's largest particle mass (fixed-point bound)
    FftPlan half, full; // plans for length M/2 (rows) and M (lines)
};

// What one launch sees: the boxes and the exclusive prefix over boxes of
// the launch's work units (blocks, or tiles for the pair kernel).
struct BoxView {
    const BoxDev *box;
    const uint32_t *map;  // nboxes+1 entries, map[0] = 0
    int nboxes;
};

// Box holding work unit u (u < map[nboxes]): the last b with map[b] <= u.
// Boxes without units are skipped.  Called with wave-uniform u.
__device__ inline int find_box(const uint32_t *__restrict__ map, int nboxes,
                               uint32_t u)
{
    int lo = 0, hi = nboxes;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (map[mid] <= u) lo = mid; else hi = mid;
    }
    return lo;
}

// ---------------------------------------------------------------- complex
template <typename T>
struct Cx {
    T x, y;
};

template <typename T>
__device__ inline Cx<T> cmul(Cx<T> a, Cx<T> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

template <typename T>
__device__ inline Cx<T> cadd(Cx<T> a, Cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ inline Cx<T> csub(Cx<T> a, Cx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ inline Cx<T> cconj(Cx<T> a) { return {a.x, -a.y}; }

// ---------------------------------------------------------------- errors
void set_error(const char *fmt, ...);

#define CUCO_HIP(call)                                                        \
    do {                                                                      \
        hipError_t err_ = (call);                                             \
        if (err_ != hipSuccess) {                                             \
            ::cuco::set_error("%s failed: %s (%s:%d)", #call,                 \
                              hipGetErrorString(err_), __FILE__, __LINE__);  \
            return CUCO_ERR_HIP;                                              \
        }                                                                     \
    } while (0)

#define CUCO_LAUNCH_CHECK()                                                   \
    do {                                                                      \
        hipError_t err_ = hipGetLastError();                                  \
        if (err_ != hipSuccess) {                                             \
            ::cuco::set_error("kernel launch failed: %s (%s:%d)",             \
                              hipGetErrorString(err_), __FILE__, __LINE__);  \
            return CUCO_ERR_HIP;                                              \
        }                                                                     \
    } while (0)

inline unsigned grid_for(int64_t n, int block)
{
    return (unsigned)((n + block - 1) / block);
}

// Workgroups go to the 8 XCDs (each with its own L2) round-robin in
// blockIdx order.  Launch xcd_grid(nb) blocks and map each to
// xcd_block(nb): XCD k then works on the k-th contiguous eighth of the
// logical blocks, so neighbouring blocks (which share cache lines) share an
// L2.  Returns -1 for the padding blocks.
constexpr unsigned kXcds = 8;
inline unsigned xcd_grid(unsigned nb) { return (nb + kXcds - 1) / kXcds * kXcds; }
__device__ inline long xcd_block(unsigned nb)
{
    const unsigned per = (nb + kXcds - 1) / kXcds;
    const unsigned b = blockIdx.x;
    const unsigned logical = (b % kXcds) * per + b / kXcds;
    return logical < nb ? (long)logical : -1;
}

// Logical block (XCD-remapped) -> (box, block inside the box).  Returns
// false for padding blocks.
__device__ inline bool box_block(const BoxView &v, int &bi, uint32_t &lb)
{
    const uint32_t total = v.map[v.nboxes];
    const long l = xcd_block(total);
    if (l < 0) return false;
    // Box and block are uniform over the workgroup; in scalar registers the
    // descriptor loads that depend on them can be scalar too.
    bi = __builtin_amdgcn_readfirstlane(find_box(v.map, v.nboxes, (uint32_t)l));
    lb = __builtin_amdgcn_readfirstlane((uint32_t)l - v.map[bi]);
    return true;
}

}  // namespace cuco
