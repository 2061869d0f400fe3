// PM long-range force: CIC deposit, 3D r2c FFT, Green's function, c2r FFT,
// four-point gradient and CIC interpolation (pm_periodic.c:67-314).
//
// Layout in HBM: real grids are M^3 x-major (index (x*M+y)*M+z), unpadded;
// the complex grid is M x M x (M/2+1), z innermost.
//
//  deposit  FP64: one thread per mesh node gathers the particles of its
//           eight neighbouring cells (sorted by cell, read through the CSR
//           table) with the reference's CIC weights (pm_periodic.c:97-132);
//           every node is written once, in a fixed order: no atomics.
//           FP32: per 16^3-cell Morton brick a node tile summed in 32-bit
//           fixed point in LDS, brick faces added in a fixed order.
//  FFT      hand-written Stockham autosort FFT in LDS, radix 4/2/3/5 (FP64,
//           and the FP32 meshes fft32.hip's register FFT does not take:
//           it takes 32..1024 in powers of two, 96, 160 and 192):
//           z rows: real length-M r2c through a length-M/2 complex FFT and the
//                   split post-process; c2r the mirror image;
//           y lines: strided batched complex FFT, lines contiguous in z so
//                   every load/store is a coalesced row segment;
//           x lines: forward FFT, multiply by the Green's function
//                   (pm_periodic.c:146-197; DC zeroed, :197), inverse FFT,
//                   all in one LDS residency.
//  interp   the gradient stencil of pm_periodic.c:260-267 at the 8 CIC
//           nodes, summed with the CIC weights (pm_periodic.c:297-306).
//           FP64: one thread per particle, the reference order.  FP32: per
//           16^3-cell brick, the potential staged in LDS.
//
// This file is compiled with -ffp-contract=off so the FP64 kernels keep the
// reference's rounding sequence for the weights and stencil (the FP32
// interpolation fuses explicitly with fmaf).
#include <algorithm>
#include <cmath>

#include "internal.h"

namespace cuco {

namespace {

constexpr int kFftThreads = 512;
// LDS tile per block: 16 KB (2048 complex) in FP32, 32 KB (2048 complex) in
// FP64.  Small tiles keep the per-thread butterfly state small, so three
// 512-thread blocks (24 waves) fit on a CU.
template <typename T>
constexpr int fft_bytes() { return sizeof(T) == 4 ? 16384 : 32768; }

template <typename T>
constexpr int fft_elems() { return fft_bytes<T>() / (int)sizeof(Cx<T>); }

// Lines per block of the strided (y, x) passes: the largest power of two
// that fits the LDS tile, so a block's element -> (line, position) split is
// a mask and a shift.
template <typename T>
__host__ __device__ inline int lines_log2(int Nf)
{
    int lg = 0;
    while ((2 << lg) * Nf <= fft_elems<T>()) ++lg;
    return lg;
}

// ------------------------------------------------------------- small DFTs
template <typename T, int R, bool INV>
__device__ inline void dft(Cx<T> *v)
{
    if (R == 2) {
        Cx<T> a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if (R == 4) {
        Cx<T> a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
        Cx<T> a2 = cadd(v[1], v[3]), a3 = csub(v[1], v[3]);
        // i*a3 = (-a3.y, a3.x)
        Cx<T> ia3 = {-a3.y, a3.x};
        v[0] = cadd(a0, a2);
        v[2] = csub(a0, a2);
        if (!INV) {
            v[1] = csub(a1, ia3);
            v[3] = cadd(a1, ia3);
        } else {
            v[1] = cadd(a1, ia3);
            v[3] = csub(a1, ia3);
        }
    } else if (R == 3) {
        const T s3 = (T)0.86602540378443864676;
        Cx<T> sum = cadd(v[1], v[2]), dif = csub(v[1], v[2]);
        Cx<T> t1 = {v[0].x - (T)0.5 * sum.x, v[0].y - (T)0.5 * sum.y};
        // t2 = -i*s3*dif (forward), +i*s3*dif (inverse)
        Cx<T> t2 = INV ? Cx<T>{-s3 * dif.y, s3 * dif.x} : Cx<T>{s3 * dif.y, -s3 * dif.x};
        v[0] = cadd(v[0], sum);
        v[1] = cadd(t1, t2);
        v[2] = csub(t1, t2);
    } else if (R == 5) {
        const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
        const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;
        Cx<T> a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
        Cx<T> a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
        Cx<T> x0 = v[0];
        Cx<T> r1 = {x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y};
        Cx<T> r2 = {x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y};
        // forward: y1 = r1 - i(s1 b1 + s2 b2), y4 = r1 + i(...)
        //          y2 = r2 - i(s2 b1 - s1 b2), y3 = r2 + i(...)
        Cx<T> q1 = {s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y};
        Cx<T> q2 = {s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y};
        Cx<T> iq1 = {-q1.y, q1.x}, iq2 = {-q2.y, q2.x};
        v[0] = {x0.x + a1.x + a2.x, x0.y + a1.y + a2.y};
        if (!INV) {
            v[1] = csub(r1, iq1);
            v[4] = cadd(r1, iq1);
            v[2] = csub(r2, iq2);
            v[3] = cadd(r2, iq2);
        } else {
            v[1] = cadd(r1, iq1);
            v[4] = csub(r1, iq1);
            v[2] = cadd(r2, iq2);
            v[3] = csub(r2, iq2);
        }
    }
}

// Element (l, j) of the LDS tile: ROWS keeps a line contiguous, COLS keeps
// the B lines of one position contiguous.
template <bool ROWS>
__device__ inline int lidx(int l, int j, int ld)
{
    return ROWS ? l * ld + j : j * ld + l;
}

// q = n / d and r = n % d for 0 <= n < 2^20 through a float reciprocal
// (inv = 1/d rounded): the estimate is off by at most one and corrected.
__device__ inline void divmod(int n, int d, float inv, int &q, int &r)
{
    q = (int)((float)n * inv);
    r = n - q * d;
    if (r < 0) { --q; r += d; }
    else if (r >= d) { ++q; r -= d; }
}

// One Stockham stage over all B lines of length Nf held in LDS.  Every
// thread reads its butterflies, the block synchronises, and every thread
// writes them to their output positions (kept from the read phase).
template <typename T, int R, bool ROWS, bool INV>
__device__ void stage(Cx<T> *lds, int Nf, int B, int lgB, int ld, int Ns,
                      const Cx<T> *__restrict__ tw, int twstride)
{
    constexpr int MAXB = (fft_elems<T>() / R + kFftThreads - 1) / kFftThreads;
    const int span = Nf / R;
    const int nb = span * B;
    const float inv_span = 1.0f / (float)span, inv_ns = 1.0f / (float)Ns;
    const int tstep = (Nf / (Ns * R)) * twstride;
    const int ostride = ROWS ? Ns : Ns * ld;  // output element stride r -> r+1
    Cx<T> v[MAXB][R];
    int out[MAXB];
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int b = threadIdx.x + q * kFftThreads;
        if (b < nb) {
            int l, j;
            if (ROWS) divmod(b, span, inv_span, l, j);
            else { l = b & (B - 1); j = b >> lgB; }
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = lds[lidx<ROWS>(l, j + r * span, ld)];
            int jq, k;
            divmod(j, Ns, inv_ns, jq, k);
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    Cx<T> w = tw[k * r * tstep];
                    if (INV) w.y = -w.y;
                    v[q][r] = cmul(v[q][r], w);
                }
            }
            dft<T, R, INV>(v[q]);
            out[q] = lidx<ROWS>(l, jq * Ns * R + k, ld);
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int b = threadIdx.x + q * kFftThreads;
        if (b < nb) {
#pragma unroll
            for (int r = 0; r < R; ++r) lds[out[q] + r * ostride] = v[q][r];
        }
    }
    __syncthreads();
}

// lgB: log2 of B for the strided passes (B is a power of two there).
template <typename T, bool ROWS, bool INV>
__device__ void run_fft(Cx<T> *lds, int Nf, int B, int lgB, int ld, const FftPlan &plan,
                        const Cx<T> *__restrict__ tw, int twstride)
{
    int Ns = 1;
    const int ns = __builtin_amdgcn_readfirstlane(plan.nstages);
    for (int s = 0; s < ns; ++s) {
        const int R = __builtin_amdgcn_readfirstlane(plan.radix[s]);
        if (R == 4) stage<T, 4, ROWS, INV>(lds, Nf, B, lgB, ld, Ns, tw, twstride);
        else if (R == 2) stage<T, 2, ROWS, INV>(lds, Nf, B, lgB, ld, Ns, tw, twstride);
        else if (R == 3) stage<T, 3, ROWS, INV>(lds, Nf, B, lgB, ld, Ns, tw, twstride);
        else stage<T, 5, ROWS, INV>(lds, Nf, B, lgB, ld, Ns, tw, twstride);
        Ns *= R;
    }
}

template <typename T>
__device__ inline Cx<T> *lds_tile()
{
    extern __shared__ __align__(16) unsigned char smem[];
    return reinterpret_cast<Cx<T> *>(smem);
}

// ------------------------------------------------------------- z rows
// Real rows of length M -> M/2+1 complex (forward, unnormalized).  One
// block holds B = (LDS tile)/(M/2) rows of one box.
template <typename T>
__global__ __launch_bounds__(kFftThreads) void k_rows_r2c(
    const T *__restrict__ in_all, Cx<T> *__restrict__ out_all, BoxView v,
    const Cx<T> *__restrict__ tw_all)
{
    int bi;
    uint32_t lb;
    if (!box_block(v, bi, lb)) return;
    const BoxDev &bx = v.box[bi];
    const int M = bx.g.M;
    const T *in = in_all + bx.grid0;
    Cx<T> *out = out_all + bx.cgrid0;
    const Cx<T> *tw = tw_all + bx.tab0;
    Cx<T> *lds = lds_tile<T>();
    const int Nf = M / 2, Zc = M / 2 + 1, nrows = M * M;
    const int B = fft_elems<T>() / Nf;
    const int row0 = (int)lb * B;
    const Cx<T> *inc = reinterpret_cast<const Cx<T> *>(in);
    for (int e = threadIdx.x; e < Nf * B; e += kFftThreads) {
        const int l = e / Nf, j = e % Nf;
        const int row = row0 + l;
        lds[l * Nf + j] = row < nrows ? inc[(size_t)row * Nf + j] : Cx<T>{0, 0};
    }
    __syncthreads();
    run_fft<T, true, false>(lds, Nf, B, 0, Nf, bx.half, tw, 2);
    for (int e = threadIdx.x; e < Zc * B; e += kFftThreads) {
        const int l = e / Zc, k = e % Zc;
        const int row = row0 + l;
        if (row >= nrows) continue;
        const Cx<T> zk = lds[l * Nf + (k % Nf)];
        const Cx<T> zm = cconj(lds[l * Nf + ((Nf - k) % Nf)]);
        const Cx<T> s = cadd(zk, zm), d = csub(zk, zm);
        const Cx<T> wd = cmul(tw[k], d);  // W_M^k (Z_k - conj Z_{N/2-k})
        Cx<T> x;
        x.x = (T)0.5 * s.x + (T)0.5 * wd.y;
        x.y = (T)0.5 * s.y - (T)0.5 * wd.x;
        out[(size_t)row * Zc + k] = x;
    }
}

// M/2+1 complex -> real rows of length M (inverse, unnormalized).
template <typename T>
__global__ __launch_bounds__(kFftThreads) void k_rows_c2r(
    const Cx<T> *__restrict__ in_all, T *__restrict__ out_all, BoxView v,
    const Cx<T> *__restrict__ tw_all)
{
    int bi;
    uint32_t lb;
    if (!box_block(v, bi, lb)) return;
    const BoxDev &bx = v.box[bi];
    const int M = bx.g.M;
    const Cx<T> *in = in_all + bx.cgrid0;
    T *out = out_all + bx.grid0;
    const Cx<T> *tw = tw_all + bx.tab0;
    Cx<T> *lds = lds_tile<T>();
    const int Nf = M / 2, Zc = M / 2 + 1, nrows = M * M;
    const int B = fft_elems<T>() / Nf;
    const int row0 = (int)lb * B;
    for (int e = threadIdx.x; e < Nf * B; e += kFftThreads) {
        const int l = e / Nf, k = e % Nf;
        const int row = row0 + l;
        Cx<T> z = {0, 0};
        if (row < nrows) {
            Cx<T> xk = in[(size_t)row * Zc + k];
            Cx<T> xm = in[(size_t)row * Zc + (Nf - k)];
            if (k == 0) { xk.y = 0; xm.y = 0; }  // DC and Nyquist are real
            const Cx<T> c = cconj(xm);
            const Cx<T> s = cadd(xk, c), d = csub(xk, c);
            const Cx<T> wd = cmul(cconj(tw[k]), d);
            z.x = s.x - wd.y;
            z.y = s.y + wd.x;
        }
        lds[l * Nf + k] = z;
    }
    __syncthreads();
    run_fft<T, true, true>(lds, Nf, B, 0, Nf, bx.half, tw, 2);
    Cx<T> *outc = reinterpret_cast<Cx<T> *>(out);
    for (int e = threadIdx.x; e < Nf * B; e += kFftThreads) {
        const int l = e / Nf, j = e % Nf;
        const int row = row0 + l;
        if (row < nrows) outc[(size_t)row * Nf + j] = lds[l * Nf + j];
    }
}

// ------------------------------------------------------------- strided lines
// Complex mesh [x][y][z], z innermost, Zc = M/2+1.  y lines (XAXIS false):
// line L = x*Zc + z, elements Zc apart.  x lines (XAXIS true): line
// L = y*Zc + z, elements M*Zc apart.  Consecutive lines are consecutive in
// memory, so each of the B lines of a block is one coalesced segment per
// element.
template <bool XAXIS>
__device__ inline size_t line_base(int L, int M, int Zc)
{
    return XAXIS ? (size_t)L : (size_t)(L / Zc) * M * Zc + (L % Zc);
}

template <typename T, bool INV, bool XAXIS>
__global__ __launch_bounds__(kFftThreads) void k_lines(
    Cx<T> *__restrict__ data_all, BoxView v, const Cx<T> *__restrict__ tw_all)
{
    int bi;
    uint32_t lb;
    if (!box_block(v, bi, lb)) return;
    const BoxDev &bx = v.box[bi];
    const int M = bx.g.M, Zc = M / 2 + 1, Nf = M;
    const int nlines = M * Zc;
    const size_t stride = XAXIS ? (size_t)M * Zc : (size_t)Zc;
    Cx<T> *data = data_all + bx.cgrid0;
    const Cx<T> *tw = tw_all + bx.tab0;
    Cx<T> *lds = lds_tile<T>();
    const int lgB = lines_log2<T>(Nf), B = 1 << lgB;
    const int line0 = (int)lb * B;
    for (int e = threadIdx.x; e < Nf * B; e += kFftThreads) {
        const int l = e & (B - 1), j = e >> lgB;
        const int L = line0 + l;
        Cx<T> x = {0, 0};
        if (L < nlines) x = data[line_base<XAXIS>(L, M, Zc) + j * stride];
        lds[j * B + l] = x;
    }
    __syncthreads();
    run_fft<T, false, INV>(lds, Nf, B, lgB, B, bx.full, tw, 1);
    for (int e = threadIdx.x; e < Nf * B; e += kFftThreads) {
        const int l = e & (B - 1), j = e >> lgB;
        const int L = line0 + l;
        if (L < nlines) data[line_base<XAXIS>(L, M, Zc) + j * stride] = lds[j * B + l];
    }
}

// x lines: forward FFT, Green's function, inverse FFT.  Line L = y*Zc + z.
template <typename T>
__global__ __launch_bounds__(kFftThreads) void k_lines_green(
    Cx<T> *__restrict__ data_all, BoxView v, const Cx<T> *__restrict__ tw_all,
    const T *__restrict__ gexp_all, const T *__restrict__ gsinc_all)
{
    int bi;
    uint32_t lb;
    if (!box_block(v, bi, lb)) return;
    const BoxDev &bx = v.box[bi];
    const int M = bx.g.M;
    Cx<T> *data = data_all + bx.cgrid0;
    const Cx<T> *tw = tw_all + bx.tab0;
    const T *gexp = gexp_all + bx.tab0;
    const T *gsinc = gsinc_all + bx.tab0;
    Cx<T> *lds = lds_tile<T>();
    const int Nf = M, Zc = M / 2 + 1;
    const int nlines = M * Zc;
    const size_t stride = (size_t)M * Zc;
    const int lgB = lines_log2<T>(Nf), B = 1 << lgB;
    const int line0 = (int)lb * B;
    for (int e = threadIdx.x; e < Nf * B; e += kFftThreads) {
        const int l = e & (B - 1), j = e >> lgB;
        const int L = line0 + l;
        lds[j * B + l] = L < nlines ? data[(size_t)L + j * stride] : Cx<T>{0, 0};
    }
    __syncthreads();
    run_fft<T, false, false>(lds, Nf, B, lgB, B, bx.full, tw, 1);
    for (int e = threadIdx.x; e < Nf * B; e += kFftThreads) {
        const int l = e & (B - 1), j = e >> lgB;
        const int L = line0 + l;
        if (L >= nlines) continue;
        const int y = L / Zc, z = L % Zc;
        const int kx = j > M / 2 ? j - M : j;
        const int ky = y > M / 2 ? y - M : y;
        const int kz = z;
        const int k2i = kx * kx + ky * ky + kz * kz;
        T g = 0;
        if (k2i > 0)
            g = -(gexp[j] * gexp[y] * gexp[z]) * (gsinc[j] * gsinc[y] * gsinc[z]) / (T)k2i;
        Cx<T> x = lds[j * B + l];
        lds[j * B + l] = {x.x * g, x.y * g};
    }
    __syncthreads();
    run_fft<T, false, true>(lds, Nf, B, lgB, B, bx.full, tw, 1);
    for (int e = threadIdx.x; e < Nf * B; e += kFftThreads) {
        const int l = e & (B - 1), j = e >> lgB;
        const int L = line0 + l;
        if (L < nlines) data[(size_t)L + j * stride] = lds[j * B + l];
    }
}

// ------------------------------------------------------------- deposit
// FP64: two passes, both deterministic and free of atomics, every node
// summed in one fixed order:
//
//  k_cic_segments: one lane per sorted particle.  Each lane forms its eight
//    CIC corner contributions m*wx*wy*wz (pm_periodic.c:123-131), and the
//    wave runs a segmented inclusive scan keyed by the cell key (particles of
//    a cell are consecutive after the sort).  A cell that lies inside one
//    wave gets its eight corner sums stored at slot = its first sorted
//    index; for cells that straddle waves every wave stores the sums of its
//    first and of its last segment.
//  k_cic_gather: one thread per mesh node, nodes walked in 4x4x16 bricks so
//    the cell-table reads of neighbouring nodes share cache lines.  A node
//    adds, in a fixed order, the corner sum of each of its eight neighbour
//    cells (a straddling cell adds its wave partials in wave order).
//
// FP32 (Morton keys): k_cic_brick and k_cic_sum, below.

struct Src64 {
    const double4 *p;
    __device__ inline void weights(uint32_t j, const GridInfo &g, double &dx,
                                   double &dy, double &dz, double &m) const
    {
        const double4 q = p[j];
        // pm_periodic.c:99-118: d = fac*x - slab, with the clamped slab.
        dx = g.fac * q.x - cell_coord(g.fac, q.x, g.M);
        dy = g.fac * q.y - cell_coord(g.fac, q.y, g.M);
        dz = g.fac * q.z - cell_coord(g.fac, q.z, g.M);
        m = q.w;
    }
};

struct Src32 {
    const float4 *p;
    __device__ inline void weights(uint32_t j, const GridInfo &g, float &dx,
                                   float &dy, float &dz, float &m) const
    {
        const float facf = (float)g.fac;
        const float4 q = p[j];
        dx = q.x * facf;
        dy = q.y * facf;
        dz = q.z * facf;
        m = q.w;
    }
};

// Waves are counted from each box's first particle, so a box's sums do not
// depend on where the box sits in the particle arrays.
__global__ __launch_bounds__(256) void k_cic_segments(
    Src64 src, const uint32_t *__restrict__ skeys, BoxView v,
    double *__restrict__ cell_sums, double *__restrict__ first_part,
    double *__restrict__ last_part)
{
    int bi;
    uint32_t lb;
    if (!box_block(v, bi, lb)) return;
    const BoxDev &bx = v.box[bi];
    const uint32_t il = lb * 256 + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const uint32_t wl = il >> 6;
    const uint32_t base = bx.p0 + (wl << 6);
    const uint32_t nb = bx.n;
    const bool valid = il < nb;
    const uint32_t i = bx.p0 + il;
    const uint32_t key = valid ? skeys[i] : 0xffffffffu;
    double val[8];
    if (valid) {
        double dx, dy, dz, m;
        src.weights(i, bx.g, dx, dy, dz, m);
        const double wx[2] = {1.0 - dx, dx}, wy[2] = {1.0 - dy, dy}, wz[2] = {1.0 - dz, dz};
#pragma unroll
        for (int c = 0; c < 8; ++c) val[c] = m * wx[c >> 2] * wy[(c >> 1) & 1] * wz[c & 1];
    } else {
#pragma unroll
        for (int c = 0; c < 8; ++c) val[c] = 0;
    }
    // Segmented inclusive scan; keys are non-decreasing along the wave, so
    // lane-o holding the same key means every lane between does too.
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t ko = __shfl_up(key, o, 64);
        const bool take = lane >= o && ko == key;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const double t = __shfl_up(val[c], o, 64);
            val[c] += take ? t : 0.0;
        }
    }
    // Both shuffles run on every lane: a shuffle inside a short-circuit
    // condition would read an inactive source lane.
    const uint32_t knext = __shfl_down(key, 1, 64);
    const uint32_t kprev = __shfl_up(key, 1, 64);
    const bool seg_end = valid && (lane == 63 || knext != key);
    // First lane of this lane's segment inside the wave.
    const uint64_t heads = __ballot(lane == 0 || kprev != key);
    const uint64_t below = heads & ((lane == 63) ? ~0ull : ((2ull << lane) - 1));
    const int start = 63 - __clzll(below);
    const uint32_t prev_key = wl > 0 ? skeys[base - 1] : 0xffffffffu;
    const uint32_t next_key = (wl << 6) + 64 < nb ? skeys[base + 64] : 0xffffffffu;
    if (seg_end) {
        const bool open_left = start == 0 && prev_key == key;
        const bool open_right = lane == 63 && next_key == key;
        const size_t slot = (size_t)bx.wave0 + wl;
        if (!open_left && !open_right) {
#pragma unroll
            for (int c = 0; c < 8; ++c) cell_sums[(size_t)(base + start) * 8 + c] = val[c];
        }
        if (start == 0) {
#pragma unroll
            for (int c = 0; c < 8; ++c) first_part[slot * 8 + c] = val[c];
        }
        if (lane == 63) {
#pragma unroll
            for (int c = 0; c < 8; ++c) last_part[slot * 8 + c] = val[c];
        }
    }
}

constexpr int kBx = 4, kBy = 4, kBz = 16;  // x-major node brick, 256 threads

inline uint32_t gather_blocks(int M)
{
    return (uint32_t)(((M + kBx - 1) / kBx) * ((M + kBy - 1) / kBy) * ((M + kBz - 1) / kBz));
}

__global__ __launch_bounds__(256) void k_cic_gather(
    const uint32_t *__restrict__ off, BoxView v, const double *__restrict__ cell_sums,
    const double *__restrict__ first_part, const double *__restrict__ last_part,
    double *__restrict__ rho_all)
{
    int bi;
    uint32_t lb;
    if (!box_block(v, bi, lb)) return;
    const BoxDev &bx = v.box[bi];
    const GridInfo &g = bx.g;
    const int M = g.M;
    const int t = threadIdx.x;
    const int nbz = (M + kBz - 1) / kBz, nby = (M + kBy - 1) / kBy;
    const int bz = (int)(lb % nbz), by = (int)((lb / nbz) % nby), bxx = (int)(lb / (nbz * nby));
    const int z = bz * kBz + (t % kBz);
    const int y = by * kBy + ((t / kBz) % kBy);
    const int x = bxx * kBx + t / (kBz * kBy);
    if (x >= M || y >= M || z >= M) return;
    const uint32_t p0 = bx.p0;
    // Corner c = (a, bb, cc): the node is the cell's upper neighbour along
    // the axes whose bit is set; cell = node - (a, bb, cc).
    uint32_t cs[8], ce[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int cx = wrap_index(x - (c >> 2), M);
        const int cy = wrap_index(y - ((c >> 1) & 1), M);
        const int cz = wrap_index(z - (c & 1), M);
        load_range(off, bx.key_base + cell_key(g, cx, cy, cz), cs[c], ce[c]);
    }
    // Cells inside one wave: one load each, issued together (the index of an
    // empty or straddling cell is redirected to a valid slot and masked).
    double val[8];
    bool straddle = false;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const bool nonempty = ce[c] > cs[c];
        const bool single = nonempty && ((cs[c] - p0) >> 6) == ((ce[c] - 1 - p0) >> 6);
        straddle |= nonempty && !single;
        const double tv = cell_sums[(single ? (size_t)cs[c] * 8 : 0) + c];
        val[c] = single ? tv : 0.0;
    }
    double sum = 0;
#pragma unroll
    for (int c = 0; c < 8; ++c) sum += val[c];
    if (straddle) {
        for (int c = 0; c < 8; ++c) {
            if (ce[c] <= cs[c]) continue;
            const uint32_t w0 = (cs[c] - p0) >> 6, w1 = (ce[c] - 1 - p0) >> 6;
            if (w0 == w1) continue;
            sum += last_part[((size_t)bx.wave0 + w0) * 8 + c];
            for (uint32_t w = w0 + 1; w <= w1; ++w)
                sum += first_part[((size_t)bx.wave0 + w) * 8 + c];
        }
    }
    rho_all[bx.grid0 + ((size_t)x * M + y) * M + z] = sum;
}

// FP32 deposit over Morton bricks of 16^3 cells (4096 consecutive keys, so
// a brick's particles are one sorted range).  A brick's cells reach the
// nodes at local coordinates 0..u of its corner per axis, u = 16 (at the
// upper mesh edge, when M is not a multiple of 16, u = M minus the corner:
// the node at M is node 0 of the next period).  A brick owns the nodes
// 0..u-1 on every axis; its nodes at u on some axis (its upper faces) are
// owned by the bricks above.
//
//  k_cic_unit: one block per work unit (a brick, or a run of at most H of
//    its particles when it holds more).  The block sums its particles' CIC
//    corners into the 17^3-node tile (a wave-level segmented scan by cell
//    first where a wave holds repeated cells, then the corner sums added in
//    32-bit fixed point with LDS atomics: integer adds commute, so the tile
//    does not depend on the order of waves or particles).  Owned nodes go
//    to the mesh (those on a lower face still without the bricks below);
//    the upper faces to the brick's three face planes.
//  k_cic_shell: one block per brick for its nodes with a local coordinate 0
//    (the lower faces): the brick's own value from the mesh plus, along
//    every subset of the axes where the coordinate is 0, the brick below at
//    its u (periodic) from its planes, summed in a fixed order.
//
// Fixed point: a brick of n particles, none heavier than m_max, puts at
// most n m_max on a node, so with S = the largest power of two where
// S n m_max <= 2^31, a node's sum (each contribution rounded to the
// nearest unit, at most n of them) stays below 2^32.  Scaling by S is exact
// in FP32; a unit is n m_max 2^-31 or less: 4.8e-7 particle masses in a
// brick of 512.
constexpr int kDepBrick = 16;
constexpr int kDepEdge = kDepBrick + 1;
constexpr int kDepNodes = kDepEdge * kDepEdge * kDepEdge;
constexpr uint32_t kDepCells = kDepBrick * kDepBrick * kDepBrick;
constexpr int kDepFace = kDepEdge * kDepEdge;
constexpr int kDepShell = 3 * kDepFace;  // floats of a brick's upper face planes

// Upper node coordinate per axis of the brick at corner o: 16, or M - o at
// the upper mesh edge.
__device__ inline int dep_upper(int o, int M) { return min(kDepBrick, M - o); }

// Position of upper-face node l in a brick's face planes (0: x = u, 17 x 17
// over (y, z); 1: y = u with x < u, over (x, z); 2: z = u with x, y < u,
// over (x, y)).
__device__ inline int upper_slot(int lx, int ly, int lz, int ux, int uy)
{
    if (lx == ux) return ly * kDepEdge + lz;
    if (ly == uy) return kDepFace + lx * kDepEdge + lz;
    return 2 * kDepFace + lx * kDepEdge + ly;
}

// Fixed-point units per unit mass of a brick of n particles (see above);
// the exponent is clamped to FP32's normal range.
__device__ inline float brick_scale(uint32_t n, double m_max)
{
    const double bound = (double)max(n, 1u) * m_max;
    if (!(bound > 0.0)) return 1.0f;
    int ex = 0;
    (void)frexp(bound, &ex);  // bound < 2^ex
    return ldexpf(1.0f, min(120, max(-120, 31 - ex)));
}

// Work units of the FP32 deposit and interpolation: a brick's particles in
// runs of at most H (a brick in a halo core can hold 10^5 particles; one
// block per brick would leave the whole launch waiting for it).  Built per
// evaluation: per brick the number of units (and, for a brick of more than
// one, a slot for its 64-bit fixed-point tile), the scan of those counts,
// then the unit -> brick list.
struct Units {
    const uint32_t *uoff;   // nbricks + 1: first unit of each brick (scanned)
    const uint32_t *brick;  // per unit: its brick (global over the boxes)
    uint32_t H;             // particles per unit at most
    uint32_t nbricks;
};

__global__ void k_brick_units(const uint32_t *__restrict__ off, BoxView v, uint32_t H,
                              uint32_t *__restrict__ ucount, uint32_t *__restrict__ slot_of,
                              uint32_t *__restrict__ slot_brick, uint32_t *__restrict__ nsplit)
{
    const uint32_t gb = blockIdx.x * blockDim.x + threadIdx.x;
    if (gb >= v.map[v.nboxes]) return;
    const int bi = find_box(v.map, v.nboxes, gb);
    const BoxDev &bx = v.box[bi];
    const uint32_t k0 = bx.key_base + (gb - v.map[bi]) * kDepCells;
    const uint32_t n = off[k0 + kDepCells] - off[k0];
    const uint32_t U = n > H ? (n + H - 1) / H : 1u;
    ucount[gb] = U;
    if (U > 1) {
        // The slot only says where the brick's tile lives, so the order in
        // which bricks get slots does not matter.
        const uint32_t slot = atomicAdd(nsplit, 1u);
        slot_of[gb] = slot;
        slot_brick[slot] = gb;
    }
}

__global__ void k_brick_units_fill(const uint32_t *__restrict__ uoff, uint32_t nbricks,
                                   uint32_t *__restrict__ brick)
{
    const uint32_t gb = blockIdx.x * blockDim.x + threadIdx.x;
    if (gb >= nbricks) return;
    for (uint32_t u = uoff[gb]; u < uoff[gb + 1]; ++u) brick[u] = gb;
}

// Unit u (wave-uniform): its box, brick inside the box, part and number of
// parts; false past the last unit.
__device__ inline bool unit_of(const Units &un, const BoxView &v, uint32_t u, int &bi,
                               uint32_t &lb, uint32_t &part, uint32_t &nparts)
{
    if (u >= __builtin_amdgcn_readfirstlane(un.uoff[un.nbricks])) return false;
    const uint32_t gb = __builtin_amdgcn_readfirstlane(un.brick[u]);
    bi = __builtin_amdgcn_readfirstlane(find_box(v.map, v.nboxes, gb));
    lb = gb - v.map[bi];
    const uint32_t u0 = __builtin_amdgcn_readfirstlane(un.uoff[gb]);
    part = u - u0;
    nparts = __builtin_amdgcn_readfirstlane(un.uoff[gb + 1]) - u0;
    return true;
}

// The brick's tile (32-bit fixed point, in LDS; inv = 1/S) to the mesh
// and the face planes: owned nodes to the mesh, upper faces to the planes;
// nodes past u get nothing.  Thread t takes the node column (ly, lz) =
// (t / 16, t % 16) along x (the mesh rows of 16 floats at a time), then the
// nodes at ly = 16 or lz = 16.  Needs 256 threads.
__device__ inline void write_brick(const uint32_t *s_node, float inv, int x0, int y0, int z0,
                                   int M, float *__restrict__ pt, float *__restrict__ rho)
{
    const int ux = dep_upper(x0, M), uy = dep_upper(y0, M), uz = dep_upper(z0, M);
    const int ly = threadIdx.x >> 4, lz = threadIdx.x & 15;
    if (ly <= uy && lz <= uz) {
        const bool yz_own = ly < uy && lz < uz;
        const size_t mm = (size_t)M * M;
        float *out = rho + (size_t)x0 * mm + (size_t)(y0 + ly) * M + (z0 + lz);
        for (int lx = 0; lx <= ux; ++lx, out += mm) {
            const float val = (float)s_node[(lx * kDepEdge + ly) * kDepEdge + lz] * inv;
            if (yz_own && lx < ux)
                *out = val;
            else
                pt[upper_slot(lx, ly, lz, ux, uy)] = val;
        }
    }
    // ly = 16 (only on the upper y face when uy = 16), then lz = 16 with
    // ly < 16 (the upper z face when uz = 16).
    for (int e = threadIdx.x; e < kDepFace + kDepEdge * kDepBrick; e += 256) {
        int lx, ly2, lz2;
        if (e < kDepFace) {
            lx = e / kDepEdge; ly2 = kDepBrick; lz2 = e - lx * kDepEdge;
        } else {
            const int f = e - kDepFace;
            lx = f / kDepBrick; ly2 = f - lx * kDepBrick; lz2 = kDepBrick;
        }
        if (lx > ux || ly2 > uy || lz2 > uz) continue;
        pt[upper_slot(lx, ly2, lz2, ux, uy)] =
            (float)s_node[(lx * kDepEdge + ly2) * kDepEdge + lz2] * inv;
    }
}

// One block per work unit.  Units are taken in launch order (round robin
// over the XCDs): a split brick's units then spread over the chip, and no
// two units share data but their split brick's tile.
__global__ __launch_bounds__(256) void k_cic_unit(Src32 src, const uint32_t *__restrict__ skeys,
                                                  const uint32_t *__restrict__ off, BoxView v,
                                                  Units un, const uint32_t *__restrict__ slot_of,
                                                  uint32_t *__restrict__ htile,
                                                  float *__restrict__ partial,
                                                  float *__restrict__ rho_all)
{
    __shared__ uint32_t s_node[kDepNodes];
    int bi;
    uint32_t lb, part, nparts;
    if (!unit_of(un, v, blockIdx.x, bi, lb, part, nparts)) return;
    const BoxDev &bx = v.box[bi];
    const GridInfo &g = bx.g;
    const int M = g.M;
    const uint32_t k0 = lb * kDepCells;
    int x0, y0, z0;
    key_cell(g, k0, x0, y0, z0);
    if (x0 >= M || y0 >= M || z0 >= M) return;  // outside the mesh: no cells
    for (int e = threadIdx.x; e < kDepNodes; e += 256) s_node[e] = 0u;
    const uint32_t b_begin = off[bx.key_base + k0];
    const uint32_t b_end = off[bx.key_base + k0 + kDepCells];
    const uint32_t p_begin = b_begin + part * un.H;
    const uint32_t p_end = min(b_end, p_begin + un.H);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const float scale = brick_scale(b_end - b_begin, bx.dep_mmax);
    for (uint32_t base = p_begin; base < p_end; base += 256) {
        const uint32_t i = base + threadIdx.x;
        const bool valid = i < p_end;
        const uint32_t key = valid ? skeys[i] : 0xffffffffu;
        float val[8];
        if (valid) {
            float dx, dy, dz, m;
            src.weights(i, g, dx, dy, dz, m);
            const float wx[2] = {1.0f - dx, dx}, wy[2] = {1.0f - dy, dy}, wz[2] = {1.0f - dz, dz};
#pragma unroll
            for (int c = 0; c < 8; ++c) val[c] = m * wx[c >> 2] * wy[(c >> 1) & 1] * wz[c & 1];
        } else {
#pragma unroll
            for (int c = 0; c < 8; ++c) val[c] = 0.f;
        }
        const uint32_t kprev = __shfl_up(key, 1, 64);
        const uint32_t knext = __shfl_down(key, 1, 64);
        // Segmented inclusive scan by cell (keys non-decreasing along the
        // wave), only where some cell repeats: then one lane per cell adds.
        if (__ballot(valid && lane > 0 && kprev == key) != 0ull) {
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t ko = __shfl_up(key, o, 64);
                const bool take = lane >= o && ko == key;
#pragma unroll
                for (int c = 0; c < 8; ++c) {
                    const float t = __shfl_up(val[c], o, 64);
                    val[c] += take ? t : 0.f;
                }
            }
        }
        if (valid && (lane == 63 || knext != key)) {
            int cx, cy, cz;
            key_cell(g, key - bx.key_base, cx, cy, cz);
            const int n0 = ((cx - x0) * kDepEdge + (cy - y0)) * kDepEdge + (cz - z0);
#pragma unroll
            for (int c = 0; c < 8; ++c)
                atomicAdd(&s_node[n0 + ((c >> 2) * kDepEdge + ((c >> 1) & 1)) * kDepEdge + (c & 1)],
                          __float2uint_rn(val[c] * scale));
        }
    }
    __syncthreads();
    if (nparts > 1) {
        // A brick of several units: every unit adds its tile into the
        // brick's tile in global memory (integer adds: the sum does not
        // depend on the order); k_cic_split writes the brick.
        uint32_t *tile = htile + (size_t)slot_of[v.map[bi] + lb] * kDepNodes;
        for (int e = threadIdx.x; e < kDepNodes; e += 256) {
            const uint32_t t = s_node[e];
            if (t) atomicAdd(&tile[e], t);
        }
        return;
    }
    write_brick(s_node, 1.0f / scale, x0, y0, z0, M,
                partial + (size_t)(v.map[bi] + lb) * kDepShell, rho_all + bx.grid0);
}

// Bricks split over several units: the summed tile (complete once
// k_cic_unit has finished) to the mesh and face planes; the tile is cleared
// for the next deposit.
__global__ __launch_bounds__(256) void k_cic_split(BoxView v, const uint32_t *__restrict__ off,
                                                   const uint32_t *__restrict__ nsplit,
                                                   const uint32_t *__restrict__ slot_brick,
                                                   uint32_t *__restrict__ htile,
                                                   float *__restrict__ partial,
                                                   float *__restrict__ rho_all)
{
    __shared__ uint32_t s_node[kDepNodes];
    const uint32_t slot = blockIdx.x;
    if (slot >= __builtin_amdgcn_readfirstlane(nsplit[0])) return;
    const uint32_t gb = __builtin_amdgcn_readfirstlane(slot_brick[slot]);
    const int bi = __builtin_amdgcn_readfirstlane(find_box(v.map, v.nboxes, gb));
    const uint32_t lb = gb - v.map[bi];
    const BoxDev &bx = v.box[bi];
    int x0, y0, z0;
    key_cell(bx.g, lb * kDepCells, x0, y0, z0);
    uint32_t *tile = htile + (size_t)slot * kDepNodes;
    for (int e = threadIdx.x; e < kDepNodes; e += 256) {
        s_node[e] = tile[e];
        tile[e] = 0u;
    }
    const uint32_t k0 = bx.key_base + lb * kDepCells;
    const float scale = brick_scale(off[k0 + kDepCells] - off[k0], bx.dep_mmax);
    __syncthreads();
    write_brick(s_node, 1.0f / scale, x0, y0, z0, bx.g.M, partial + (size_t)gb * kDepShell,
                rho_all + bx.grid0);
}

__global__ __launch_bounds__(256) void k_cic_shell(const float *__restrict__ partial, BoxView v,
                                                   float *__restrict__ rho_all)
{
    int bi;
    uint32_t lb;
    if (!box_block(v, bi, lb)) return;
    const BoxDev &bx = v.box[bi];
    const GridInfo &g = bx.g;
    const int M = g.M;
    int o[3];
    key_cell(g, lb * kDepCells, o[0], o[1], o[2]);
    if (o[0] >= M || o[1] >= M || o[2] >= M) return;
    // Per axis: the brick below (the last one of the mesh at the edge) and
    // the local coordinate the node has in its tile.
    const int last = ((M - 1) / kDepBrick) * kDepBrick;
    int below[3], lbel[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        below[d] = o[d] == 0 ? last : o[d] - kDepBrick;
        lbel[d] = o[d] == 0 ? M - last : kDepBrick;
    }
    const int ux = dep_upper(o[0], M), uy = dep_upper(o[1], M), uz = dep_upper(o[2], M);
    // Face planes of the brick below along the set bits of c (bit 2: x,
    // 1: y, 0: z).
    const float *tile[8];
    tile[0] = nullptr;
#pragma unroll
    for (int c = 1; c < 8; ++c) {
        const uint32_t brick =
            cell_key(g, (c & 4) ? below[0] : o[0], (c & 2) ? below[1] : o[1],
                     (c & 1) ? below[2] : o[2]) / kDepCells;
        tile[c] = partial + (size_t)(v.map[bi] + brick) * kDepShell;
    }
    float *rho = rho_all + bx.grid0;
    // Lower-face nodes: x = 0 (16x16), then y = 0 with x > 0 (15x16), then
    // z = 0 with x, y > 0 (15x15).
    constexpr int kFx = kDepBrick * kDepBrick, kFy = (kDepBrick - 1) * kDepBrick;
    constexpr int kFaces = kFx + kFy + (kDepBrick - 1) * (kDepBrick - 1);
    for (int e = threadIdx.x; e < kFaces; e += 256) {
        int l[3];
        if (e < kFx) {
            l[0] = 0; l[1] = e / kDepBrick; l[2] = e % kDepBrick;
        } else if (e < kFx + kFy) {
            const int f = e - kFx;
            l[0] = 1 + f / kDepBrick; l[1] = 0; l[2] = f % kDepBrick;
        } else {
            const int f = e - kFx - kFy;
            l[0] = 1 + f / (kDepBrick - 1); l[1] = 1 + f % (kDepBrick - 1); l[2] = 0;
        }
        if (l[0] >= ux || l[1] >= uy || l[2] >= uz) continue;
        float *node = rho + ((size_t)(o[0] + l[0]) * M + (o[1] + l[1])) * M + (o[2] + l[2]);
        float sum = *node;  // the brick's own particles
#pragma unroll
        for (int c = 1; c < 8; ++c) {
            if (((c & 4) && l[0]) || ((c & 2) && l[1]) || ((c & 1) && l[2])) continue;
            // The node in the brick below: on its upper face along the axes of c.
            const int ax = (c & 4) ? lbel[0] : l[0], ay = (c & 2) ? lbel[1] : l[1],
                      az = (c & 1) ? lbel[2] : l[2];
            sum += tile[c][upper_slot(ax, ay, az, (c & 4) ? lbel[0] : ux,
                                      (c & 2) ? lbel[1] : uy)];
        }
        *node = sum;
    }
}

// ------------------------------------------------------------- interpolation
// The 4-point stencil (pm_periodic.c:260-267) at the node s+a of a line of
// potential values v[k] = phi(s-2+k), k = 0..5.
template <typename T>
__device__ inline T stencil_line(const T *v, int a, T fac)
{
    return fac * (((T)4.0 / (T)3) * (v[a + 1] - v[a + 3]) - ((T)1.0 / (T)6) * (v[a] - v[a + 4]));
}

// The 72 potential values the 8 CIC nodes' stencils read, each once: along
// x for the 4 (y, z) node pairs, along y for the 4 (x, z), along z for the
// 4 (x, y).
template <typename T>
struct Lines {
    T x[2][2][6], y[2][2][6], z[2][2][6];
};

template <typename T>
__device__ inline void load_lines(const T *__restrict__ phi, int M, int sx, int sy,
                                  int sz, Lines<T> &L)
{
    const size_t MM = (size_t)M;
    int xs[6], ys[6], zs[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        xs[k] = wrap_index(sx - 2 + k, M);
        ys[k] = wrap_index(sy - 2 + k, M);
        zs[k] = wrap_index(sz - 2 + k, M);
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                L.x[a][b][k] = phi[((size_t)xs[k] * MM + ys[2 + a]) * MM + zs[2 + b]];
                L.y[a][b][k] = phi[((size_t)xs[2 + a] * MM + ys[k]) * MM + zs[2 + b]];
                L.z[a][b][k] = phi[((size_t)xs[2 + a] * MM + ys[2 + b]) * MM + zs[k]];
            }
}

// pm_periodic.c:297-306: the stencil at the 8 nodes with the CIC weights,
// same term order.  Node (a, b, c) = (sx+a, sy+b, sz+c).
template <typename T, int DIM>
__device__ inline T interp_dim(const Lines<T> &L, T dx, T dy, T dz, T fac)
{
    auto st = [&](int a, int b, int c) -> T {
        if (DIM == 0) return stencil_line(L.x[b][c], a, fac);
        if (DIM == 1) return stencil_line(L.y[a][c], b, fac);
        return stencil_line(L.z[a][b], c, fac);
    };
    T acc = st(0, 0, 0) * ((T)1.0 - dx) * ((T)1.0 - dy) * ((T)1.0 - dz);
    acc += st(0, 1, 0) * ((T)1.0 - dx) * dy * ((T)1.0 - dz);
    acc += st(0, 0, 1) * ((T)1.0 - dx) * ((T)1.0 - dy) * dz;
    acc += st(0, 1, 1) * ((T)1.0 - dx) * dy * dz;
    acc += st(1, 0, 0) * (dx) * ((T)1.0 - dy) * ((T)1.0 - dz);
    acc += st(1, 1, 0) * (dx) * dy * ((T)1.0 - dz);
    acc += st(1, 0, 1) * (dx) * ((T)1.0 - dy) * dz;
    acc += st(1, 1, 1) * (dx) * dy * dz;
    return acc;
}

__global__ __launch_bounds__(256) void k_interp64(const double *__restrict__ phi_all,
                                                  const double4 *__restrict__ spos,
                                                  const uint32_t *__restrict__ perm,
                                                  BoxView v,
                                                  double4 *__restrict__ acc)
{
    int bi;
    uint32_t lb;
    if (!box_block(v, bi, lb)) return;
    const BoxDev &bx = v.box[bi];
    const uint32_t il = lb * 256 + threadIdx.x;
    if (il >= bx.n) return;
    const uint32_t i = bx.p0 + il;
    const GridInfo &g = bx.g;
    const int M = g.M;
    const double fac = bx.pm_fac;
    const double *phi = phi_all + bx.grid0;
    const double4 q = spos[i];
    const int sx = cell_coord(g.fac, q.x, M), sy = cell_coord(g.fac, q.y, M),
              sz = cell_coord(g.fac, q.z, M);
    const double dx = g.fac * q.x - sx, dy = g.fac * q.y - sy, dz = g.fac * q.z - sz;
    Lines<double> L;
    load_lines(phi, M, sx, sy, sz, L);
    double4 a;
    a.x = interp_dim<double, 0>(L, dx, dy, dz, fac);
    a.y = interp_dim<double, 1>(L, dx, dy, dz, fac);
    a.z = interp_dim<double, 2>(L, dx, dy, dz, fac);
    a.w = 0.0;
    acc[perm[i]] = a;
}

// FP32: one block per Morton brick of 16^3 cells (4096 consecutive keys,
// so the brick's particles are one sorted range).  The block stages the
// potential at the 21^3 nodes its particles' stencils can reach (2 below,
// 3 above the brick, periodic) in LDS once; every particle then reads its
// 72 values from there.  The same stencil and CIC weights as the FP64
// kernel, evaluated as nested linear interpolations with fused
// multiply-adds (FP32 needs no reference rounding order).
constexpr int kBrick = 16;
constexpr int kBrickNodes = kBrick + 5;

// Unscaled stencil at node a of a line v[0..5] (v[k] = phi(s-2+k)).
__device__ inline float stencil32(const float *v, int a)
{
    return fmaf(4.0f / 3.0f, v[a + 1] - v[a + 3], -(1.0f / 6.0f) * (v[a] - v[a + 4]));
}

__device__ inline float lerp32(float g0, float g1, float d) { return fmaf(d, g1 - g0, g0); }

// Component DIM at offsets (dx, dy, dz) inside the cell: the stencil at the
// 8 CIC nodes, weighted (1-d) / d along each axis.
template <int DIM>
__device__ inline float interp32(const Lines<float> &L, float dx, float dy, float dz)
{
    float g[2][2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int c = 0; c < 2; ++c)
                g[a][b][c] = DIM == 0 ? stencil32(L.x[b][c], a)
                           : DIM == 1 ? stencil32(L.y[a][c], b)
                                      : stencil32(L.z[a][b], c);
    const float s00 = lerp32(g[0][0][0], g[0][0][1], dz), s01 = lerp32(g[0][1][0], g[0][1][1], dz);
    const float s10 = lerp32(g[1][0][0], g[1][0][1], dz), s11 = lerp32(g[1][1][0], g[1][1][1], dz);
    return lerp32(lerp32(s00, s01, dy), lerp32(s10, s11, dy), dx);
}

__global__ __launch_bounds__(256) void k_interp32(const float *__restrict__ phi_all,
                                                  const float4 *__restrict__ spos,
                                                  const uint32_t *__restrict__ skeys,
                                                  const uint32_t *__restrict__ perm,
                                                  const uint32_t *__restrict__ off, BoxView v,
                                                  Units un, uint32_t max_units,
                                                  float4 *__restrict__ acc)
{
    constexpr int E = kBrickNodes;
    constexpr uint32_t kCells = kBrick * kBrick * kBrick;
    __shared__ float s_phi[E * E * E];
    // Neighbouring bricks stage overlapping potential nodes: XCD k takes the
    // k-th contiguous eighth of the units actually present, so neighbours
    // share an L2 (the grid has at least 8 ceil(total/8) blocks).
    const uint32_t total = __builtin_amdgcn_readfirstlane(un.uoff[un.nbricks]);
    const uint32_t per = (total + kXcds - 1) / kXcds;
    const uint32_t ul = (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
    (void)max_units;
    int bi;
    uint32_t lb, part, nparts;
    if (!unit_of(un, v, ul, bi, lb, part, nparts)) return;
    const BoxDev &bx = v.box[bi];
    const GridInfo &g = bx.g;
    const uint32_t k0 = lb * kCells;
    const uint32_t b_end = off[bx.key_base + k0 + kCells];
    const uint32_t p_begin = off[bx.key_base + k0] + part * un.H;
    const uint32_t p_end = min(b_end, p_begin + un.H);
    if (p_begin >= p_end) return;  // empty brick (or one outside the mesh)
    const int M = g.M;
    int x0, y0, z0;
    key_cell(g, k0, x0, y0, z0);
    const float *phi = phi_all + bx.grid0;
    // A thread owns node columns (iy, iz) and walks each along x.
    for (int col = threadIdx.x; col < E * E; col += 256) {
        const int iy = col / E, iz = col - iy * E;
        const size_t yz = (size_t)wrap_index(y0 - 2 + iy, M) * M + wrap_index(z0 - 2 + iz, M);
        const size_t MM = (size_t)M * M;
#pragma unroll
        for (int ix = 0; ix < E; ++ix)
            s_phi[(ix * E + iy) * E + iz] = phi[(size_t)wrap_index(x0 - 2 + ix, M) * MM + yz];
    }
    __syncthreads();
    const float facf = (float)g.fac, fac = (float)bx.pm_fac;
    for (uint32_t i = p_begin + threadIdx.x; i < p_end; i += 256) {
        const float4 q = spos[i];
        int sx, sy, sz;
        key_cell(g, skeys[i] - bx.key_base, sx, sy, sz);
        const int lx = sx - x0, ly = sy - y0, lz = sz - z0;  // 0..kBrick-1
        Lines<float> L;
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int k = 0; k < 6; ++k) {
                    L.x[a][b][k] = s_phi[((lx + k) * E + ly + 2 + a) * E + lz + 2 + b];
                    L.y[a][b][k] = s_phi[((lx + 2 + a) * E + ly + k) * E + lz + 2 + b];
                    L.z[a][b][k] = s_phi[((lx + 2 + a) * E + ly + 2 + b) * E + lz + k];
                }
        const float dx = q.x * facf, dy = q.y * facf, dz = q.z * facf;
        acc[perm[i]] = make_float4(fac * interp32<0>(L, dx, dy, dz), fac * interp32<1>(L, dx, dy, dz),
                                   fac * interp32<2>(L, dx, dy, dz), 0.f);
    }
}

template <typename T>
size_t lds_bytes() { return (size_t)fft_bytes<T>(); }

template <typename T, typename K, typename... Args>
int launch_fft_kernel(K kernel, unsigned grid, hipStream_t s, Args... args)
{
    if (grid == 0) return CUCO_OK;
    kernel<<<grid, kFftThreads, lds_bytes<T>(), s>>>(args...);
    CUCO_LAUNCH_CHECK();
    return CUCO_OK;
}

// Byte layout of the FP32 deposit scratch: the brick face planes, the
// work-unit tables, and the 32-bit tiles of bricks split into several units
// (at most capacity/H + 1 such bricks: each holds more than H particles).
struct DepLayout {
    uint32_t nbricks, H, max_units, max_slots;
    size_t partial, ucount, slot_of, brick, scan, ctr, slot_brick, htile, bytes;
};

DepLayout dep_layout(const BoxSet &bs)
{
    DepLayout d{};
    d.nbricks = bs.total[kMapBrick];
    const uint32_t cap = std::max(bs.capacity, 1u);
    // At most 8192 split bricks whatever the particle count, and at least
    // 2048 particles (8 passes of the block) per unit.
    uint32_t H = std::max<uint32_t>(2048u, (cap + 8191u) / 8192u);
    d.H = (H + 255u) / 256u * 256u;
    d.max_slots = cap / d.H + 1;
    d.max_units = d.nbricks + d.max_slots;
    size_t o = 0;
    auto take = [&o](size_t bytes) {
        const size_t at = o;
        o = (o + bytes + 255) / 256 * 256;
        return at;
    };
    d.partial = take(sizeof(float) * kDepShell * (size_t)d.nbricks);
    d.ucount = take(sizeof(uint32_t) * ((size_t)d.nbricks + 1));
    d.slot_of = take(sizeof(uint32_t) * (size_t)d.nbricks);
    d.brick = take(sizeof(uint32_t) * (size_t)d.max_units);
    d.scan = take(sizeof(uint32_t) * scan_scratch_words(d.nbricks + 1));
    d.ctr = take(16);
    d.slot_brick = take(sizeof(uint32_t) * (size_t)d.max_slots);
    d.htile = take(sizeof(uint32_t) * kDepNodes * (size_t)d.max_slots);
    d.bytes = o;
    return d;
}

Units units_view(const DepLayout &d, const void *scratch)
{
    const char *base = static_cast<const char *>(scratch);
    return Units{reinterpret_cast<const uint32_t *>(base + d.ucount),
                 reinterpret_cast<const uint32_t *>(base + d.brick), d.H, d.nbricks};
}

}  // namespace

bool fft_size_supported(int M)
{
    if (M < 4 || M % 2 != 0 || M > fft_elems<double>()) return false;
    return make_plan(M).nstages >= 0 && make_plan(M / 2).nstages >= 0;
}

template <typename T>
int pm_deposit(const void *spos, const uint32_t *skeys, const uint32_t *off,
               const BoxSet &bs, int mode, T *rho, void *scratch, hipStream_t s)
{
    if constexpr (sizeof(T) == 4) {
        const DepLayout d = dep_layout(bs);
        if (d.nbricks == 0) return CUCO_OK;
        char *base = static_cast<char *>(scratch);
        float *partial = reinterpret_cast<float *>(base + d.partial);
        uint32_t *ucount = reinterpret_cast<uint32_t *>(base + d.ucount);
        uint32_t *slot_of = reinterpret_cast<uint32_t *>(base + d.slot_of);
        uint32_t *brick = reinterpret_cast<uint32_t *>(base + d.brick);
        uint32_t *ctr = reinterpret_cast<uint32_t *>(base + d.ctr);
        int rc = launch_fill_u32(ctr, 0u, 1, s);
        if (rc) return rc;
        uint32_t *slot_brick = reinterpret_cast<uint32_t *>(base + d.slot_brick);
        uint32_t *htile = reinterpret_cast<uint32_t *>(base + d.htile);
        k_brick_units<<<grid_for(d.nbricks, 256), 256, 0, s>>>(
            off, bs.view(mode, kMapBrick), d.H, ucount, slot_of, slot_brick, ctr);
        CUCO_LAUNCH_CHECK();
        // counts -> first unit per brick, in place; the total after the last
        rc = exclusive_scan_u32(ucount, ucount, d.nbricks,
                                reinterpret_cast<uint32_t *>(base + d.scan), ucount + d.nbricks, s);
        if (rc) return rc;
        k_brick_units_fill<<<grid_for(d.nbricks, 256), 256, 0, s>>>(ucount, d.nbricks, brick);
        CUCO_LAUNCH_CHECK();
        k_cic_unit<<<d.max_units, 256, 0, s>>>(
            Src32{static_cast<const float4 *>(spos)}, skeys, off, bs.view(mode, kMapBrick),
            units_view(d, scratch), slot_of, htile, partial, reinterpret_cast<float *>(rho));
        CUCO_LAUNCH_CHECK();
        k_cic_split<<<d.max_slots, 256, 0, s>>>(bs.view(mode, kMapBrick), off, ctr, slot_brick,
                                                htile, partial, reinterpret_cast<float *>(rho));
        CUCO_LAUNCH_CHECK();
        k_cic_shell<<<xcd_grid(bs.total[kMapBrick]), 256, 0, s>>>(
            partial, bs.view(mode, kMapBrick), reinterpret_cast<float *>(rho));
        CUCO_LAUNCH_CHECK();
    } else {
        double *cell_sums = static_cast<double *>(scratch);
        double *first_part = cell_sums + (size_t)bs.capacity * 8;
        double *last_part = first_part + (size_t)bs.wave_slots * 8;
        if (bs.total[kMapPart] > 0) {
            k_cic_segments<<<xcd_grid(bs.total[kMapPart]), 256, 0, s>>>(
                Src64{static_cast<const double4 *>(spos)}, skeys, bs.view(mode, kMapPart),
                cell_sums, first_part, last_part);
            CUCO_LAUNCH_CHECK();
        }
        k_cic_gather<<<xcd_grid(bs.total[kMapGatherX]), 256, 0, s>>>(
            off, bs.view(mode, kMapGatherX), cell_sums, first_part, last_part, rho);
        CUCO_LAUNCH_CHECK();
    }
    return CUCO_OK;
}

int pm_deposit_scratch_init(const BoxSet &bs, void *scratch, hipStream_t s)
{
    // The split-brick tiles start at zero; every deposit leaves them at zero
    // again.
    const DepLayout d = dep_layout(bs);
    char *base = static_cast<char *>(scratch);
    return launch_fill_u32(reinterpret_cast<uint32_t *>(base + d.htile), 0u,
                           (d.bytes - d.htile) / sizeof(uint32_t), s);
}

size_t pm_deposit_scratch_bytes(const BoxSet &bs, int fp64)
{
    if (!fp64) return dep_layout(bs).bytes;
    return sizeof(double) * 8 * ((size_t)bs.capacity + (size_t)bs.wave_slots * 2 + 64);
}


uint32_t pm_blocks(int M, int map, int fft_split)
{
    if (map >= kMapRowsR && map < kMapBrick) {  // a register-FFT size's own map
        const bool rows = map < kMapLinesR;
        const int size = map - (rows ? kMapRowsR : kMapLinesR);
        return fft32_mask_bit(M) == (1u << size) ? fft32_blocks(M, rows, fft_split) : 0u;
    }
    const uint32_t Zc = (uint32_t)(M / 2 + 1);
    switch (map) {
    case kMapGatherX: return gather_blocks(M);
    case kMapRowsD: { const uint32_t B = fft_elems<double>() / (M / 2); return ((uint32_t)M * M + B - 1) / B; }
    case kMapRowsF: {
        if (fft32_supported(M)) return 0;
        const uint32_t B = fft_elems<float>() / (M / 2);
        return ((uint32_t)M * M + B - 1) / B;
    }
    case kMapLinesD: { const uint32_t B = 1u << lines_log2<double>(M); return ((uint32_t)M * Zc + B - 1) / B; }
    case kMapLinesF: {
        if (fft32_supported(M)) return 0;
        const uint32_t B = 1u << lines_log2<float>(M);
        return ((uint32_t)M * Zc + B - 1) / B;
    }
    case kMapBrick: {  // 16^3-cell Morton bricks of the padded key cube
        int bits = 0;
        while ((1 << bits) < M) ++bits;
        return bits >= 4 ? 1u << (3 * (bits - 4)) : 1u;
    }
    default: return 0;
    }
}

template <typename T>
int pm_fft_forward(const T *rho, Cx<T> *c, const BoxSet &bs, const PmTables &tb,
                   hipStream_t s)
{
    const bool d = sizeof(T) == 8;
    const int mr = d ? kMapRowsD : kMapRowsF, ml = d ? kMapLinesD : kMapLinesF;
    const Cx<T> *tw = static_cast<const Cx<T> *>(tb.tw);
    int rc = launch_fft_kernel<T>(k_rows_r2c<T>, xcd_grid(bs.total[mr]), s, rho, c,
                                  bs.view(0, mr), tw);
    if (rc) return rc;
    rc = launch_fft_kernel<T>(k_lines<T, false, false>, xcd_grid(bs.total[ml]), s, c,
                              bs.view(0, ml), tw);
    if (rc) return rc;
    if constexpr (sizeof(T) == 4) return fft32_forward(rho, c, bs, tb, s);
    return CUCO_OK;
}

template <typename T>
int pm_green_x(Cx<T> *c, const BoxSet &bs, const PmTables &tb, hipStream_t s)
{
    const int ml = sizeof(T) == 8 ? kMapLinesD : kMapLinesF;
    const int rc = launch_fft_kernel<T>(k_lines_green<T>, xcd_grid(bs.total[ml]), s, c,
                                        bs.view(0, ml), static_cast<const Cx<T> *>(tb.tw),
                                        static_cast<const T *>(tb.gexp),
                                        static_cast<const T *>(tb.gsinc));
    if (rc) return rc;
    if constexpr (sizeof(T) == 4) return fft32_green_x(c, bs, tb, s);
    return CUCO_OK;
}

template <typename T>
int pm_fft_inverse(Cx<T> *c, T *phi, const BoxSet &bs, const PmTables &tb,
                   hipStream_t s)
{
    const bool d = sizeof(T) == 8;
    const int mr = d ? kMapRowsD : kMapRowsF, ml = d ? kMapLinesD : kMapLinesF;
    const Cx<T> *tw = static_cast<const Cx<T> *>(tb.tw);
    int rc = launch_fft_kernel<T>(k_lines<T, true, false>, xcd_grid(bs.total[ml]), s, c,
                                  bs.view(0, ml), tw);
    if (rc) return rc;
    rc = launch_fft_kernel<T>(k_rows_c2r<T>, xcd_grid(bs.total[mr]), s,
                              (const Cx<T> *)c, phi, bs.view(0, mr), tw);
    if (rc) return rc;
    if constexpr (sizeof(T) == 4) return fft32_inverse(c, phi, bs, tb, s);
    return CUCO_OK;
}

template <typename T>
int pm_interp(const T *phi, const void *spos, const uint32_t *skeys,
              const uint32_t *perm, const uint32_t *off, const BoxSet &bs, int mode,
              const void *dep_scratch, void *acc, hipStream_t s)
{
    if (bs.total[kMapPart] == 0) return CUCO_OK;
    if (sizeof(T) == 8) {
        k_interp64<<<xcd_grid(bs.total[kMapPart]), 256, 0, s>>>(
            reinterpret_cast<const double *>(phi), static_cast<const double4 *>(spos),
            perm, bs.view(mode, kMapPart), static_cast<double4 *>(acc));
    } else {
        const DepLayout d = dep_layout(bs);
        if (d.nbricks == 0) return CUCO_OK;
        k_interp32<<<xcd_grid(d.max_units), 256, 0, s>>>(
            reinterpret_cast<const float *>(phi), static_cast<const float4 *>(spos),
            skeys, perm, off, bs.view(mode, kMapBrick), units_view(d, dep_scratch),
            d.max_units, static_cast<float4 *>(acc));
    }
    CUCO_LAUNCH_CHECK();
    return CUCO_OK;
}

template <typename T>
int pm_fft_x(Cx<T> *c, const BoxSet &bs, bool inverse, const PmTables &tb,
             hipStream_t s)
{
    const int ml = sizeof(T) == 8 ? kMapLinesD : kMapLinesF;
    const Cx<T> *tw = static_cast<const Cx<T> *>(tb.tw);
    const int rc = inverse
                       ? launch_fft_kernel<T>(k_lines<T, true, true>, xcd_grid(bs.total[ml]), s,
                                              c, bs.view(0, ml), tw)
                       : launch_fft_kernel<T>(k_lines<T, false, true>, xcd_grid(bs.total[ml]), s,
                                              c, bs.view(0, ml), tw);
    if (rc) return rc;
    if constexpr (sizeof(T) == 4) return fft32_x(c, bs, inverse, tb, s);
    return CUCO_OK;
}

#define CUCO_PM_INST(T)                                                              \
    template int pm_fft_x<T>(Cx<T> *, const BoxSet &, bool, const PmTables &,        \
                             hipStream_t);                                           \
    template int pm_deposit<T>(const void *, const uint32_t *, const uint32_t *,     \
                               const BoxSet &, int, T *, void *, hipStream_t);       \
    template int pm_fft_forward<T>(const T *, Cx<T> *, const BoxSet &,               \
                                   const PmTables &, hipStream_t);                   \
    template int pm_green_x<T>(Cx<T> *, const BoxSet &, const PmTables &,            \
                               hipStream_t);                                         \
    template int pm_fft_inverse<T>(Cx<T> *, T *, const BoxSet &, const PmTables &,   \
                                   hipStream_t);                                     \
    template int pm_interp<T>(const T *, const void *, const uint32_t *,             \
                              const uint32_t *, const uint32_t *, const BoxSet &,    \
                              int, const void *, void *, hipStream_t);
CUCO_PM_INST(double)
CUCO_PM_INST(float)
#undef CUCO_PM_INST

}  // namespace cuco
