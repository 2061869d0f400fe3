// FP32 short-range pair force -- the chosen path.
//
// Same pair law as the reference (gravlist_kernel.cu:20-88 / gravlist.cpp:
// 279-327: minimum image via cell offsets, cut at rcut, spline softening
// inside h, erfc table), re-organised for CDNA4:
//
//  * Particles are sorted by a Morton cell key, so consecutive particles are
//    spatially compact at every density.  A tile is <= 64 consecutive
//    particles: a run of 64 inside one Morton block of 2^s cells per
//    dimension where blocks are dense, several neighbouring sparse blocks of
//    one 2^(s+1) super-block where they are not; one wavefront owns one
//    tile, one lane one target particle.
//  * Positions are FP32 relative to their cell corner (the subtraction is
//    done in FP64 by the reorder kernel).  A tile works in a frame anchored
//    at the corner of its first particle's cell: every coordinate is the
//    cell-relative value plus an integer number of cell sizes, so pair
//    separations are small numbers with small absolute error.
//  * For each tile the wave enumerates the cells within rcut of the tile's
//    bounding box (64 cells per step), reads their ranges from the CSR cell
//    table, and streams the j particles through a filter (distance to the
//    bounding box <= rcut) into an LDS buffer.  Each 64-entry buffer is
//    broadcast to all lanes and evaluated with the erfc table in LDS.
//  * Each lane sums a 64-j chunk in plain FP32 and adds the chunk sum into
//    its total with Kahan compensation.
//  * A persistent grid pulls tiles from an atomic work counter, so dense
//    halo tiles and void tiles balance across the chip.
#include <algorithm>
#include <cmath>

#include "internal.h"

namespace cuco {

namespace {

constexpr int kWaves = 4;
constexpr int kThreads = 64 * kWaves;
constexpr int kRanges = 128;   // cell ranges per wave before the j stream
constexpr int kJBuf = 128;     // staged j particles per wave
constexpr int kTab = 4096;     // LDS erfc table; zero from the cut bin on
constexpr int kTabSym = 2048;  // its copy in the each-pair-once kernel

__device__ inline float wave_min(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ inline float wave_max(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// Wave-uniform copies in scalar registers.  Every loop bound and branch
// condition of the tile loop goes through these, so the compiler emits
// uniform (scalar) control flow: a lane-divergent view of an in-fact
// uniform loop exit lets it split the loop into per-lane nests around the
// cross-lane operations, which then read lanes that are not active.
__device__ inline uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline float uni(float v)
{
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

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

// Scaled coordinate of a particle relative to its tile's anchor cell: the
// FP32 cell-relative coordinate p times s, plus dq cells.  Targets and
// staged candidates both come through here, as one explicit multiply-add
// (this file is compiled with contraction on, and the two sides written as
// separate expressions were once contracted differently): a target and its
// own copy among the candidates then have bit-identical coordinates, so
// the self pair's separation is exactly zero and its term vanishes.
__device__ inline float rel_coord(float p, int dq, float s, float cellf)
{
    return fmaf(p, s, (float)dq * cellf);
}

__device__ inline uint32_t pack_q(int qx, int qy, int qz)
{
    return ((uint32_t)(qx + 512) << 20) | ((uint32_t)(qy + 512) << 10) |
           (uint32_t)(qz + 512);
}

__device__ inline void unpack_q(uint32_t p, int &qx, int &qy, int &qz)
{
    qx = (int)(p >> 20) - 512;
    qy = (int)((p >> 10) & 1023u) - 512;
    qz = (int)(p & 1023u) - 512;
}

__device__ inline int unwrap_delta(int d, int M)
{
    if (d > M / 2) d -= M;
    if (d < -M / 2) d += M;
    return d;
}

// Newtonian-or-spline factor of one pair (gravlist_kernel.cu:47-56 with
// 1/u^3 = h^3 rinv^3); fac enters as m*rinv^3.
template <bool UNIFORM>
__device__ inline float spline_fac(float r, float fac, float mj, const P32 &pp)
{
    const float u = r * pp.h_inv;
    const float m = UNIFORM ? 1.0f : mj;
    const float lo = m * pp.h3_inv * (10.66667f + u * u * (32.0f * u - 38.4f));
    const float hi = m * pp.h3_inv *
                         (21.33333f - 48.0f * u + 38.4f * u * u - 10.66667f * u * u * u) -
                     0.06667f * fac;
    return r < pp.h ? (u < 0.5f ? lo : hi) : fac;
}

// Byte offset of a pair's table entry, bin floor(r / 4) (lengths are in
// quarter bins).  Far pairs (r >= 1): 2^23 - 0.5 + r, rounded to an
// integer by the multiply-add, is 2^23 + floor(r) (an exact integer r can
// round down by one: the lower bin of an edge), and its mantissa with the
// two low bits cleared is 4 floor(r / 4).  Near pairs, where r may be
// below 1, convert r itself.
__device__ inline uint32_t far_bin_bytes(float r2, float rinv)
{
    return __builtin_bit_cast(uint32_t, fmaf(r2, rinv, 8388607.5f)) & 0x7ffffcu;
}

__device__ inline uint32_t near_bin_bytes(float r) { return (uint32_t)(int)r & ~3u; }

__device__ inline float tab_at(const float *__restrict__ tab, uint32_t bytes)
{
    return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(tab) + bytes);
}

// Pair sum of one lane against cnt staged j particles.
//
// GENERIC=false is the reference configuration: rcut sits on a table bin
// edge (RCUT/ASMTH = 4.5 puts it on bin 750) so the table's zeros do the cut.
// GENERIC keeps the explicit r^2 <= rcut^2 mask and clamps the bin.  CLAMP:
// the tile is wide enough for a candidate to reach past the table (the
// caller's per-tile test), so the bin is clamped.  UNIFORM folds the common
// particle mass into the LDS table.
//
// Candidates go four at a time: the four chains (LDS reads, rsq, table
// lookup) are branch-free and interleave; the softened law is applied in
// one wave-level branch per group, taken only when some lane of the wave
// has a pair inside h.  NEAR=false: no staged j is within h of the tile's
// bounding box (chunk_near), so no pair can be and the test is left out.
template <bool GENERIC, bool UNIFORM, bool NEAR, bool CLAMP>
__device__ inline void interact(const float4 *__restrict__ sj, int cnt,
                                const P32 &pp, const float *__restrict__ tab,
                                float xi, float yi, float zi, float &sx,
                                float &sy, float &sz)
{
    int k = 0;
    for (; k + 4 <= cnt; k += 4) {
        float dx[4], dy[4], dz[4], r[4], fac[4], t[4], r2[4], mj[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float4 p = sj[k + q];
            dx[q] = p.x - xi;
            dy[q] = p.y - yi;
            dz[q] = p.z - zi;
            mj[q] = p.w;
            // Self and coincident pairs: dx = dy = dz = 0, eps keeps rsq
            // finite and the term is exactly zero (gravlist_kernel.cu:64-69).
            r2[q] = fmaf(dx[q], dx[q], fmaf(dy[q], dy[q], fmaf(dz[q], dz[q], pp.eps)));
            const float rinv = __builtin_amdgcn_rsqf(r2[q]);
            const float rinv3 = rinv * rinv * rinv;
            fac[q] = UNIFORM ? rinv3 : mj[q] * rinv3;
            uint32_t b;
            if (NEAR) {
                r[q] = r2[q] * rinv;
                b = near_bin_bytes(r[q]);
            } else {
                b = far_bin_bytes(r2[q], rinv);
            }
            if (GENERIC || CLAMP) b = min(b, 4u * (kTab - 1));
            t[q] = tab_at(tab, b);
        }
        if (NEAR) {
            const bool close = (r[0] < pp.h) | (r[1] < pp.h) | (r[2] < pp.h) | (r[3] < pp.h);
            if (close) {
#pragma unroll
                for (int q = 0; q < 4; ++q) fac[q] = spline_fac<UNIFORM>(r[q], fac[q], mj[q], pp);
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            float tv = t[q];
            if (GENERIC) tv = r2[q] <= pp.rcut2 ? tv : 0.0f;
            const float f = fac[q] * tv;
            sx = fmaf(dx[q], f, sx);
            sy = fmaf(dy[q], f, sy);
            sz = fmaf(dz[q], f, sz);
        }
    }
    for (; k < cnt; ++k) {
        const float4 p = sj[k];
        const float dx = p.x - xi, dy = p.y - yi, dz = p.z - zi;
        const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, pp.eps)));
        const float rinv = __builtin_amdgcn_rsqf(r2);
        const float r = r2 * rinv;
        const float rinv3 = rinv * rinv * rinv;
        float fac = UNIFORM ? rinv3 : p.w * rinv3;
        if (NEAR) fac = spline_fac<UNIFORM>(r, fac, p.w, pp);
        uint32_t b = NEAR ? near_bin_bytes(r) : far_bin_bytes(r2, rinv);
        if (GENERIC || CLAMP) b = min(b, 4u * (kTab - 1));
        float tv = tab_at(tab, b);
        if (GENERIC) tv = r2 <= pp.rcut2 ? tv : 0.0f;
        const float f = fac * tv;
        sx = fmaf(dx, f, sx);
        sy = fmaf(dy, f, sy);
        sz = fmaf(dz, f, sz);
    }
}

__device__ inline void kahan_add(float &sum, float &comp, float v)
{
    const float y = v - comp;
    const float t = sum + y;
    comp = (t - sum) - y;
    sum = t;
}

// ---------------------------------------------------------------------------
// Each pair once (Newton's third law).  A tile takes only the j with a larger
// sorted index than its first particle, and inside the tile only j > i; the
// force on j is minus the pair term, summed over the wave's lanes as it travels with j (DPP)
// and added to j's total in 64-bit fixed point (an integer sum, so the result
// does not depend on the order in which tiles run).

// The j side rides along with j: at step s lane L evaluates the pair with
// staged j number (L + s) mod 64 and subtracts the pair term from its
// accumulator, which then moves one lane down (lane L takes lane L+1's), so
// the accumulator of j is always in the lane that meets j next; after 64
// steps lane L holds the j-side sum of staged j number L.  The move is a
// DPP wave rotation (one full-rate vector instruction per component); a
// backward permute through the LDS crossbar instead takes those three
// instructions off the vector ALU but chains every step on the LDS round
// trip, and measured 70.5 against 47.9 ms at halo256.
__device__ inline float rotate_down(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x134, 0xF, 0xF, false));
}

// Pair sums of one lane against the 64 staged j.  GEN: rcut is not on a
// table bin edge, so pairs past it are masked by r^2 and the bin clamped;
// otherwise the table's zero tail makes the cut.  UNI: the common particle
// mass is folded into the table; otherwise the target side takes j's mass
// (the staged w) and the j side the target's (mi).  With MASKED a pair counts only when j's
// sorted index exceeds i's (iid = 0xffffffff for lanes without a target;
// unused slots carry index 0 and a position inside the tile).
// rot / rotid hold the 64 staged j twice (slots 64..127 repeat 0..63), so
// lane L reads slot L + s with no wrap.
template <bool GEN, bool UNI, bool MASKED, bool NEAR, bool CLAMP>
__device__ inline void interact_sym(const float4 *__restrict__ rot,
                                    const uint32_t *__restrict__ rotid,
                                    const P32 &pp, const float *__restrict__ tab,
                                    float xi, float yi, float zi, float mi, uint32_t iid,
                                    float &sx, float &sy, float &sz, float &jx,
                                    float &jy, float &jz)
{
    const int lane = threadIdx.x & 63;
    const float4 *rl = rot + lane;
    const uint32_t *il = rotid + lane;
    auto group = [&](int s) {
        float dx[4], dy[4], dz[4], r[4], fac[4], t[4], mj[4];
        bool keep[4], inside[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float4 p = rl[s + q];
            mj[q] = p.w;
            if (MASKED) keep[q] = il[s + q] > iid;
            dx[q] = p.x - xi;
            dy[q] = p.y - yi;
            dz[q] = p.z - zi;
            const float r2 = fmaf(dx[q], dx[q], fmaf(dy[q], dy[q], fmaf(dz[q], dz[q], pp.eps)));
            const float rinv = __builtin_amdgcn_rsqf(r2);
            fac[q] = rinv * rinv * rinv;
            if (GEN) inside[q] = r2 <= pp.rcut2;
            // The table's zero tail from the cut bin on.  CLAMP: bins past
            // the LDS copy read its last (zero) entry; without it the tile's
            // extent keeps every bin inside (the per-tile clamp test).
            uint32_t b;
            if (NEAR) {
                r[q] = r2 * rinv;
                b = near_bin_bytes(r[q]);
            } else {
                b = far_bin_bytes(r2, rinv);
            }
            if (CLAMP || GEN) b = min(b, 4u * (kTabSym - 1));
            t[q] = tab_at(tab, b);
        }
        if (NEAR) {
            const bool close = (r[0] < pp.h) | (r[1] < pp.h) | (r[2] < pp.h) | (r[3] < pp.h);
            if (close) {
#pragma unroll
                for (int q = 0; q < 4; ++q) fac[q] = spline_fac<true>(r[q], fac[q], 1.0f, pp);
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            float f = fac[q] * t[q];
            if (MASKED) f = keep[q] ? f : 0.0f;
            if (GEN) f = inside[q] ? f : 0.0f;
            const float fi = UNI ? f : f * mj[q];
            const float fj = UNI ? f : f * mi;
            sx = fmaf(dx[q], fi, sx);
            sy = fmaf(dy[q], fi, sy);
            sz = fmaf(dz[q], fi, sz);
            jx = rotate_down(fmaf(-dx[q], fj, jx));
            jy = rotate_down(fmaf(-dy[q], fj, jy));
            jz = rotate_down(fmaf(-dz[q], fj, jz));
        }
    };
    // The common case (no masks, far, no clamp) is unrolled whole: the LDS
    // reads then take their offsets from the instruction and the loop keeps
    // no counter (1-3% at halo256).
    if (!MASKED && !NEAR && !CLAMP) {
#pragma unroll
        for (int s = 0; s < 64; s += 4) group(s);
    } else {
        for (int s = 0; s < 64; s += 4) group(s);
    }
}

// Two targets per lane (a tile of up to 128): at each step the staged j
// meets both, its position and index are read once, and the j-side sum
// takes both terms before the one rotation -- 20.5 vector instructions per
// pair slot instead of 22.  Same masks and laws as interact_sym.
template <bool GEN, bool UNI, bool MASKED, bool NEAR, bool CLAMP>
__device__ inline void interact_sym2(const float4 *__restrict__ rot,
                                     const uint32_t *__restrict__ rotid, const P32 &pp,
                                     const float *__restrict__ tab, float xa, float ya, float za,
                                     float ma, uint32_t iida, float xb, float yb, float zb,
                                     float mb, uint32_t iidb, float &sxa, float &sya,
                                     float &sza, float &sxb, float &syb, float &szb,
                                     float &jx, float &jy, float &jz)
{
    const int lane = threadIdx.x & 63;
    const float4 *rl = rot + lane;
    const uint32_t *il = rotid + lane;
    auto group = [&](int s) {
        float dxa[2], dya[2], dza[2], ra[2], fa[2], ta[2];
        float dxb[2], dyb[2], dzb[2], rb[2], fb[2], tb[2], mj[2];
        bool keepa[2], keepb[2], ina[2], inb[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const float4 p = rl[s + q];
            mj[q] = p.w;
            if (MASKED) {
                const uint32_t id = il[s + q];
                keepa[q] = id > iida;
                keepb[q] = id > iidb;
            }
            dxa[q] = p.x - xa;
            dya[q] = p.y - ya;
            dza[q] = p.z - za;
            dxb[q] = p.x - xb;
            dyb[q] = p.y - yb;
            dzb[q] = p.z - zb;
            const float r2a = fmaf(dxa[q], dxa[q], fmaf(dya[q], dya[q], fmaf(dza[q], dza[q], pp.eps)));
            const float r2b = fmaf(dxb[q], dxb[q], fmaf(dyb[q], dyb[q], fmaf(dzb[q], dzb[q], pp.eps)));
            const float rinva = __builtin_amdgcn_rsqf(r2a);
            const float rinvb = __builtin_amdgcn_rsqf(r2b);
            fa[q] = rinva * rinva * rinva;
            fb[q] = rinvb * rinvb * rinvb;
            if (GEN) {
                ina[q] = r2a <= pp.rcut2;
                inb[q] = r2b <= pp.rcut2;
            }
            uint32_t ba, bb;
            if (NEAR) {
                ra[q] = r2a * rinva;
                rb[q] = r2b * rinvb;
                ba = near_bin_bytes(ra[q]);
                bb = near_bin_bytes(rb[q]);
            } else {
                ba = far_bin_bytes(r2a, rinva);
                bb = far_bin_bytes(r2b, rinvb);
            }
            if (CLAMP || GEN) {
                ba = min(ba, 4u * (kTabSym - 1));
                bb = min(bb, 4u * (kTabSym - 1));
            }
            ta[q] = tab_at(tab, ba);
            tb[q] = tab_at(tab, bb);
        }
        if (NEAR) {
            const bool close = (ra[0] < pp.h) | (ra[1] < pp.h) | (rb[0] < pp.h) | (rb[1] < pp.h);
            if (close) {
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    fa[q] = spline_fac<true>(ra[q], fa[q], 1.0f, pp);
                    fb[q] = spline_fac<true>(rb[q], fb[q], 1.0f, pp);
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            float f1 = fa[q] * ta[q], f2 = fb[q] * tb[q];
            if (MASKED) {
                f1 = keepa[q] ? f1 : 0.0f;
                f2 = keepb[q] ? f2 : 0.0f;
            }
            if (GEN) {
                f1 = ina[q] ? f1 : 0.0f;
                f2 = inb[q] ? f2 : 0.0f;
            }
            const float fia = UNI ? f1 : f1 * mj[q], fja = UNI ? f1 : f1 * ma;
            const float fib = UNI ? f2 : f2 * mj[q], fjb = UNI ? f2 : f2 * mb;
            sxa = fmaf(dxa[q], fia, sxa);
            sya = fmaf(dya[q], fia, sya);
            sza = fmaf(dza[q], fia, sza);
            sxb = fmaf(dxb[q], fib, sxb);
            syb = fmaf(dyb[q], fib, syb);
            szb = fmaf(dzb[q], fib, szb);
            jx = rotate_down(fmaf(-dxb[q], fjb, fmaf(-dxa[q], fja, jx)));
            jy = rotate_down(fmaf(-dyb[q], fjb, fmaf(-dya[q], fja, jy)));
            jz = rotate_down(fmaf(-dzb[q], fjb, fmaf(-dza[q], fja, jz)));
        }
    };
    if (!MASKED && !NEAR && !CLAMP) {
#pragma unroll
        for (int s = 0; s < 64; s += 2) group(s);
    } else {
        for (int s = 0; s < 64; s += 2) group(s);
    }
}

// True when some staged j (lanes with in set) lies within h of the tile's
// bounding box; otherwise every pair of the chunk has r >= h and takes the
// plain law.  Called with h = max(softening, 1 quarter bin), so a far
// chunk's pairs also have r >= 1, which far_bin_bytes needs.  Wave-uniform.
__device__ inline bool chunk_near(float4 q, bool in, float bminx, float bminy, float bminz,
                                  float bmaxx, float bmaxy, float bmaxz, float h)
{
    const float gx = fmaxf(0.f, fmaxf(bminx - q.x, q.x - bmaxx));
    const float gy = fmaxf(0.f, fmaxf(bminy - q.y, q.y - bmaxy));
    const float gz = fmaxf(0.f, fmaxf(bminz - q.z, q.z - bmaxz));
    const bool near = in && gx * gx + gy * gy + gz * gz < h * h;
    return __builtin_amdgcn_readfirstlane((uint32_t)(__ballot(near) != 0ull)) != 0u;
}

// Tiles of all boxes form one work list; tv.map holds each box's first
// tile, p32s[b] the scaled constants of box b.
// SYM: each pair once (section above); the tile's own sums go to iacc
// (sorted order, unscaled) and the j side to jacc (three 64-bit fixed-point
// planes of n_sorted entries); a second kernel combines them.
// TPL = 2 (each pair once only): a tile holds up to 128 targets, lane L
// targets L and 64 + L.
template <bool GENERIC, bool UNIFORM, bool SYM, int TPL>
__global__ __launch_bounds__(kThreads, SYM ? (TPL == 2 ? 4 : 5) : 1) void k_pp32(
    const float4 *__restrict__ spos, const uint32_t *__restrict__ skeys,
    const uint32_t *__restrict__ perm, const uint32_t *__restrict__ off,
    BoxView tv, const P32 *__restrict__ p32s, float mass,
    const float *__restrict__ table32,
    const uint32_t *__restrict__ tstart, const uint32_t *__restrict__ tcount,
    uint32_t *__restrict__ counters, float4 *__restrict__ acc,
    unsigned long long *__restrict__ cand,
    const int32_t *__restrict__ active_end, int32_t ti_current,
    float4 *__restrict__ iacc, unsigned long long *__restrict__ jacc,
    uint32_t n_sorted, const uint32_t *__restrict__ tlist, float coarse_cells)
{
    // SYM trades half of the range list for the doubled rotation buffer.
    constexpr int kR = SYM ? kRanges / 2 : kRanges;
    constexpr int kT = SYM ? kTabSym : kTab;
    __shared__ float s_tab[kT];
    __shared__ uint32_t s_rs[kWaves][kR];
    __shared__ uint32_t s_rp[kWaves][kR + 1];
    __shared__ uint32_t s_rq[kWaves][kR];
    __shared__ float4 s_j[kWaves][kJBuf];
    __shared__ uint32_t s_jid[kWaves][SYM ? kJBuf : 1];
    __shared__ float4 s_rot[kWaves][SYM ? 128 : 1];
    __shared__ uint32_t s_rotid[kWaves][SYM ? 128 : 1];

    for (int k = threadIdx.x; k < kT; k += kThreads)
        s_tab[k] = k < kTabPadded ? (UNIFORM ? mass * table32[k] : table32[k]) : 0.0f;
    __syncthreads();

    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t *rs = s_rs[w];
    uint32_t *rp = s_rp[w];
    uint32_t *rq = s_rq[w];
    float4 *sj = s_j[w];
    uint32_t *sjid = s_jid[w];
    float4 *rot = s_rot[w];
    uint32_t *rotid = s_rotid[w];
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint32_t ntiles = counters[0];
    const float inf = __builtin_huge_valf();

    for (;;) {
        // Every lane takes part in the atomic (lane 0 adds 1, the rest 0), so
        // the loop header carries no lane-divergent branch; lane 0's return
        // value is the tile index whatever order the adds are applied in.
        uint32_t t = atomicAdd(&counters[1], lane == 0 ? 1u : 0u);
        t = __builtin_amdgcn_readlane(t, 0);
        if (t >= ntiles) break;

        const int bi = find_box(tv.map, tv.nboxes, t);
        const P32 pp = p32s[bi];
        const GridInfo &g = tv.box[bi].g;
        const uint32_t key_base = tv.box[bi].key_base;
        const int M = pp.M;
        const uint32_t i0 = tstart[t];
        const uint32_t cnt = tcount[t];
        const bool valid = (uint32_t)lane < cnt;
        // tlist: the tile is a run of the active-target list (ia: its first
        // target, the anchor); else a run of the sorted particles.
        const uint32_t ia = tlist ? tlist[i0] : i0;
        const uint32_t i = tlist ? tlist[i0 + (valid ? (uint32_t)lane : 0u)]
                                 : i0 + (valid ? (uint32_t)lane : 0u);
        // Active set: a tile without an active target is skipped; in the
        // others every lane computes and only active lanes store.
        bool act = valid;
        if (active_end) act = valid && active_end[perm[i]] == ti_current;
        if (uni((uint32_t)(__ballot(act) != 0ull)) == 0u) continue;

        int c0x, c0y, c0z, cx, cy, cz;
        key_cell(g, skeys[ia] - key_base, c0x, c0y, c0z);
        key_cell(g, skeys[i] - key_base, cx, cy, cz);
        const float4 pi = spos[i];
        const float xi = rel_coord(pi.x, unwrap_delta(cx - c0x, M), pp.s, pp.cellf);
        const float yi = rel_coord(pi.y, unwrap_delta(cy - c0y, M), pp.s, pp.cellf);
        const float zi = rel_coord(pi.z, unwrap_delta(cz - c0z, M), pp.s, pp.cellf);
        // The lane's second target (TPL = 2): 64 + L of the tile.
        const bool valid2 = TPL == 2 && (uint32_t)lane + 64u < cnt;
        const uint32_t i2 = valid2 ? i0 + 64u + (uint32_t)lane : i;
        float xi2 = xi, yi2 = yi, zi2 = zi, mi2 = pi.w;
        if (TPL == 2) {
            int cx2, cy2, cz2;
            key_cell(g, skeys[i2] - key_base, cx2, cy2, cz2);
            const float4 p2 = spos[i2];
            xi2 = rel_coord(p2.x, unwrap_delta(cx2 - c0x, M), pp.s, pp.cellf);
            yi2 = rel_coord(p2.y, unwrap_delta(cy2 - c0y, M), pp.s, pp.cellf);
            zi2 = rel_coord(p2.z, unwrap_delta(cz2 - c0z, M), pp.s, pp.cellf);
            mi2 = p2.w;
        }

        const float bminx = uni(wave_min(fminf(valid ? xi : inf, valid2 ? xi2 : inf)));
        const float bminy = uni(wave_min(fminf(valid ? yi : inf, valid2 ? yi2 : inf)));
        const float bminz = uni(wave_min(fminf(valid ? zi : inf, valid2 ? zi2 : inf)));
        const float bmaxx = uni(wave_max(fmaxf(valid ? xi : -inf, valid2 ? xi2 : -inf)));
        const float bmaxy = uni(wave_max(fmaxf(valid ? yi : -inf, valid2 ? yi2 : -inf)));
        const float bmaxz = uni(wave_max(fmaxf(valid ? zi : -inf, valid2 ? zi2 : -inf)));

        const int lqx = (int)floorf((bminx - pp.rcut) * pp.inv_cellf - 0.01f);
        const int lqy = (int)floorf((bminy - pp.rcut) * pp.inv_cellf - 0.01f);
        const int lqz = (int)floorf((bminz - pp.rcut) * pp.inv_cellf - 0.01f);
        const int nqx = (int)floorf((bmaxx + pp.rcut) * pp.inv_cellf + 0.01f) - lqx + 1;
        const int nqy = (int)floorf((bmaxy + pp.rcut) * pp.inv_cellf + 0.01f) - lqy + 1;
        const int nqz = (int)floorf((bmaxz + pp.rcut) * pp.inv_cellf + 0.01f) - lqz + 1;
        // A tile spans at most one Morton super-block plus the cut on each
        // side; anything larger means non-finite input and the tile is
        // skipped.
        const bool sane = nqx > 0 && nqy > 0 && nqz > 0 && nqx <= 2 * M &&
                          nqy <= 2 * M && nqz <= 2 * M;
        // Enumeration unit: one cell, or for a wide tile (a merge of sparse
        // blocks) an aligned 2x2x2-cell Morton block, whose cells form one
        // contiguous key range: eight times fewer units to test, and the
        // staging takes each candidate's cell from the low bits of its key.
        // Blocks need an even mesh (they wrap onto blocks) and a reach of at
        // most M cells per axis; coarse_cells is the tile extent (in cells)
        // from which they are used (both variants).
        const float ext = fmaxf(bmaxx - bminx, fmaxf(bmaxy - bminy, bmaxz - bminz));
        const bool coarse =
            uni((uint32_t)(sane && (M & 1) == 0 && nqx <= M && nqy <= M && nqz <= M &&
                                  ext >= coarse_cells * pp.cellf)) != 0u;
        const int lv = coarse ? 1 : 0;
        // Units (absolute, unwrapped) covering cells c0 + lq .. c0 + lq + nq - 1.
        const int ux0 = (c0x + lqx) >> lv, uy0 = (c0y + lqy) >> lv, uz0 = (c0z + lqz) >> lv;
        const int nux = ((c0x + lqx + nqx - 1) >> lv) - ux0 + 1;
        const int nuy = ((c0y + lqy + nqy - 1) >> lv) - uy0 + 1;
        const int nuz = ((c0z + lqz + nqz - 1) >> lv) - uz0 + 1;
        const uint32_t total_q = uni(sane ? (uint32_t)nux * nuy * nuz : 0u);

        float ax = 0.f, ay = 0.f, az = 0.f, kx = 0.f, ky = 0.f, kz = 0.f;
        float ax2 = 0.f, ay2 = 0.f, az2 = 0.f, kx2 = 0.f, ky2 = 0.f, kz2 = 0.f;
        // Every candidate lies within sqrt(filter2) of the tile's box, so a
        // pair's distance is at most that plus the box diagonal; below the
        // last bin of the LDS table (with margin for rounding) the bin needs
        // no clamp.
        const float ex = bmaxx - bminx, ey = bmaxy - bminy, ez = bmaxz - bminz;
        const bool clamp =
            !(sqrtf(ex * ex + ey * ey + ez * ez) + sqrtf(pp.filter2) + 1.0f < 4.0f * (kT - 2));
        // SYM: the j side of one staged chunk, added to jacc (fixed point).
        const uint32_t iid = valid ? i : 0xffffffffu;
        const uint32_t iid2 = valid2 ? i2 : 0xffffffffu;
        auto eval_sym = [&](int nj, float &sx_, float &sy_, float &sz_, float &sx2_, float &sy2_,
                            float &sz2_) {
            float jx = 0.f, jy = 0.f, jz = 0.f;
            // Copy the chunk twice into the rotation buffer; unused slots of a
            // partial chunk get index 0 (always masked) and a position one
            // quarter bin below the tile's box (a bounded far-pair bin).
            float4 mine = make_float4(bminx - 1.0f, bminy, bminz, 0.f);
            uint32_t mid = 0u;
            if (lane < nj) {
                mine = sj[lane];
                mid = sjid[lane];
            }
            rot[lane] = mine;
            rot[lane + 64] = mine;
            rotid[lane] = mid;
            rotid[lane + 64] = mid;
            wave_sync();
            // Pairs inside the tile, lanes without a target, unused slots:
            // mask by index.
            const bool inside = __ballot(lane < nj && mid < i0 + cnt) != 0ull;
            const bool near = chunk_near(mine, lane < nj, bminx, bminy, bminz, bmaxx, bmaxy,
                                         bmaxz, pp.hnear);
            const bool masked = nj < 64 || cnt < 64u * TPL || inside;
#define CUCO_SYM(MK, NR, CL)                                                                    \
    if (TPL == 2)                                                                               \
        interact_sym2<GENERIC, UNIFORM, MK, NR, CL>(rot, rotid, pp, s_tab, xi, yi, zi, pi.w, iid, \
                                                    xi2, yi2, zi2, mi2, iid2, sx_, sy_, sz_,    \
                                                    sx2_, sy2_, sz2_, jx, jy, jz);              \
    else                                                                                        \
        interact_sym<GENERIC, UNIFORM, MK, NR, CL>(rot, rotid, pp, s_tab, xi, yi, zi, pi.w, iid, \
                                                   sx_, sy_, sz_, jx, jy, jz)
            if (clamp) {
                if (masked) {
                    if (near) {
                        CUCO_SYM(true, true, true);
                    } else {
                        CUCO_SYM(true, false, true);
                    }
                } else if (near) {
                    CUCO_SYM(false, true, true);
                } else {
                    CUCO_SYM(false, false, true);
                }
            } else {
                if (masked) {
                    if (near) {
                        CUCO_SYM(true, true, false);
                    } else {
                        CUCO_SYM(true, false, false);
                    }
                } else if (near) {
                    CUCO_SYM(false, true, false);
                } else {
                    CUCO_SYM(false, false, false);
                }
            }
#undef CUCO_SYM
            if ((uint32_t)lane < (uint32_t)nj) {
                const uint32_t j = mid;
                atomicAdd(&jacc[j], (unsigned long long)__double2ll_rn((double)jx * pp.jscale));
                atomicAdd(&jacc[(size_t)n_sorted + j], (unsigned long long)__double2ll_rn((double)jy * pp.jscale));
                atomicAdd(&jacc[2 * (size_t)n_sorted + j], (unsigned long long)__double2ll_rn((double)jz * pp.jscale));
            }
        };
        auto eval_default = [&](int nj, float &sx_, float &sy_, float &sz_) {
            const float4 mine = lane < nj ? sj[lane] : make_float4(0.f, 0.f, 0.f, 0.f);
            const bool near =
                chunk_near(mine, lane < nj, bminx, bminy, bminz, bmaxx, bmaxy, bmaxz, pp.hnear);
#define CUCO_DEF(NR, CL) \
    interact<GENERIC, UNIFORM, NR, CL>(sj, nj, pp, s_tab, xi, yi, zi, sx_, sy_, sz_)
            if (clamp) {
                if (near) CUCO_DEF(true, true);
                else CUCO_DEF(false, true);
            } else if (near) {
                CUCO_DEF(true, false);
            } else {
                CUCO_DEF(false, false);
            }
#undef CUCO_DEF
        };
        uint32_t jc = 0;
        uint32_t cursor = 0;
        unsigned long long ncand = 0;

        // Unit number cursor + lane as digits (dx, dy, dz) of the mixed radix
        // (nux, nuy, nuz), advanced by the digits of 64 each round (no
        // division per unit); wx0.. = the wrapped unit of digit 0.
        uint32_t dz = 0, dy = 0, dx = 0, sz = 0, sy = 0, sx = 0;
        uint32_t wx0 = 0, wy0 = 0, wz0 = 0;
        const int Mb = M >> lv;
        if (total_q) {
            const uint32_t nz = (uint32_t)nuz, ny = (uint32_t)nuy;
            dz = (uint32_t)lane % nz;
            dy = ((uint32_t)lane / nz) % ny;
            dx = ((uint32_t)lane / nz) / ny;
            sz = 64u % nz;
            sy = (64u / nz) % ny;
            sx = (64u / nz) / ny;
            wx0 = (uint32_t)((ux0 % Mb + Mb) % Mb);
            wy0 = (uint32_t)((uy0 % Mb + Mb) % Mb);
            wz0 = (uint32_t)((uz0 % Mb + Mb) % Mb);
        }
        // A digit is below nu <= 2 Mb + 1 (cells: nq <= 2M; blocks: nq <= M),
        // so two conditional subtractions wrap.
        const uint32_t Mu = (uint32_t)Mb;
        auto wrap2 = [Mu](uint32_t u) {
            u = u >= Mu ? u - Mu : u;
            return u >= Mu ? u - Mu : u;
        };
        // First cell of digit 0 relative to the anchor cell, and the unit's
        // width: cells, and scaled length.
        const int qbx = (ux0 << lv) - c0x, qby = (uy0 << lv) - c0y, qbz = (uz0 << lv) - c0z;
        const int ub = 1 << lv;
        const float uw = (float)ub * pp.cellf;
        const uint32_t span = 1u << (3 * lv);  // keys per unit
        // SYM: a unit whose keys all precede the tile's first particle's
        // holds only particles before it (keys ascend in sorted order).
        const uint32_t key_first = skeys[ia];

        while (cursor < total_q) {
            // Enumerate cells into the range list.
            uint32_t nr = 0;
            // Wide tiles keep filling the range list while a whole round
            // fits (a round that does not fit is redone for the next batch):
            // their rounds find few non-empty units each.
            while (cursor < total_q && (coarse || nr + 64 <= (uint32_t)kR)) {
                bool ok = cursor + lane < total_q;
                uint32_t s = 0, c = 0, packed = 0;
                if (ok) {
                    const int qz = (int)dz * ub + qbz;
                    const int qy = (int)dy * ub + qby;
                    const int qx = (int)dx * ub + qbx;
                    const float lx = qx * pp.cellf, ly = qy * pp.cellf, lz = qz * pp.cellf;
                    const float gx = fmaxf(0.f, fmaxf(lx - bmaxx, bminx - (lx + uw)));
                    const float gy = fmaxf(0.f, fmaxf(ly - bmaxy, bminy - (ly + uw)));
                    const float gz = fmaxf(0.f, fmaxf(lz - bmaxz, bminz - (lz + uw)));
                    ok = gx * gx + gy * gy + gz * gz <= pp.filter2;
                    // FP32 keys are Morton keys; a block's key is its cells'
                    // keys without their 3 low bits.
                    const uint32_t key = key_base + (((spread3(wrap2(wx0 + dx)) << 2) |
                                                      (spread3(wrap2(wy0 + dy)) << 1) |
                                                      spread3(wrap2(wz0 + dz))) << (3 * lv));
                    if (SYM) ok = ok && key + span > key_first;
                    if (ok) {
                        uint32_t e1;
                        if (coarse) {
                            s = off[key];
                            e1 = off[key + 8];
                        } else {
                            load_range(off, key, s, e1);
                        }
                        c = e1 - s;
                        if (SYM) {
                            // only j after the tile's first particle
                            const uint32_t e = s + c, lo = i0 + 1;
                            s = max(s, lo);
                            c = e > s ? e - s : 0u;
                        }
                        ok = c > 0;
                        packed = pack_q(qx, qy, qz);
                    }
                }
                const uint64_t bal = __ballot(ok);
                const uint32_t nnew = uni((uint32_t)__popcll(bal));
                if (nr + nnew > (uint32_t)kR) break;
                if (ok) {
                    const uint32_t pos = nr + (uint32_t)__popcll(bal & lt);
                    rs[pos] = s;
                    rp[pos] = c;
                    rq[pos] = packed;
                }
                nr += nnew;
                cursor += 64;
                dz += sz;
                const uint32_t carry_z = dz >= (uint32_t)nuz ? 1u : 0u;
                dz -= carry_z * (uint32_t)nuz;
                dy += sy + carry_z;
                const uint32_t carry_y = dy >= (uint32_t)nuy ? 1u : 0u;
                dy -= carry_y * (uint32_t)nuy;
                dx += sx + carry_y;
            }
            wave_sync();
            // Exclusive prefix of the range lengths.
            uint32_t carry = 0;
            for (uint32_t b0 = 0; b0 < nr; b0 += 64) {
                const uint32_t idx = b0 + lane;
                const uint32_t v = idx < nr ? rp[idx] : 0u;
                const uint32_t inc = wave_incl_scan(v);
                if (idx < nr) rp[idx] = carry + inc - v;
                carry = uni(carry + (uint32_t)__builtin_amdgcn_readlane(inc, 63));
            }
            if (lane == 0) rp[nr] = carry;
            wave_sync();
            const uint32_t total_j = carry;

            // Stream, filter and evaluate.  Candidate slots are taken 64 at a
            // time; rc is the range holding the window's first slot f0.
            // Ranges are non-empty, so the ranges rc+1.. start at strictly
            // increasing slots after f0: lane L looks up where range rc+1+L
            // starts, marks that slot of the window (a permute: each lane
            // sends a 1 to the lane of its start slot; lanes without a
            // start in the window send 0 to slot 0, which no range after rc
            // can start at), and slot s then lies in range rc + (number of
            // marks at or below s).
            uint32_t rc = 0;
            for (uint32_t f0 = 0; f0 < total_j; f0 += 64) {
                const uint32_t f = f0 + lane;
                bool ok = f < total_j;
                const uint32_t rn = rc + 1 + (uint32_t)lane;
                const uint32_t st = rn < nr ? rp[rn] - f0 : 0xffffffffu;
                const bool mark = st < 64u;
                const int got = __builtin_amdgcn_ds_permute(mark ? (int)(st * 4u) : 0, mark ? 1 : 0);
                const uint64_t starts = __ballot(got != 0);
                const uint32_t lo =
                    rc + __builtin_amdgcn_mbcnt_hi((uint32_t)(starts >> 32),
                                                   __builtin_amdgcn_mbcnt_lo((uint32_t)starts, 0u)) +
                    (got != 0 ? 1u : 0u);
                rc = uni(rc + (uint32_t)__popcll(__ballot(st <= 64u)));
                float4 qj = make_float4(0.f, 0.f, 0.f, 0.f);
                uint32_t jsorted = 0;
                if (ok) {
                    const uint32_t j = rs[lo] + (f - rp[lo]);
                    jsorted = j;
                    int qx, qy, qz;
                    unpack_q(rq[lo], qx, qy, qz);
                    if (coarse) {
                        // the cell inside the 2x2x2 block: the key's low bits
                        const uint32_t kj = skeys[j];
                        qx += (int)((kj >> 2) & 1u);
                        qy += (int)((kj >> 1) & 1u);
                        qz += (int)(kj & 1u);
                    }
                    const float4 pj = spos[j];
                    qj = make_float4(rel_coord(pj.x, qx, pp.s, pp.cellf),
                                     rel_coord(pj.y, qy, pp.s, pp.cellf),
                                     rel_coord(pj.z, qz, pp.s, pp.cellf), pj.w);
                    const float gx = fmaxf(0.f, fmaxf(bminx - qj.x, qj.x - bmaxx));
                    const float gy = fmaxf(0.f, fmaxf(bminy - qj.y, qj.y - bmaxy));
                    const float gz = fmaxf(0.f, fmaxf(bminz - qj.z, qj.z - bmaxz));
                    ok = gx * gx + gy * gy + gz * gz <= pp.filter2;
                }
                const uint64_t bal = __ballot(ok);
                if (ok) {
                    const uint32_t slot = jc + (uint32_t)__popcll(bal & lt);
                    sj[slot] = qj;
                    if (SYM) sjid[slot] = jsorted;
                }
                jc = uni(jc + (uint32_t)__popcll(bal));
                wave_sync();
                if (jc >= 64) {
                    float sx = 0.f, sy = 0.f, sz = 0.f, sx2 = 0.f, sy2 = 0.f, sz2 = 0.f;
                    if (SYM) {
                        eval_sym(64, sx, sy, sz, sx2, sy2, sz2);
                    } else {
                        eval_default(64, sx, sy, sz);
                    }
                    kahan_add(ax, kx, sx);
                    kahan_add(ay, ky, sy);
                    kahan_add(az, kz, sz);
                    if (TPL == 2) {
                        kahan_add(ax2, kx2, sx2);
                        kahan_add(ay2, ky2, sy2);
                        kahan_add(az2, kz2, sz2);
                    }
                    if (cand) ncand += 64ull * cnt;
                    const uint32_t rem = jc - 64;
                    float4 tmp = make_float4(0.f, 0.f, 0.f, 0.f);
                    uint32_t tid = 0;
                    if ((uint32_t)lane < rem) {
                        tmp = sj[64 + lane];
                        if (SYM) tid = sjid[64 + lane];
                    }
                    wave_sync();
                    if ((uint32_t)lane < rem) {
                        sj[lane] = tmp;
                        if (SYM) sjid[lane] = tid;
                    }
                    jc = rem;
                    wave_sync();
                }
            }
        }
        if (jc > 0) {
            float sx = 0.f, sy = 0.f, sz = 0.f, sx2 = 0.f, sy2 = 0.f, sz2 = 0.f;
            if (SYM) {
                eval_sym((int)jc, sx, sy, sz, sx2, sy2, sz2);
            } else {
                eval_default((int)jc, sx, sy, sz);
            }
            kahan_add(ax, kx, sx);
            kahan_add(ay, ky, sy);
            kahan_add(az, kz, sz);
            if (TPL == 2) {
                kahan_add(ax2, kx2, sx2);
                kahan_add(ay2, ky2, sy2);
                kahan_add(az2, kz2, sz2);
            }
            if (cand) ncand += (unsigned long long)jc * cnt;
            jc = 0;
        }
        wave_sync();
        if (SYM) {
            if (valid) iacc[i] = make_float4(ax, ay, az, 0.f);
            if (TPL == 2 && valid2) iacc[i2] = make_float4(ax2, ay2, az2, 0.f);
        } else if (act) {
            acc[perm[i]] = make_float4(ax * pp.out_scale, ay * pp.out_scale,
                                       az * pp.out_scale, 0.f);
        }
        if (cand) atomicAdd(cand, lane == 0 ? ncand : 0ull);
    }
}

// Tiles of one super-block: the 8 Morton blocks (2^shift cells a side) of
// one block a level up.  Consecutive blocks whose particles add up to at
// most run (64, or 128 for two targets per lane) share a tile (in sparse
// regions a block alone would leave most of the wave idle); a block with
// more than run particles is cut into runs of that length.  A tile is a
// contiguous range of sorted particles inside one super-block.  Returns the
// number of tiles; writes them from t0 on when tstart is given.
__device__ inline uint32_t super_tiles(const uint32_t *__restrict__ off, uint32_t sb, int shift3,
                                       uint32_t run, uint32_t *__restrict__ tstart,
                                       uint32_t *__restrict__ tcount, uint32_t t0)
{
    const size_t base = (size_t)sb << (shift3 + 3);
    uint32_t t = t0, gs = off[base], gn = 0;
    // An empty super-block (voids, and the padding of a Morton key space
    // around a mesh that is not a power of two) has no tiles: two reads
    // instead of nine.
    if (off[base + ((size_t)8 << shift3)] == gs) return 0;
    auto emit = [&](uint32_t start, uint32_t n) {
        if (tstart) {
            tstart[t] = start;
            tcount[t] = n;
        }
        ++t;
    };
    for (int k = 0; k < 8; ++k) {
        const uint32_t bs = off[base + ((size_t)k << shift3)];
        const uint32_t be = off[base + ((size_t)(k + 1) << shift3)];
        const uint32_t c = be - bs;
        if (c > run) {
            if (gn) emit(gs, gn);
            gn = 0;
            for (uint32_t p = bs; p < be; p += run) emit(p, min(run, be - p));
        } else if (gn + c > run) {
            emit(gs, gn);
            gs = bs;
            gn = c;
        } else {
            if (gn == 0) gs = bs;
            gn += c;
        }
    }
    if (gn) emit(gs, gn);
    return t - t0;
}

__global__ void k_block_tiles(const uint32_t *__restrict__ off, uint32_t nsuper, int shift3,
                              uint32_t run, uint32_t *__restrict__ super_count)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nsuper) super_count[b] = super_tiles(off, b, shift3, run, nullptr, nullptr, 0);
}

__global__ void k_fill_tiles(const uint32_t *__restrict__ off,
                             const uint32_t *__restrict__ tile_off, uint32_t nsuper,
                             int shift3, uint32_t run, uint32_t *__restrict__ tstart,
                             uint32_t *__restrict__ tcount)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nsuper) (void)super_tiles(off, b, shift3, run, tstart, tcount, tile_off[b]);
}

// First tile of each box: boxes start on super-block boundaries (a box's
// key range is 8^bits with bits > shift), so it is the scanned tile offset
// of the box's first super-block.
__global__ void k_box_tiles(const uint32_t *__restrict__ tile_off,
                            const BoxDev *__restrict__ boxes, int nboxes,
                            int shift3, const uint32_t *__restrict__ counters,
                            uint32_t *__restrict__ tile_map)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nboxes) tile_map[b] = tile_off[boxes[b].key_base >> (shift3 + 3)];
    if (b == nboxes) tile_map[b] = counters[0];
}

// Active targets of an integer-timeline step, packed into their own tiles.
// flag: 1 where the particle at sorted index s is active; scanned in place
// (exclusive, scan[n] = total), scan[s+1] - scan[s] is that flag again.
__global__ void k_act_flags(const uint32_t *__restrict__ perm,
                            const int32_t *__restrict__ active_end, int32_t ti_current,
                            uint32_t n, uint32_t *__restrict__ flag)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) flag[s] = active_end[perm[s]] == ti_current ? 1u : 0u;
}

__global__ void k_act_list(const uint32_t *__restrict__ scan, uint32_t n,
                           uint32_t *__restrict__ list)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n && scan[s + 1] > scan[s]) list[scan[s]] = s;
}

// Active targets of super-block sb: list positions [a0, a0 + count).
__device__ inline uint32_t super_active(const uint32_t *__restrict__ off,
                                        const uint32_t *__restrict__ scan, uint32_t sb,
                                        int shift3, uint32_t &a0)
{
    const size_t base = (size_t)sb << (shift3 + 3);
    a0 = scan[off[base]];
    return scan[off[base + ((size_t)1 << (shift3 + 3))]] - a0;
}

__global__ void k_block_tiles_act(const uint32_t *__restrict__ off,
                                  const uint32_t *__restrict__ scan, uint32_t nsuper,
                                  int shift3, uint32_t *__restrict__ super_count)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nsuper) return;
    uint32_t a0;
    super_count[b] = (super_active(off, scan, b, shift3, a0) + 63u) / 64u;
}

__global__ void k_fill_tiles_act(const uint32_t *__restrict__ off,
                                 const uint32_t *__restrict__ scan,
                                 const uint32_t *__restrict__ tile_off, uint32_t nsuper,
                                 int shift3, uint32_t *__restrict__ tstart,
                                 uint32_t *__restrict__ tcount)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nsuper) return;
    uint32_t a0;
    const uint32_t c = super_active(off, scan, b, shift3, a0);
    uint32_t t = tile_off[b];
    for (uint32_t p = 0; p < c; p += 64u, ++t) {
        tstart[t] = a0 + p;
        tcount[t] = min(64u, c - p);
    }
}

// Each-pair-once result: own sums plus the fixed-point j sums, scaled to
// accelerations and scattered to original order.
__global__ void k_pp32_combine(const float4 *__restrict__ iacc,
                               const unsigned long long *__restrict__ jacc,
                               uint32_t n_sorted, const uint32_t *__restrict__ perm,
                               BoxView v, const P32 *__restrict__ p32s,
                               float4 *__restrict__ acc)
{
    int bi;
    uint32_t lb;
    if (!box_block(v, bi, lb)) return;
    const BoxDev &bx = v.box[bi];
    const uint32_t il = lb * 256 + threadIdx.x;
    if (il >= bx.n) return;
    const uint32_t i = bx.p0 + il;
    const double inv = p32s[bi].jinv;
    const float sc = p32s[bi].out_scale;
    const float4 a = iacc[i];
    const float x = a.x + (float)((double)(long long)jacc[i] * inv);
    const float y = a.y + (float)((double)(long long)jacc[(size_t)n_sorted + i] * inv);
    const float z = a.z + (float)((double)(long long)jacc[2 * (size_t)n_sorted + i] * inv);
    acc[perm[i]] = make_float4(x * sc, y * sc, z * sc, 0.f);
}

}  // namespace

namespace {

int build_active_tiles(const uint32_t *off, const uint32_t *perm, const int32_t *active_end,
                       int32_t ti_current, uint32_t n, const BoxSet &bs,
                       const ActiveTiles &at, hipStream_t s)
{
    const int shift3 = 3 * at.t.block_shift;
    const uint32_t nsuper = at.nkeys >> (shift3 + 3);
    if (nsuper == 0 || nsuper > at.t.nblocks || n > at.capacity) {
        set_error("active tiles: %u super-blocks / %u particles exceed capacity", nsuper, n);
        return CUCO_ERR_INVALID_ARG;
    }
    k_act_flags<<<grid_for(n, 256), 256, 0, s>>>(perm, active_end, ti_current, n, at.scan);
    CUCO_LAUNCH_CHECK();
    int rc = exclusive_scan_u32(at.scan, at.scan, n, at.nscratch, at.scan + n, s);
    if (rc) return rc;
    k_act_list<<<grid_for(n, 256), 256, 0, s>>>(at.scan, n, at.list);
    CUCO_LAUNCH_CHECK();
    k_block_tiles_act<<<grid_for(nsuper, 256), 256, 0, s>>>(off, at.scan, nsuper, shift3,
                                                            at.t.block_tiles);
    CUCO_LAUNCH_CHECK();
    rc = exclusive_scan_u32(at.t.block_tiles, at.t.block_tiles, nsuper, at.t.scan_scratch,
                            at.t.counters, s);
    if (rc) return rc;
    k_fill_tiles_act<<<grid_for(nsuper, 256), 256, 0, s>>>(off, at.scan, at.t.block_tiles,
                                                           nsuper, shift3, at.t.tile_start,
                                                           at.t.tile_count);
    CUCO_LAUNCH_CHECK();
    k_box_tiles<<<grid_for(bs.nboxes + 1, 64), 64, 0, s>>>(
        at.t.block_tiles, bs.dev[kKeyMorton], bs.nboxes, shift3, at.t.counters, at.tile_map);
    CUCO_LAUNCH_CHECK();
    return CUCO_OK;
}

}  // namespace

int launch_build_tiles(const uint32_t *off, const BoxSet &bs, uint32_t nkeys,
                       TileInfo &t, hipStream_t s)
{
    const int shift3 = 3 * t.block_shift;
    const uint32_t nsuper = nkeys >> (shift3 + 3);
    // t.nblocks is the capacity (blocks of the smallest tunable shift).
    if (nsuper == 0 || nsuper > t.nblocks) {
        set_error("tile build: %u Morton super-blocks exceed capacity %u", nsuper, t.nblocks);
        return CUCO_ERR_INVALID_ARG;
    }
    if (t.run != 64u && t.run != 128u) {
        set_error("tile build: run length %u (64 or 128)", t.run);
        return CUCO_ERR_INVALID_ARG;
    }
    k_block_tiles<<<grid_for(nsuper, 256), 256, 0, s>>>(off, nsuper, shift3, t.run,
                                                        t.block_tiles);
    CUCO_LAUNCH_CHECK();
    int rc = exclusive_scan_u32(t.block_tiles, t.block_tiles, nsuper, t.scan_scratch,
                                t.counters, s);
    if (rc) return rc;
    k_fill_tiles<<<grid_for(nsuper, 256), 256, 0, s>>>(off, t.block_tiles, nsuper, shift3,
                                                       t.run, t.tile_start, t.tile_count);
    CUCO_LAUNCH_CHECK();
    k_box_tiles<<<grid_for(bs.nboxes + 1, 64), 64, 0, s>>>(
        t.block_tiles, bs.dev[kKeyMorton], bs.nboxes, shift3, t.counters, bs.tile_map);
    CUCO_LAUNCH_CHECK();
    return CUCO_OK;
}

P32 make_p32(const PairParams &pp, const GridInfo &g, double uniform_mass,
             double mass_max, uint32_t n_box)
{
    const double sc = 4.0 * pp.asmthfac;  // quarter bins
    P32 p;
    p.s = (float)sc;
    p.rcut = (float)(pp.rcut * sc);
    p.rcut2 = (float)(pp.rcut2 * sc * sc);
    p.filter2 = (float)(pp.rcut2 * sc * sc * (1.0 + 1e-4));
    p.h = (float)(pp.h * sc);
    p.h_inv = (float)(1.0 / (pp.h * sc));
    p.h3_inv = p.h_inv * p.h_inv * p.h_inv;
    p.hnear = std::max(p.h, 1.0f);
    p.cellf = (float)(g.cell * sc);
    p.inv_cellf = (float)(1.0 / (g.cell * sc));
    p.eps = 1e-12f;
    p.out_scale = (float)(pp.G * sc * sc);
    p.mass = (float)uniform_mass;
    p.M = g.M;
    // Fixed-point scale of the each-pair-once j sums: a pair term is at most
    // m_max * 10.67 / h^2 (scaled units; the spline's peak, the table is <= 1),
    // so n such terms stay below 2^62 / scale.
    const double pair_max = std::max(mass_max, 1e-30) * 10.67 / ((double)p.h * p.h);
    const double bound = std::max(1.0, (double)n_box) * pair_max;
    int e2 = 0;
    std::frexp(std::ldexp(1.0, 62) / bound, &e2);
    p.jscale = (float)std::ldexp(1.0, std::min(e2 - 1, 126));
    p.jinv = 1.0f / p.jscale;
    return p;
}

bool pp32_table_cut(const PairParams &pp)
{
    // The table alone can make the cut when rcut sits on a bin edge; bins a
    // wide tile's candidates could reach past the LDS table are clamped per
    // tile in the kernel.
    const double edge = pp.asmthfac * pp.rcut;
    return fabs(edge - nearbyint(edge)) < 1e-9 * edge;
}

int launch_pp32(const float4 *spos, const uint32_t *skeys,
                const uint32_t *perm, const uint32_t *off, const BoxSet &bs,
                const P32 *p32s, bool table_cut, double uniform_mass,
                const float *table32, const TileInfo &t, float4 *acc,
                unsigned long long *cand_counter, int num_workgroups,
                const int32_t *active_end, int32_t ti_current,
                const PairSym &sym, const ActiveTiles *at, uint32_t n_sorted,
                hipStream_t s)
{
    if (bs.total[kMapPart] == 0) return CUCO_OK;
    const bool generic = !table_cut;
    const bool uniform = uniform_mass > 0;
    const float mass = (float)uniform_mass;
    const BoxView tv{bs.dev[kKeyMorton], bs.tile_map, bs.nboxes};
    int rc = launch_fill_u32(t.counters + 1, 0u, 1, s);
    if (rc) return rc;
    const int nwg = num_workgroups > 0 ? num_workgroups : 2048;
    // Each pair once: all targets (not an active set).
    const bool symmetric = sym.on && sym.iacc && sym.jacc && !active_end;
    if (symmetric) {
        rc = launch_fill_u32(reinterpret_cast<uint32_t *>(sym.jacc), 0u,
                             (size_t)6 * sym.n_sorted, s);
        if (rc) return rc;
#define CUCO_SYM32(GENERIC, UNIFORM, TPL)                                               \
    k_pp32<GENERIC, UNIFORM, true, TPL><<<nwg, kThreads, 0, s>>>(                       \
        spos, skeys, perm, off, tv, p32s, mass, table32, t.tile_start, t.tile_count,    \
        t.counters, acc, cand_counter, nullptr, 0, sym.iacc, sym.jacc, sym.n_sorted,   \
        nullptr, t.coarse_cells)
        if (t.run == 128u) {
            if (generic) {
                if (uniform) CUCO_SYM32(true, true, 2); else CUCO_SYM32(true, false, 2);
            } else {
                if (uniform) CUCO_SYM32(false, true, 2); else CUCO_SYM32(false, false, 2);
            }
        } else {
            if (generic) {
                if (uniform) CUCO_SYM32(true, true, 1); else CUCO_SYM32(true, false, 1);
            } else {
                if (uniform) CUCO_SYM32(false, true, 1); else CUCO_SYM32(false, false, 1);
            }
        }
#undef CUCO_SYM32
        CUCO_LAUNCH_CHECK();
        k_pp32_combine<<<xcd_grid(bs.total[kMapPart]), 256, 0, s>>>(
            sym.iacc, sym.jacc, sym.n_sorted, perm, bs.view(kKeyMorton, kMapPart), p32s, acc);
        CUCO_LAUNCH_CHECK();
        return CUCO_OK;
    }
    // Active set with its own tiles: every listed target is active.
    const bool packed = active_end && at;
    const TileInfo &tt = packed ? at->t : t;
    // The kernel that visits every ordered pair holds one target per lane.
    if (tt.run != 64u) {
        set_error("pair force: tiles of %u targets need the each-pair-once kernel", tt.run);
        return CUCO_ERR_INVALID_ARG;
    }
    const BoxView tva{bs.dev[kKeyMorton], packed ? at->tile_map : bs.tile_map, bs.nboxes};
    const int32_t *ae = packed ? nullptr : active_end;
    const uint32_t *tl = packed ? at->list : nullptr;
    if (packed) {
        rc = build_active_tiles(off, perm, active_end, ti_current, n_sorted, bs, *at, s);
        if (rc) return rc;
        rc = launch_fill_u32(at->t.counters + 1, 0u, 1, s);
        if (rc) return rc;
    }
#define CUCO_PP32(GENERIC, UNIFORM)                                                     \
    k_pp32<GENERIC, UNIFORM, false, 1><<<nwg, kThreads, 0, s>>>(                        \
        spos, skeys, perm, off, tva, p32s, mass, table32, tt.tile_start, tt.tile_count, \
        tt.counters, acc, cand_counter, ae, ti_current, nullptr, nullptr, 0u, tl, t.coarse_cells)
    if (generic) {
        if (uniform) CUCO_PP32(true, true); else CUCO_PP32(true, false);
    } else {
        if (uniform) CUCO_PP32(false, true); else CUCO_PP32(false, false);
    }
#undef CUCO_PP32
    CUCO_LAUNCH_CHECK();
    return CUCO_OK;
}

}  // namespace cuco
