// FP32 3D FFTs held in registers, for power-of-two meshes (M = 32..1024) and
// meshes of 3 or 5 times a power of two (M = 96, 160, 192):
// the transforms of pm_periodic.c:140 (forward r2c), :146-197 (Green's
// function on the x lines) and :201 (c2r), for the FP32 path.
//
// A line of length N = N1 x N2 is spread over N1 threads, N2 values each
// (the four-step split, j = g + N1*m, k = k2 + N2*k1):
//
//   X[k2 + N2*k1] = sum_g W_N1^(g*k1) * W_N^(g*k2) * sum_m x[g + N1*m] W_N2^(m*k2)
//
// Thread g loads x[g + N1*m] for every m straight from HBM, runs the
// length-N2 DFT in registers, applies W_N^(g*k2), and swaps through LDS once
// so that it holds, for its k2 = g + N1*u, all N1 partial sums; the
// length-N1 DFTs then give X[k2 + N2*k1], stored straight to HBM.  One LDS
// exchange per pass (the Stockham kernels need one per radix stage), no
// twiddle loads inside the small DFTs (their constants are compile-time).
//
// Blocks of 256 threads (128 where 256 would need over 64 KB of LDS for the
// exchange) hold threads/N1 lines.  In the strided passes (y and x
// lines) consecutive lanes take consecutive lines, so each load and store of
// a wave covers 128-byte runs of the mesh (16 lines x 8 bytes at N1 = 16);
// in the z-row passes consecutive lanes take consecutive elements of a row.
//
// The real<->half-complex split of the z rows (length-M real rows through a
// length-M/2 complex FFT) and the Green's function factors are the ones of
// the generic kernels, evaluated in FP32.
#include "internal.h"

namespace cuco {

namespace {

constexpr int kRegThreads = 256;

// ------------------------------------------------------------ constants
// cos and sin of 2*pi*k/n at compile time: exact octant reduction in
// integers, then Taylor series on [0, pi/4] (12 terms: below 1e-20).
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double taylor_cos(double x)
{
    double t = 1.0, s = 1.0;
    for (int i = 1; i <= 12; ++i) {
        t *= -x * x / ((2.0 * i - 1.0) * (2.0 * i));
        s += t;
    }
    return s;
}

constexpr double taylor_sin(double x)
{
    double t = x, s = x;
    for (int i = 1; i <= 12; ++i) {
        t *= -x * x / ((2.0 * i) * (2.0 * i + 1.0));
        s += t;
    }
    return s;
}

// cos(pi*a/(2n)) and sin(pi*a/(2n)) for 0 <= a <= n (angle in [0, pi/2]).
constexpr double quad_cos(int a, int n)
{
    return 2 * a <= n ? taylor_cos(kPi * a / (2.0 * n)) : taylor_sin(kPi * (n - a) / (2.0 * n));
}
constexpr double quad_sin(int a, int n)
{
    return 2 * a <= n ? taylor_sin(kPi * a / (2.0 * n)) : taylor_cos(kPi * (n - a) / (2.0 * n));
}

// cos/sin(2*pi*k/n): quadrant q = floor(4k/n), remainder a = 4k - q*n in
// [0, n), angle within the quadrant pi*a/(2n).
constexpr double ct_cos(int k, int n)
{
    const int q = (4 * k) / n, a = 4 * k - q * n;
    const double c = quad_cos(a, n), s = quad_sin(a, n);
    return q == 0 ? c : q == 1 ? -s : q == 2 ? -c : s;
}
constexpr double ct_sin(int k, int n)
{
    const int q = (4 * k) / n, a = 4 * k - q * n;
    const double c = quad_cos(a, n), s = quad_sin(a, n);
    return q == 0 ? s : q == 1 ? c : q == 2 ? -s : -c;
}

static_assert(ct_cos(0, 8) == 1.0 && ct_sin(2, 8) == 1.0 && ct_cos(4, 8) == -1.0,
              "compile-time twiddles");
static_assert(ct_cos(1, 8) - 0.70710678118654752440 < 4e-16 &&
                  0.70710678118654752440 - ct_cos(1, 8) < 4e-16 &&
                  ct_sin(5, 24) - 0.96592582628906828675 < 4e-16 &&
                  0.96592582628906828675 - ct_sin(5, 24) < 4e-16,
              "compile-time twiddles");

using C = Cx<float>;

__device__ __forceinline__ C c_add(C a, C b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ C c_sub(C a, C b) { return {a.x - b.x, a.y - b.y}; }

// x * W_R^K, W = exp(-2 pi i/R) forward, exp(+2 pi i/R) inverse.
template <int K, int R, bool INV>
__device__ __forceinline__ C twiddle(C x)
{
    constexpr int k = K % R;
    if constexpr (k == 0) {
        return x;
    } else if constexpr (4 * k == R) {  // -i forward, +i inverse
        return INV ? C{-x.y, x.x} : C{x.y, -x.x};
    } else if constexpr (2 * k == R) {
        return {-x.x, -x.y};
    } else if constexpr (4 * k == 3 * R) {  // +i forward, -i inverse
        return INV ? C{x.y, -x.x} : C{-x.y, x.x};
    } else {
        constexpr float c = (float)ct_cos(k, R);
        constexpr float s = INV ? (float)ct_sin(k, R) : -(float)ct_sin(k, R);
        return {x.x * c - x.y * s, x.x * s + x.y * c};
    }
}

// ------------------------------------------------------------ register DFTs
// Length-R DFT of v[0..R) in place, natural order in and out (radix-2
// decimation in time down to lengths 2, 4, 3 or 5; every index and twiddle
// is a compile-time constant, so v stays in registers).
template <int R, bool INV>
struct RegDft;

// Odd lengths (3, 5) by the definition X[k] = sum_j v[j] W_R^(j k).
template <int R, bool INV>
struct DirectDft {
    template <int K, int J>
    __device__ __forceinline__ static C term(const C *u)
    {
        if constexpr (J + 1 == R)
            return twiddle<(J * K) % R, R, INV>(u[J]);
        else
            return c_add(twiddle<(J * K) % R, R, INV>(u[J]), term<K, J + 1>(u));
    }
    template <int K>
    __device__ __forceinline__ static void out(C *v, const C *u)
    {
        if constexpr (K < R) {
            v[K] = term<K, 0>(u);
            out<K + 1>(v, u);
        }
    }
    __device__ __forceinline__ static void run(C *v)
    {
        C u[R];
#pragma unroll
        for (int i = 0; i < R; ++i) u[i] = v[i];
        out<0>(v, u);
    }
};

template <int K, int R, bool INV>
struct Combine {
    __device__ __forceinline__ static void run(C *v, const C *e, const C *o)
    {
        if constexpr (K < R / 2) {
            const C t = twiddle<K, R, INV>(o[K]);
            v[K] = c_add(e[K], t);
            v[K + R / 2] = c_sub(e[K], t);
            Combine<K + 1, R, INV>::run(v, e, o);
        }
    }
};

template <int R, bool INV>
struct RegDft {
    __device__ __forceinline__ static void run(C *v)
    {
        C e[R / 2], o[R / 2];
#pragma unroll
        for (int i = 0; i < R / 2; ++i) {
            e[i] = v[2 * i];
            o[i] = v[2 * i + 1];
        }
        RegDft<R / 2, INV>::run(e);
        RegDft<R / 2, INV>::run(o);
        Combine<0, R, INV>::run(v, e, o);
    }
};

template <bool INV>
struct RegDft<2, INV> {
    __device__ __forceinline__ static void run(C *v)
    {
        const C a = v[0], b = v[1];
        v[0] = c_add(a, b);
        v[1] = c_sub(a, b);
    }
};

template <bool INV>
struct RegDft<3, INV> : DirectDft<3, INV> {};

template <bool INV>
struct RegDft<5, INV> : DirectDft<5, INV> {};

template <bool INV>
struct RegDft<4, INV> {
    __device__ __forceinline__ static void run(C *v)
    {
        const C a0 = c_add(v[0], v[2]), a1 = c_sub(v[0], v[2]);
        const C a2 = c_add(v[1], v[3]), a3 = twiddle<1, 4, INV>(c_sub(v[1], v[3]));
        v[0] = c_add(a0, a2);
        v[2] = c_sub(a0, a2);
        v[1] = c_add(a1, a3);
        v[3] = c_sub(a1, a3);
    }
};

// ------------------------------------------------------------ shapes
constexpr bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

// Threads per line N1, a power of two (the rest, N2 = N/N1, a multiple of
// N1, in registers).  Power-of-two lines: S = 1 takes half as many threads
// per line (more values each, more lines per block).  Lines of 3 or 5 times
// a power of two (48, 80, 96, 160, 192): 8 threads where N/8 is a multiple
// of 8, else 4; S does not apply.
constexpr int reg_n1(int N, int S)
{
    return is_pow2(N) ? (N <= 32 ? 4 : N <= 128 ? 8 : N <= 512 ? 16 : 32) >> S
                      : ((N / 8) % 8 == 0 ? 8 : 4);
}

// Threads per block: 256, or 128 for a line length that is not a power of
// two whose 256/N1 lines would need more than 64 KB of exchange slots.
constexpr int reg_threads(int N, int S)
{
    return !is_pow2(N) && sizeof(Cx<float>) * (size_t)(kRegThreads / reg_n1(N, S)) * N > 65536
               ? kRegThreads / 2
               : kRegThreads;
}

template <int N, int S = 0>
struct Shape {
    static constexpr int N1 = reg_n1(N, S);
    static constexpr int N2 = N / N1;
    static constexpr int T = reg_threads(N, S);  // threads per block
    static constexpr int B = T / N1;             // lines per block
    // LDS exchange swizzle: slot k2*N1 + (g ^ ((k2 + Q*l) mod N1)) of line l
    // spreads every wave's accesses evenly over the banks.
    static constexpr int Q = N1 >= 32 ? 4 : 1;
    static constexpr size_t lds = sizeof(C) * (size_t)B * N;
    static_assert(N2 % N1 == 0, "N2 must be a multiple of N1");
    static_assert(N1 * N2 == N && (is_pow2(N) || lds <= 65536), "register FFT shape");
};

template <int N, int S>
__device__ __forceinline__ int xslot(int l, int g, int k2)
{
    constexpr int N1 = Shape<N, S>::N1;
    return l * N + k2 * N1 + (g ^ ((k2 + Shape<N, S>::Q * l) & (N1 - 1)));
}

__device__ __forceinline__ C conj_if(C w, bool inv) { return inv ? C{w.x, -w.y} : w; }

// First half of the four-step on one line.  In: x[m] = element g + N1*m.
// Runs the length-N2 DFTs, applies W_N^(g*k2) and writes the partial sums
// to the exchange slots of line l.  tw: W_N^k, k < N, at stride ts (the
// box's table holds W_M^k; rows use N = M/2, ts = 2).
template <int N, bool INV, int S>
__device__ __forceinline__ void step_a(C *x, C *lds, int l, int g,
                                       const C *__restrict__ tw, int ts)
{
    constexpr int N1 = Shape<N, S>::N1, N2 = Shape<N, S>::N2;
    RegDft<N2, INV>::run(x);
#pragma unroll
    for (int k2 = 1; k2 < N2; ++k2) {
        const C w = conj_if(tw[g * k2 * ts], INV);  // g * k2 < N
        x[k2] = {x[k2].x * w.x - x[k2].y * w.y, x[k2].x * w.y + x[k2].y * w.x};
    }
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) lds[xslot<N, S>(l, g, k2)] = x[k2];
}

// The full four-step.  Out: y[u*N1 + k1] = X[k2 + N2*k1] with
// k2 = g + N1*u.  The block synchronises inside.
template <int N, bool INV, int S>
__device__ __forceinline__ void four_step(C *x, C *y, C *lds, int l, int g,
                                          const C *__restrict__ tw, int ts)
{
    constexpr int N1 = Shape<N, S>::N1, N2 = Shape<N, S>::N2;
    step_a<N, INV, S>(x, lds, l, g, tw, ts);
    __syncthreads();
#pragma unroll
    for (int u = 0; u < N2 / N1; ++u) {
        const int k2 = g + N1 * u;
        C z[N1];
#pragma unroll
        for (int gp = 0; gp < N1; ++gp) z[gp] = lds[xslot<N, S>(l, gp, k2)];
        RegDft<N1, INV>::run(z);
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1) y[u * N1 + k1] = z[k1];
    }
}

template <int N>
__device__ __forceinline__ C *lds_area()
{
    extern __shared__ __align__(16) unsigned char smem[];
    return reinterpret_cast<C *>(smem);
}

// ------------------------------------------------------------ z rows
// Real rows of length M = 2N -> N+1 complex, forward.  Rows mapping: lane
// group of N1 threads per row.
template <int N, int S>
__global__ __launch_bounds__((Shape<N, S>::T)) void k_rrows_r2c(const float *__restrict__ in_all,
                                                           C *__restrict__ out_all, BoxView v,
                                                           const C *__restrict__ tw_all)
{
    constexpr int N1 = Shape<N, S>::N1, N2 = Shape<N, S>::N2, B = Shape<N, S>::B;
    constexpr int M = 2 * N, Zc = N + 1, nrows = M * M;
    int bi;
    uint32_t lb;
    if (!box_block(v, bi, lb)) return;
    const BoxDev &bx = v.box[bi];
    if (bx.g.M != 2 * N) return;  // another mesh size's launch
    const C *in = reinterpret_cast<const C *>(in_all + bx.grid0);
    C *out = out_all + bx.cgrid0;
    const C *tw = tw_all + bx.tab0;
    C *lds = lds_area<N>();
    const int l = threadIdx.x / N1, g = threadIdx.x % N1;
    const int row0 = (int)lb * B, row = row0 + l;
    const bool ok = row < nrows;
    C x[N2], y[N2];
#pragma unroll
    for (int m = 0; m < N2; ++m)
        x[m] = ok ? in[(size_t)row * N + g + N1 * m] : C{0.f, 0.f};
    four_step<N, false, S>(x, y, lds, l, g, tw, 2);
    __syncthreads();
    // Natural order per row, then the split post-process.
    C *rl = lds + l * N;
#pragma unroll
    for (int u = 0; u < N2 / N1; ++u)
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1) rl[g + N1 * u + N2 * k1] = y[u * N1 + k1];
    __syncthreads();
    for (int e = threadIdx.x; e < Zc * B; e += Shape<N, S>::T) {
        const int ll = e / Zc, k = e - ll * Zc;
        if (row0 + ll >= nrows) break;
        const C *r = lds + ll * N;
        const C zk = r[k == N ? 0 : k];
        const C zm = r[k == 0 ? 0 : N - k];
        const C s = {zk.x + zm.x, zk.y - zm.y}, d = {zk.x - zm.x, zk.y + zm.y};
        const C w = tw[k];  // W_M^k (Z_k - conj Z_{N-k})
        const C wd = {w.x * d.x - w.y * d.y, w.x * d.y + w.y * d.x};
        out[(size_t)(row0 + ll) * Zc + k] = {0.5f * s.x + 0.5f * wd.y, 0.5f * s.y - 0.5f * wd.x};
    }
}

// N+1 complex -> real rows of length 2N, inverse (unnormalized).
template <int N, int S>
__global__ __launch_bounds__((Shape<N, S>::T)) void k_rrows_c2r(const C *__restrict__ in_all,
                                                           float *__restrict__ out_all, BoxView v,
                                                           const C *__restrict__ tw_all)
{
    constexpr int N1 = Shape<N, S>::N1, N2 = Shape<N, S>::N2, B = Shape<N, S>::B;
    constexpr int M = 2 * N, Zc = N + 1, nrows = M * M;
    int bi;
    uint32_t lb;
    if (!box_block(v, bi, lb)) return;
    const BoxDev &bx = v.box[bi];
    if (bx.g.M != 2 * N) return;  // another mesh size's launch
    const C *in = in_all + bx.cgrid0;
    C *out = reinterpret_cast<C *>(out_all + bx.grid0);
    const C *tw = tw_all + bx.tab0;
    C *lds = lds_area<N>();
    const int l = threadIdx.x / N1, g = threadIdx.x % N1;
    const int row = (int)lb * B + l;
    const bool ok = row < nrows;
    C x[N2], y[N2];
#pragma unroll
    for (int m = 0; m < N2; ++m) {
        const int k = g + N1 * m;
        C z = {0.f, 0.f};
        if (ok) {
            C xk = in[(size_t)row * Zc + k];
            C xm = in[(size_t)row * Zc + (N - k)];
            if (k == 0) {  // DC and Nyquist are real
                xk.y = 0.f;
                xm.y = 0.f;
            }
            const C c = {xm.x, -xm.y};
            const C s = c_add(xk, c), d = c_sub(xk, c);
            const C w = tw[k];
            const C wd = {w.x * d.x + w.y * d.y, w.x * d.y - w.y * d.x};  // conj(W) d
            z = {s.x - wd.y, s.y + wd.x};
        }
        x[m] = z;
    }
    four_step<N, true, S>(x, y, lds, l, g, tw, 2);
    if (!ok) return;
#pragma unroll
    for (int u = 0; u < N2 / N1; ++u)
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1)
            out[(size_t)row * N + g + N1 * u + N2 * k1] = y[u * N1 + k1];
}

// ------------------------------------------------------------ strided lines
// Complex mesh [x][y][z], Zc = M/2+1.  y lines: line L = x*Zc + z, elements
// Zc apart; x lines: line L = y*Zc + z, elements M*Zc apart.  Lines mapping:
// lane l of a block takes line lb*B + l.
template <bool XAXIS>
__device__ __forceinline__ size_t rline_base(int L, int M, int Zc)
{
    return XAXIS ? (size_t)L : (size_t)(L / Zc) * M * Zc + (L % Zc);
}

template <int N, bool INV, bool XAXIS, int S>
__global__ __launch_bounds__((Shape<N, S>::T)) void k_rlines(C *__restrict__ data_all, BoxView v,
                                                        const C *__restrict__ tw_all)
{
    constexpr int N1 = Shape<N, S>::N1, N2 = Shape<N, S>::N2, B = Shape<N, S>::B;
    constexpr int M = N, Zc = M / 2 + 1, nlines = M * Zc;
    constexpr size_t stride = XAXIS ? (size_t)M * Zc : (size_t)Zc;
    int bi;
    uint32_t lb;
    if (!box_block(v, bi, lb)) return;
    const BoxDev &bx = v.box[bi];
    if (bx.g.M != N) return;  // another mesh size's launch
    C *data = data_all + bx.cgrid0;
    const C *tw = tw_all + bx.tab0;
    C *lds = lds_area<N>();
    const int l = threadIdx.x % B, g = threadIdx.x / B;
    const int L = (int)lb * B + l;
    const bool ok = L < nlines;
    const size_t base = rline_base<XAXIS>(ok ? L : 0, M, Zc);
    C x[N2], y[N2];
#pragma unroll
    for (int m = 0; m < N2; ++m)
        x[m] = ok ? data[base + (size_t)(g + N1 * m) * stride] : C{0.f, 0.f};
    four_step<N, INV, S>(x, y, lds, l, g, tw, 1);
    if (!ok) return;
#pragma unroll
    for (int u = 0; u < N2 / N1; ++u)
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1)
            data[base + (size_t)(g + N1 * u + N2 * k1) * stride] = y[u * N1 + k1];
}

// x lines: forward FFT, Green's function (pm_periodic.c:146-197, DC zeroed
// at :197), inverse FFT, in one residency.  Line L = y*Zc + z.  After the
// forward exchange each thread owns the k2 = g + N1*u columns: it finishes
// their forward DFTs, multiplies by the Green's function, runs the inverse
// length-N1 DFTs and applies W_N^(-g'*k2) in place -- the same exchange
// slots it read, so no other thread's slots are touched -- and after one
// more barrier reads its own g back for the inverse length-N2 DFTs.
template <int N, int S>
__global__ __launch_bounds__((Shape<N, S>::T), (N >= 1024 || (S && N >= 512)) ? 1 : 2) void k_rgreen(C *__restrict__ data_all, BoxView v,
                                                        const C *__restrict__ tw_all,
                                                        const float *__restrict__ gexp_all,
                                                        const float *__restrict__ gsinc_all,
                                                        const float *__restrict__ gprod_all)
{
    constexpr int N1 = Shape<N, S>::N1, N2 = Shape<N, S>::N2, B = Shape<N, S>::B;
    constexpr int M = N, Zc = M / 2 + 1, nlines = M * Zc;
    constexpr size_t stride = (size_t)M * Zc;
    int bi;
    uint32_t lb;
    if (!box_block(v, bi, lb)) return;
    const BoxDev &bx = v.box[bi];
    if (bx.g.M != N) return;  // another mesh size's launch
    C *data = data_all + bx.cgrid0;
    const C *tw = tw_all + bx.tab0;
    const float *gexp = gexp_all + bx.tab0;
    const float *gsinc = gsinc_all + bx.tab0;
    const float *gprod = gprod_all + bx.tab0;
    C *lds = lds_area<N>();
    const int l = threadIdx.x % B, g = threadIdx.x / B;
    const int L = (int)lb * B + l;
    const bool ok = L < nlines;
    const int Lc = ok ? L : 0;
    C x[N2];
#pragma unroll
    for (int m = 0; m < N2; ++m)
        x[m] = ok ? data[(size_t)Lc + (size_t)(g + N1 * m) * stride] : C{0.f, 0.f};
    step_a<N, false, S>(x, lds, l, g, tw, 1);
    const int yy = Lc / Zc, z = Lc - yy * Zc;
    const int ky = yy > M / 2 ? yy - M : yy;
    // -exp(-k^2 a^2) / sinc^4 / k^2, the y and z factors taken once per line.
    const float fyz = -(gexp[yy] * gexp[z]) * (gsinc[yy] * gsinc[z]);
    const int k2yz = ky * ky + z * z;
    __syncthreads();
#pragma unroll
    for (int u = 0; u < N2 / N1; ++u) {
        const int k2 = g + N1 * u;
        C q[N1];
#pragma unroll
        for (int gp = 0; gp < N1; ++gp) q[gp] = lds[xslot<N, S>(l, gp, k2)];
        RegDft<N1, false>::run(q);
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1) {
            const int j = k2 + N2 * k1;
            const int kx = j > M / 2 ? j - M : j;
            const int k2i = kx * kx + k2yz;
            const float gr =
                k2i > 0 ? gprod[j] * fyz * __builtin_amdgcn_rcpf((float)k2i) : 0.f;
            q[k1] = {q[k1].x * gr, q[k1].y * gr};
        }
        RegDft<N1, true>::run(q);
#pragma unroll
        for (int gp = 0; gp < N1; ++gp) {
            const C w = conj_if(tw[gp * k2], true);  // gp * k2 < N
            lds[xslot<N, S>(l, gp, k2)] = {q[gp].x * w.x - q[gp].y * w.y,
                                        q[gp].x * w.y + q[gp].y * w.x};
        }
    }
    __syncthreads();
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) x[k2] = lds[xslot<N, S>(l, g, k2)];
    RegDft<N2, true>::run(x);
    if (!ok) return;
#pragma unroll
    for (int m = 0; m < N2; ++m) data[(size_t)Lc + (size_t)(g + N1 * m) * stride] = x[m];
}

// ------------------------------------------------------------ launchers
// The mesh sizes with register FFTs; bit i of BoxSet::fft32_mask, and the
// i-th row and line maps, stand for kRegMeshes[i].
constexpr int kRegMeshes[] = {32, 64, 128, 256, 512, 1024, 96, 160, 192};
constexpr int kNRegMeshes = (int)(sizeof(kRegMeshes) / sizeof(kRegMeshes[0]));
static_assert(kNRegMeshes == kRegFftSizes, "one block map per register-FFT size");

constexpr int reg_index(int M)
{
    for (int i = 0; i < kNRegMeshes; ++i)
        if (kRegMeshes[i] == M) return i;
    return -1;
}
template <int M>
constexpr int rows_map = kMapRowsR + reg_index(M);
template <int M>
constexpr int lines_map = kMapLinesR + reg_index(M);

template <typename K, typename... Args>
int launch_reg(K kernel, size_t lds, int threads, unsigned blocks, hipStream_t s,
               Args... args)
{
    if (blocks == 0) return CUCO_OK;
    kernel<<<xcd_grid(blocks), threads, lds, s>>>(args...);
    CUCO_LAUNCH_CHECK();
    return CUCO_OK;
}

template <int M, int S>
int forward_m(const float *rho, C *c, const BoxSet &bs, const C *tw, hipStream_t s)
{
    int rc = launch_reg(k_rrows_r2c<M / 2, S>, Shape<M / 2, S>::lds, Shape<M / 2, S>::T, bs.total[rows_map<M>], s,
                        rho, c, bs.view(0, rows_map<M>), tw);
    if (rc) return rc;
    return launch_reg(k_rlines<M, false, false, S>, Shape<M, S>::lds, Shape<M, S>::T, bs.total[lines_map<M>], s,
                      c, bs.view(0, lines_map<M>), tw);
}

template <int M, int S>
int inverse_m(C *c, float *phi, const BoxSet &bs, const C *tw, hipStream_t s)
{
    int rc = launch_reg(k_rlines<M, true, false, S>, Shape<M, S>::lds, Shape<M, S>::T, bs.total[lines_map<M>], s,
                        c, bs.view(0, lines_map<M>), tw);
    if (rc) return rc;
    return launch_reg(k_rrows_c2r<M / 2, S>, Shape<M / 2, S>::lds, Shape<M / 2, S>::T, bs.total[rows_map<M>], s,
                      (const C *)c, phi, bs.view(0, rows_map<M>), tw);
}

template <int M, int S>
int green_m(C *c, const BoxSet &bs, const PmTables &tb, hipStream_t s)
{
    return launch_reg(k_rgreen<M, S>, Shape<M, S>::lds, Shape<M, S>::T, bs.total[lines_map<M>], s, c,
                      bs.view(0, lines_map<M>), static_cast<const C *>(tb.tw),
                      static_cast<const float *>(tb.gexp), static_cast<const float *>(tb.gsinc),
                      static_cast<const float *>(tb.gprod));
}

template <int M, int S>
int x_m(C *c, const BoxSet &bs, bool inverse, const C *tw, hipStream_t s)
{
    if (inverse)
        return launch_reg(k_rlines<M, true, true, S>, Shape<M, S>::lds, Shape<M, S>::T, bs.total[lines_map<M>], s,
                          c, bs.view(0, lines_map<M>), tw);
    return launch_reg(k_rlines<M, false, true, S>, Shape<M, S>::lds, Shape<M, S>::T, bs.total[lines_map<M>], s, c,
                      bs.view(0, lines_map<M>), tw);
}

// One launch per mesh size present in the set.
template <typename F>
int each_mesh(const BoxSet &bs, F &&f)
{
    for (int i = 0; i < kNRegMeshes; ++i)
        if (bs.fft32_mask & (1u << i)) {
            const int rc = f(kRegMeshes[i]);
            if (rc) return rc;
        }
    return CUCO_OK;
}

// The split variant (S = 1) of a mesh size: its own shape for powers of
// two, the S = 0 shape otherwise.
template <int M>
constexpr int kSplitS = is_pow2(M) ? 1 : 0;

#define CUCO_FFT32_CASES(M, CALL)              \
    switch (M) {                               \
    case 32: return CALL(32);                  \
    case 64: return CALL(64);                  \
    case 128: return CALL(128);                \
    case 256: return CALL(256);                \
    case 512: return CALL(512);                \
    case 1024: return CALL(1024);              \
    case 96: return CALL(96);                  \
    case 160: return CALL(160);                \
    case 192: return CALL(192);                \
    default: return (int)CUCO_ERR_UNSUPPORTED; \
    }

}  // namespace

bool fft32_supported(int M)
{
    return fft32_mask_bit(M) != 0;
}

uint32_t fft32_mask_bit(int M)
{
    for (int i = 0; i < kNRegMeshes; ++i)
        if (kRegMeshes[i] == M) return 1u << i;
    return 0;
}

uint32_t fft32_blocks(int M, bool rows, int split)
{
    if (!fft32_supported(M)) return 0;
    const int N = rows ? M / 2 : M;
    const int S = split && is_pow2(M) ? 1 : 0;
    const uint32_t B = (uint32_t)(reg_threads(N, S) / reg_n1(N, S));
    const uint32_t lines = rows ? (uint32_t)M * M : (uint32_t)M * (M / 2 + 1);
    return (lines + B - 1) / B;
}

int fft32_forward(const float *rho, Cx<float> *c, const BoxSet &bs, const PmTables &tb,
                  hipStream_t s)
{
    const C *tw = static_cast<const C *>(tb.tw);
    return each_mesh(bs, [&](int M) -> int {
#define CALL(m) (bs.fft_split ? forward_m<m, kSplitS<m>>(rho, c, bs, tw, s) : forward_m<m, 0>(rho, c, bs, tw, s))
        CUCO_FFT32_CASES(M, CALL)
#undef CALL
    });
}

int fft32_inverse(Cx<float> *c, float *phi, const BoxSet &bs, const PmTables &tb,
                  hipStream_t s)
{
    const C *tw = static_cast<const C *>(tb.tw);
    return each_mesh(bs, [&](int M) -> int {
#define CALL(m) (bs.fft_split ? inverse_m<m, kSplitS<m>>(c, phi, bs, tw, s) : inverse_m<m, 0>(c, phi, bs, tw, s))
        CUCO_FFT32_CASES(M, CALL)
#undef CALL
    });
}

int fft32_green_x(Cx<float> *c, const BoxSet &bs, const PmTables &tb, hipStream_t s)
{
    return each_mesh(bs, [&](int M) -> int {
#define CALL(m) (bs.fft_split ? green_m<m, kSplitS<m>>(c, bs, tb, s) : green_m<m, 0>(c, bs, tb, s))
        CUCO_FFT32_CASES(M, CALL)
#undef CALL
    });
}

int fft32_x(Cx<float> *c, const BoxSet &bs, bool inverse, const PmTables &tb, hipStream_t s)
{
    const C *tw = static_cast<const C *>(tb.tw);
    return each_mesh(bs, [&](int M) -> int {
#define CALL(m) (bs.fft_split ? x_m<m, kSplitS<m>>(c, bs, inverse, tw, s) : x_m<m, 0>(c, bs, inverse, tw, s))
        CUCO_FFT32_CASES(M, CALL)
#undef CALL
    });
}

#undef CUCO_FFT32_CASES

}  // namespace cuco
