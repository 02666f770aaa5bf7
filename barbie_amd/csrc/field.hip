// Seeded non-Gaussian redshift-space input field on the GPU: the step before
// the estimator, the reference's generator (generate_field.c:22-107) with a
// counter-based random stream instead of GSL's mt19937 + ziggurat:
//
//  1. Fill_Modes (ngrf_2d_funcs.c:341-505): every Fourier mode of the r2c
//     half space gets Re, Im ~ N(0, V P(k) / 2); the kz = 0 and kz = NN/2
//     planes are Hermitian (a mode and its mirror are conjugate, the eight
//     self-conjugate modes are real), the k = 0 mode is zero;
//  2. Make_Image (:512-531): divided by V, complex-to-real 3-D transform;
//  3. Non_Gauss (:553-577): delta += fnl (delta^2 - <delta^2>);
//  4. Get_Modes (:534-549): times LL^3, real-to-complex 3-D transform;
//  5. RSD (:734-765): times (1 + beta mu^2), mu = kz / |k|.
//
// Randomness: mode c (its index in the r2c buffer) draws one block of
// Philox4x64-10 (Salmon, Moraes, Dror and Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC 2011) with counter (c, 0, 0, 0) and key (seed, 0):
// the first two 64-bit words give two uniforms u = (w >> 11) 2^-53 and a
// Box-Muller pair r (cos 2 pi u2, sin 2 pi u2), r = sqrt(-2 ln(1 - u1)).
// The logarithm and the sine and cosine are this file's own series,
// evaluated with explicit fused multiply-adds and contraction off, so that a
// plain C restatement of the same steps gives the same bits on the CPU.
//
// The transforms are batched 1-D Stockham FFTs in LDS (fft_lds.hpp) along
// each axis, in FP64: the two complex axes through tiles of 4 columns, the
// real axis a row at a time (Hermitian extension for complex-to-real).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "barbie.h"
#include "device.hpp"
#include "fft_lds.hpp"
#include "plan.hpp"

namespace {

using u64 = unsigned long long;

// ------------------------------------------------------------ Philox4x64 --
__device__ __forceinline__ void philox4x64_10(u64 c0, u64 c1, u64 c2, u64 c3, u64 k0, u64 k1,
                                              u64 out[4])
{
    const u64 M0 = 0xD2E7470EE14C6C93ULL, M1 = 0xCA5A826395121157ULL;
    const u64 W0 = 0x9E3779B97F4A7C15ULL, W1 = 0xBB67AE8584CAA73BULL;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const u64 lo0 = M0 * c0, hi0 = __umul64hi(M0, c0);
        const u64 lo1 = M1 * c2, hi1 = __umul64hi(M1, c2);
        const u64 n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += W0;
        k1 += W1;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// ------------------------------------------------- log, sin, cos (own) --
// ln x for x in (0, 1]: x = m 2^e with m in [sqrt(1/2), sqrt(2)), s =
// (m - 1)/(m + 1), ln m = 2 atanh s = 2 s sum_n s^(2n) / (2n + 1) (s^2 <
// 0.0295, 12 terms), ln x = e ln2 + ln m with ln 2 split in two parts.
__device__ __forceinline__ double gen_log(double x)
{
#pragma clang fp contract(off)
    int e;
    double m = frexp(x, &e);   // m in [0.5, 1)
    if (m < 0.70710678118654752440) {
        m = m * 2.0;
        e = e - 1;
    }
    const double f = m - 1.0;
    const double s = f / (m + 1.0);
    const double z = s * s;
    double p = 1.0 / 23.0;
    p = __builtin_fma(p, z, 1.0 / 21.0);
    p = __builtin_fma(p, z, 1.0 / 19.0);
    p = __builtin_fma(p, z, 1.0 / 17.0);
    p = __builtin_fma(p, z, 1.0 / 15.0);
    p = __builtin_fma(p, z, 1.0 / 13.0);
    p = __builtin_fma(p, z, 1.0 / 11.0);
    p = __builtin_fma(p, z, 1.0 / 9.0);
    p = __builtin_fma(p, z, 1.0 / 7.0);
    p = __builtin_fma(p, z, 1.0 / 5.0);
    p = __builtin_fma(p, z, 1.0 / 3.0);
    const double lm = __builtin_fma(2.0 * s * z, p, 2.0 * s);
    const double ln2_hi = 6.93147180369123816490e-01;   // 0x3FE62E42FEE00000
    const double ln2_lo = 1.90821492927058770002e-10;
    return __builtin_fma((double)e, ln2_hi, __builtin_fma((double)e, ln2_lo, lm));
}

// sin and cos of a in [0, pi/4] by their Taylor series (x^23 / 23! < 1e-22).
__device__ __forceinline__ void gen_sincos_small(double a, double &s, double &c)
{
#pragma clang fp contract(off)
    const double z = a * a;
    double ps = -1.0 / 51090942171709440000.0;          // -1/21!
    ps = __builtin_fma(ps, z, 1.0 / 121645100408832000.0);   // 1/19!
    ps = __builtin_fma(ps, z, -1.0 / 355687428096000.0);     // -1/17!
    ps = __builtin_fma(ps, z, 1.0 / 1307674368000.0);        // 1/15!
    ps = __builtin_fma(ps, z, -1.0 / 6227020800.0);          // -1/13!
    ps = __builtin_fma(ps, z, 1.0 / 39916800.0);             // 1/11!
    ps = __builtin_fma(ps, z, -1.0 / 362880.0);              // -1/9!
    ps = __builtin_fma(ps, z, 1.0 / 5040.0);                 // 1/7!
    ps = __builtin_fma(ps, z, -1.0 / 120.0);                 // -1/5!
    ps = __builtin_fma(ps, z, 1.0 / 6.0);                    // 1/3!
    s = __builtin_fma(-a * z, ps, a);   // a - a z (1/3! - ...)
    double pc = 1.0 / 2432902008176640000.0;            // 1/20!
    pc = __builtin_fma(pc, z, -1.0 / 6402373705728000.0);    // -1/18!
    pc = __builtin_fma(pc, z, 1.0 / 20922789888000.0);       // 1/16!
    pc = __builtin_fma(pc, z, -1.0 / 87178291200.0);         // -1/14!
    pc = __builtin_fma(pc, z, 1.0 / 479001600.0);            // 1/12!
    pc = __builtin_fma(pc, z, -1.0 / 3628800.0);             // -1/10!
    pc = __builtin_fma(pc, z, 1.0 / 40320.0);                // 1/8!
    pc = __builtin_fma(pc, z, -1.0 / 720.0);                 // -1/6!
    pc = __builtin_fma(pc, z, 1.0 / 24.0);                   // 1/4!
    pc = __builtin_fma(pc, z, -0.5);                         // -1/2!
    c = __builtin_fma(z, pc, 1.0);
}

// sin and cos of 2 pi u, u in [0, 1) a multiple of 2^-53: quadrant q =
// floor(4u), f = 4u - q (exact); angle f pi/2, or its complement (1 - f) pi/2
// (exact) when f > 1/2, so that the series sees at most pi/4.
__device__ __forceinline__ void gen_sincos_2pi(double u, double &s, double &c)
{
#pragma clang fp contract(off)
    const double t = 4.0 * u;
    const double qd = floor(t);
    const int q = (int)qd;
    const double f = t - qd;
    const double pio2_hi = 1.57079632679489655800e+00;   // 0x3FF921FB54442D18
    const double pio2_lo = 6.12323399573676603587e-17;
    double ss, cc;
    if (f > 0.5) {
        const double g = 1.0 - f;
        gen_sincos_small(__builtin_fma(g, pio2_hi, g * pio2_lo), cc, ss);
    } else {
        gen_sincos_small(__builtin_fma(f, pio2_hi, f * pio2_lo), ss, cc);
    }
    switch (q & 3) {
    case 0: s = ss; c = cc; break;
    case 1: s = cc; c = -ss; break;
    case 2: s = -ss; c = -cc; break;
    default: s = -cc; c = ss; break;
    }
}

// The standard normal pair of mode counter c.
__device__ __forceinline__ void gen_normal_pair(u64 c, u64 seed, double &g0, double &g1)
{
#pragma clang fp contract(off)
    u64 w[4];
    philox4x64_10(c, 0, 0, 0, seed, 0, w);
    const double u1 = (double)(w[0] >> 11) * (1.0 / 9007199254740992.0);
    const double u2 = (double)(w[1] >> 11) * (1.0 / 9007199254740992.0);
    const double r = sqrt(-2.0 * gen_log(1.0 - u1));
    double s, co;
    gen_sincos_2pi(u2, s, co);
    g0 = r * co;
    g1 = r * s;
}

// --------------------------------------------------------------- kernels --
struct GenParams {
    int NN, H;
    u64 seed;
    double vol;   // V for the field (Make_Image divides by it), 1 for the modes stage
};

// Fill_Modes (and Make_Image's 1/V): one thread per mode of [NN][NN][H].
// amp[k2]: sqrt(V P(k) / 2) by k2 = ia^2 + ja^2 + kz^2 (0 at k = 0 and
// beyond the cut).
__global__ __launch_bounds__(256) void k_fill_modes(double2 *__restrict__ A,
                                                    const double *__restrict__ amp, GenParams P)
{
#pragma clang fp contract(off)
    const long long n = (long long)P.NN * P.NN * P.H;
    const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    const int NN = P.NN, H = P.H;
    const int k = (int)(c % H);
    const long long ij = c / H;
    const int j = (int)(ij % NN), i = (int)(ij / NN);
    const long long ia = i <= NN / 2 ? i : i - NN, ja = j <= NN / 2 ? j : j - NN;
    const double a = amp[ia * ia + ja * ja + (long long)k * k];
    double g0, g1;
    double2 v;
    if (k == 0 || k == NN / 2) {
        // Hermitian plane: mode (i, j) and its mirror (-i, -j) are conjugate;
        // the smaller index draws, the larger takes the conjugate.
        const int im = (NN - i) % NN, jm = (NN - j) % NN;
        const long long cm = ((long long)im * NN + jm) * H + k;
        if (cm == c) {
            gen_normal_pair((u64)c, P.seed, g0, g1);
            v = make_double2(a * g0, 0.0);
        } else if (c < cm) {
            gen_normal_pair((u64)c, P.seed, g0, g1);
            v = make_double2(a * g0, a * g1);
        } else {
            gen_normal_pair((u64)cm, P.seed, g0, g1);
            v = make_double2(a * g0, -(a * g1));
        }
    } else {
        gen_normal_pair((u64)c, P.seed, g0, g1);
        v = make_double2(a * g0, a * g1);
    }
    A[c] = make_double2(v.x / P.vol, v.y / P.vol);
}

constexpr int kColTile = 4;   // columns per workgroup of the axis passes

// Transforms of length N along an axis with element stride es (double2),
// lines (o, t) starting at o * os + t, t in [0, nt) contiguous; one
// workgroup per (o, tile of kColTile t).  sign > 0: exp(+i); sign < 0:
// exp(-i) (by conjugation around the backward transform).
__global__ __launch_bounds__(256) void k_axis_fft(double2 *__restrict__ buf, bb::FftDesc fd,
                                                  const double2 *__restrict__ twg, long long es,
                                                  long long os, int nt, int sign)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int N = fd.N;
    double2 *tw = reinterpret_cast<double2 *>(smem);
    double2 *a = tw + N, *b = a + kColTile * N;
    for (int t = threadIdx.x; t < N; t += blockDim.x) tw[t] = twg[t];
    const int o = blockIdx.y, t0 = blockIdx.x * kColTile;
    const int nc = min(kColTile, nt - t0);
    double2 *base = buf + (long long)o * os + t0;
    for (int e = threadIdx.x; e < N * kColTile; e += blockDim.x) {
        const int col = e % kColTile, idx = e / kColTile;
        double2 v = make_double2(0.0, 0.0);
        if (col < nc) v = base[(long long)idx * es + col];
        if (sign < 0) v.y = -v.y;
        a[col * N + idx] = v;
    }
    __syncthreads();
    double2 *r = bb::lds_fft(a, b, kColTile, fd, tw);
    for (int e = threadIdx.x; e < N * kColTile; e += blockDim.x) {
        const int col = e % kColTile, idx = e / kColTile;
        if (col < nc) {
            double2 v = r[col * N + idx];
            if (sign < 0) v.y = -v.y;
            base[(long long)idx * es + col] = v;
        }
    }
}

constexpr int kRowTile = 4;   // rows per workgroup of the real-axis passes

// Complex-to-real along the last axis, in place: row (i, j) holds H complex
// modes and becomes NN reals (the Hermitian extension X[N - k] = conj X[k],
// backward transform, real part).
__global__ __launch_bounds__(256) void k_c2r_rows(double *__restrict__ buf, bb::FftDesc fd,
                                                  const double2 *__restrict__ twg, long long nrows)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int N = fd.N, H = N / 2 + 1;
    double2 *tw = reinterpret_cast<double2 *>(smem);
    double2 *a = tw + N, *b = a + kRowTile * N;
    for (int t = threadIdx.x; t < N; t += blockDim.x) tw[t] = twg[t];
    const long long r0 = (long long)blockIdx.x * kRowTile;
    for (int e = threadIdx.x; e < kRowTile * N; e += blockDim.x) {
        const int rr = e / N, k = e % N;
        double2 v = make_double2(0.0, 0.0);
        if (r0 + rr < nrows) {
            const double2 *row = reinterpret_cast<const double2 *>(buf + (r0 + rr) * (long long)(N + 2));
            if (k < H) {
                v = row[k];
            } else {
                v = row[N - k];
                v.y = -v.y;
            }
        }
        a[rr * N + k] = v;
    }
    __syncthreads();
    double2 *r = bb::lds_fft(a, b, kRowTile, fd, tw);
    for (int e = threadIdx.x; e < kRowTile * N; e += blockDim.x) {
        const int rr = e / N, k = e % N;
        if (r0 + rr < nrows) buf[(r0 + rr) * (long long)(N + 2) + k] = r[rr * N + k].x;
    }
}

// Real-to-complex along the last axis, in place: NN reals of row (i, j) ->
// the H modes of exp(-2 pi i k x / N), each times `mul` (LL^3 of Get_Modes).
__global__ __launch_bounds__(256) void k_r2c_rows(double *__restrict__ buf, bb::FftDesc fd,
                                                  const double2 *__restrict__ twg, long long nrows,
                                                  double mul)
{
#pragma clang fp contract(off)
    extern __shared__ __align__(16) unsigned char smem[];
    const int N = fd.N, H = N / 2 + 1;
    double2 *tw = reinterpret_cast<double2 *>(smem);
    double2 *a = tw + N, *b = a + kRowTile * N;
    for (int t = threadIdx.x; t < N; t += blockDim.x) tw[t] = twg[t];
    const long long r0 = (long long)blockIdx.x * kRowTile;
    for (int e = threadIdx.x; e < kRowTile * N; e += blockDim.x) {
        const int rr = e / N, k = e % N;
        double x = 0.0;
        if (r0 + rr < nrows) x = buf[(r0 + rr) * (long long)(N + 2) + k] * mul;
        a[rr * N + k] = make_double2(x, 0.0);   // conj of a real input is itself
    }
    __syncthreads();
    double2 *r = bb::lds_fft(a, b, kRowTile, fd, tw);
    for (int e = threadIdx.x; e < kRowTile * H; e += blockDim.x) {
        const int rr = e / H, k = e % H;
        if (r0 + rr < nrows) {
            double2 v = r[rr * N + k];
            v.y = -v.y;   // exp(-i): the conjugate of the backward transform of a real row
            reinterpret_cast<double2 *>(buf + (r0 + rr) * (long long)(N + 2))[k] = v;
        }
    }
}

constexpr int kSumBlocks = 1024;

// Sum of delta^2 over the NN^3 reals (row padding skipped): per-workgroup
// partial sums, then one fixed-order sum (deterministic).
__global__ __launch_bounds__(256) void k_sumsq(const double *__restrict__ buf, int N,
                                               double *__restrict__ part)
{
    __shared__ double red[4];
    const long long n = (long long)N * N * N;
    double acc = 0.0;
    for (long long x = (long long)blockIdx.x * 256 + threadIdx.x; x < n;
         x += (long long)gridDim.x * 256) {
        const long long row = x / N, k = x - row * N;
        const double v = buf[row * (N + 2) + k];
        acc += v * v;
    }
    acc = bb::wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

__global__ __launch_bounds__(256) void k_sum_parts(const double *__restrict__ part, int n,
                                                   double *__restrict__ out)
{
    __shared__ double red[4];
    double acc = 0.0;
    for (int t = threadIdx.x; t < n; t += 256) acc += part[t];
    acc = bb::wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) out[0] = ((red[0] + red[1]) + red[2]) + red[3];
}

// Non_Gauss: delta += fnl (delta^2 - var), var = sum / NN^3.
__global__ __launch_bounds__(256) void k_nongauss(double *__restrict__ buf, int N,
                                                  const double *__restrict__ sum, double fnl)
{
#pragma clang fp contract(off)
    const long long n = (long long)N * N * N;
    const double var = sum[0] / ((double)N * N * N);
    for (long long x = (long long)blockIdx.x * 256 + threadIdx.x; x < n;
         x += (long long)gridDim.x * 256) {
        const long long row = x / N, k = x - row * N;
        const double v = buf[row * (N + 2) + k];
        buf[row * (N + 2) + k] = v + fnl * (v * v - var);
    }
}

// RSD: Delta(k) *= 1 + beta mu^2, mu = kz / |k| (ngrf_2d_funcs.c:749-760).
__global__ __launch_bounds__(256) void k_rsd(double2 *__restrict__ A, int NN, double beta)
{
#pragma clang fp contract(off)
    const int H = NN / 2 + 1;
    const long long n = (long long)NN * NN * H;
    const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
    if (c >= n || c == 0) return;
    const int k = (int)(c % H);
    const long long ij = c / H;
    const int j = (int)(ij % NN), i = (int)(ij / NN);
    const double ia = i > NN / 2 ? NN - i : i, ja = j > NN / 2 ? NN - j : j;
    const double kk = sqrt(1. * (ia * ia + ja * ja + (double)k * k));
    const double myu = k / kk;
    const double f = 1. + beta * myu * myu;
    double2 v = A[c];
    A[c] = make_double2(v.x * f, v.y * f);
}

// Per-device stream, tables and partial sums, kept until the process ends;
// calls are serialised.
struct GenState {
    hipStream_t stream = nullptr;
};
std::mutex gen_mu;
std::vector<GenState> gen_state;

int check(hipError_t e, const char *what)
{
    if (e == hipSuccess) return BB_OK;
    (void)hipGetLastError();
    return bb::set_error(BB_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace

extern "C" int bb_generate_field(int NN, unsigned long long seed, const bb_field_params *fp,
                                 double *field_r2c, int on_device)
{
    using bb::set_error;
    if (!fp || !field_r2c) return set_error(BB_E_INVALID, "null argument");
    bb::FftDesc fd;
    if (NN < 4 || (NN & 1) || NN > 1024 || !bb::fft_desc(NN, fd))
        return set_error(BB_E_INVALID, "NN must be even, 4..1024, with prime factors 2, 3, 5");
    if (fp->stage != 0 && fp->stage != 1) return set_error(BB_E_INVALID, "stage must be 0 or 1");
    if (!(fp->LL > 0) || !(fp->kunit > 0) || !std::isfinite(fp->amp) || !std::isfinite(fp->index) ||
        !std::isfinite(fp->fnl) || !std::isfinite(fp->beta))
        return set_error(BB_E_INVALID, "field parameters must be finite, LL and kunit positive");
    int dev = 0, ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1 || hipGetDevice(&dev) != hipSuccess) {
        (void)hipGetLastError();
        return set_error(BB_E_DEVICE, "no HIP device");
    }
    std::lock_guard<std::mutex> lk(gen_mu);
    if ((int)gen_state.size() < ndev) gen_state.resize(ndev);
    GenState &st = gen_state[dev];
    int rc = BB_OK;
    if (!st.stream && (rc = check(hipStreamCreateWithFlags(&st.stream, hipStreamNonBlocking),
                                  "stream")))
        return rc;
    hipStream_t s = st.stream;
    const int H = NN / 2 + 1;
    const long long nmodes = (long long)NN * NN * H, nrows = (long long)NN * NN;
    const double V = std::pow(NN * fp->LL, 3.0);

    // amp[k2] = sqrt(V/2 P(k)), P(k) = amp k^index with k = kunit sqrt(k2),
    // zero at k = 0 and for k >= kcut (kcut <= 0: no cut).
    const long long nk2 = 3LL * (NN / 2) * (NN / 2) + 1;
    std::vector<double> amp((size_t)nk2, 0.0);
    for (long long k2 = 1; k2 < nk2; ++k2) {
        const double k = fp->kunit * std::sqrt((double)k2);
        if (fp->kcut > 0 && !(k < fp->kcut)) continue;
        amp[(size_t)k2] = std::sqrt(V / 2. * (fp->amp * std::pow(k, fp->index)));
    }
    const std::vector<double2> tw = bb::fft_twiddles(NN);

    double2 *d_buf = nullptr, *d_tw = nullptr;
    double *d_amp = nullptr, *d_part = nullptr;
    auto fail = [&](int code, const std::string &msg) {
        if (!rc) rc = set_error(code, msg);
    };
    if (on_device) {
        d_buf = reinterpret_cast<double2 *>(field_r2c);
    } else if (hipMalloc(&d_buf, sizeof(double2) * nmodes) != hipSuccess) {
        (void)hipGetLastError();
        d_buf = nullptr;
        fail(BB_E_DEVMEM, "field buffer");
    }
    if (!rc && (hipMalloc(&d_amp, sizeof(double) * nk2) != hipSuccess ||
                hipMalloc(&d_tw, sizeof(double2) * NN) != hipSuccess ||
                hipMalloc(&d_part, sizeof(double) * (kSumBlocks + 1)) != hipSuccess)) {
        (void)hipGetLastError();
        fail(BB_E_DEVMEM, "generator tables");
    }
    if (!rc) rc = check(hipMemcpyAsync(d_amp, amp.data(), sizeof(double) * nk2,
                                       hipMemcpyHostToDevice, s), "table upload");
    if (!rc) rc = check(hipMemcpyAsync(d_tw, tw.data(), sizeof(double2) * NN,
                                       hipMemcpyHostToDevice, s), "twiddle upload");
    if (!rc) {
        GenParams P{NN, H, seed, fp->stage == 1 ? 1.0 : V};
        const unsigned nb = (unsigned)((nmodes + 255) / 256);
        hipLaunchKernelGGL(k_fill_modes, dim3(nb), dim3(256), 0, s, d_buf, d_amp, P);
        rc = check(hipGetLastError(), "mode fill");
    }
    if (!rc && fp->stage == 0) {
        const size_t lds_col = sizeof(double2) * ((size_t)NN + 2 * (size_t)kColTile * NN);
        const size_t lds_row = sizeof(double2) * ((size_t)NN + 2 * (size_t)kRowTile * NN);
        (void)hipFuncSetAttribute((const void *)k_axis_fft,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_col);
        (void)hipFuncSetAttribute((const void *)k_c2r_rows,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_row);
        (void)hipFuncSetAttribute((const void *)k_r2c_rows,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_row);
        const dim3 gcol((H + kColTile - 1) / kColTile, NN);
        const unsigned grow = (unsigned)((nrows + kRowTile - 1) / kRowTile);
        const long long sI = (long long)NN * H, sJ = H;
        // Make_Image: backward transform along i, j (complex), then k (c2r)
        hipLaunchKernelGGL(k_axis_fft, gcol, dim3(256), lds_col, s, d_buf, fd, d_tw, sI, sJ, H, +1);
        hipLaunchKernelGGL(k_axis_fft, gcol, dim3(256), lds_col, s, d_buf, fd, d_tw, sJ, sI, H, +1);
        hipLaunchKernelGGL(k_c2r_rows, dim3(grow), dim3(256), lds_row, s, (double *)d_buf, fd,
                           d_tw, nrows);
        // Non_Gauss
        hipLaunchKernelGGL(k_sumsq, dim3(kSumBlocks), dim3(256), 0, s, (const double *)d_buf, NN,
                           d_part);
        hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, s, d_part, kSumBlocks,
                           d_part + kSumBlocks);
        hipLaunchKernelGGL(k_nongauss, dim3(2048), dim3(256), 0, s, (double *)d_buf, NN,
                           d_part + kSumBlocks, fp->fnl);
        // Get_Modes: times LL^3, forward transform along k (r2c), then j, i
        hipLaunchKernelGGL(k_r2c_rows, dim3(grow), dim3(256), lds_row, s, (double *)d_buf, fd,
                           d_tw, nrows, std::pow(fp->LL, 3.));
        hipLaunchKernelGGL(k_axis_fft, gcol, dim3(256), lds_col, s, d_buf, fd, d_tw, sJ, sI, H, -1);
        hipLaunchKernelGGL(k_axis_fft, gcol, dim3(256), lds_col, s, d_buf, fd, d_tw, sI, sJ, H, -1);
        // RSD
        const unsigned nb = (unsigned)((nmodes + 255) / 256);
        hipLaunchKernelGGL(k_rsd, dim3(nb), dim3(256), 0, s, d_buf, NN, fp->beta);
        rc = check(hipGetLastError(), "generator transforms");
    }
    if (!rc && !on_device)
        rc = check(hipMemcpyAsync(field_r2c, d_buf, sizeof(double2) * nmodes,
                                  hipMemcpyDeviceToHost, s), "field download");
    if (!rc) rc = check(hipStreamSynchronize(s), "generator kernels failed");
    else (void)hipStreamSynchronize(s);
    if (!on_device && d_buf) (void)hipFree(d_buf);
    if (d_amp) (void)hipFree(d_amp);
    if (d_tw) (void)hipFree(d_tw);
    if (d_part) (void)hipFree(d_part);
    return rc;
}
