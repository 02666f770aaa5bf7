// Plane gather and batched ring-masked 2-D inverse FFTs (the plane-ring
// cache).  Replaces Df / wrapid_fast (funcs.c:88-109, 537-566) and the ring
// IFFTs of calCR (funcs.c:367-382).
//
// D_{p,r}(x) = sum_{k in ring r} Delta(kx, ky, kz=p) exp(+2 pi i k.x / NN),
// unnormalised (FFTW_BACKWARD), with the DC mode of plane 0 zeroed.
//
// Two passes per field, each a batch of 1-D Stockham FFTs in LDS (FP64):
//  - row pass: for each row i that can hold masked modes, gather the plane
//    row, apply the ring mask (integer radius bounds), transform along j;
//  - column pass: transform along i (rows outside the ring's band are zero),
//    write the field to the cache (FP64, or rounded to FP32).
#include <hip/hip_runtime.h>

#include "device.hpp"

namespace bb {

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// multiply by +i
__device__ __forceinline__ double2 cmuli(double2 a) { return make_double2(-a.y, a.x); }

// Length-R DFT with exp(+2 pi i q s / R), in place, for R = 2, 3, 4, 5.
template <int R>
__device__ __forceinline__ void butterfly(double2 *v);
template <>
__device__ __forceinline__ void butterfly<4>(double2 *v)
{
    double2 a = v[0], b = v[1], c = v[2], d = v[3];
    double2 s0 = cadd(a, c), d0 = csub(a, c), s1 = cadd(b, d), d1 = cmuli(csub(b, d));
    v[0] = cadd(s0, s1);
    v[1] = cadd(d0, d1);
    v[2] = csub(s0, s1);
    v[3] = csub(d0, d1);
}
template <>
__device__ __forceinline__ void butterfly<2>(double2 *v)
{
    double2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}
template <>
__device__ __forceinline__ void butterfly<3>(double2 *v)
{
    const double h = 0.86602540378443864676;  // sqrt(3)/2
    double2 a = v[0], t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
    double2 m = make_double2(a.x - 0.5 * t1.x, a.y - 0.5 * t1.y);
    double2 j = make_double2(-h * t2.y, h * t2.x);  // i*h*t2
    v[0] = cadd(a, t1);
    v[1] = cadd(m, j);
    v[2] = csub(m, j);
}
template <>
__device__ __forceinline__ void butterfly<5>(double2 *v)
{
    const double c1 = 0.30901699437494742410, c2 = -0.80901699437494742410;
    const double s1 = 0.95105651629515357212, s2 = 0.58778525229247312917;
    double2 a = v[0];
    double2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
    double2 u1 = csub(v[1], v[4]), u2 = csub(v[2], v[3]);
    double2 m1 = make_double2(a.x + c1 * t1.x + c2 * t2.x, a.y + c1 * t1.y + c2 * t2.y);
    double2 m2 = make_double2(a.x + c2 * t1.x + c1 * t2.x, a.y + c2 * t1.y + c1 * t2.y);
    // i*(s1*u1 + s2*u2) and i*(s2*u1 - s1*u2)
    double2 n1 = make_double2(-(s1 * u1.y + s2 * u2.y), s1 * u1.x + s2 * u2.x);
    double2 n2 = make_double2(-(s2 * u1.y - s1 * u2.y), s2 * u1.x - s1 * u2.x);
    v[0] = cadd(a, cadd(t1, t2));
    v[1] = cadd(m1, n1);
    v[4] = csub(m1, n1);
    v[2] = cadd(m2, n2);
    v[3] = csub(m2, n2);
}

// One Stockham stage of radix R over nb transforms of length N: reads
// x[i + t N/R], twiddles by W_{pR}^{t k} (k = i mod p), writes
// y[(i-k) R + k + t p].
template <int R>
__device__ __forceinline__ void fft_stage(const double2 *a, double2 *b, int nb, int N, int p,
                                          const double2 *__restrict__ tw)
{
    const unsigned NR = (unsigned)(N / R);
    const unsigned tstep = (unsigned)(N / (p * R));
    const unsigned total = (unsigned)nb * NR;
    for (unsigned idx = threadIdx.x; idx < total; idx += blockDim.x) {
        const unsigned bt = idx / NR;
        const unsigned i = idx - bt * NR;
        const unsigned k = i % (unsigned)p;
        const double2 *x = a + bt * (unsigned)N;
        double2 *y = b + bt * (unsigned)N;
        double2 v[R];
#pragma unroll
        for (int t = 0; t < R; ++t) v[t] = x[i + t * NR];
        if (k > 0) {
#pragma unroll
            for (int t = 1; t < R; ++t) v[t] = cmul(v[t], tw[t * k * tstep]);
        }
        butterfly<R>(v);
        const unsigned j = (i - k) * R + k;
#pragma unroll
        for (int t = 0; t < R; ++t) y[j + t * (unsigned)p] = v[t];
    }
}

// nb transforms of length N stored back to back in a (LDS); b is scratch of
// the same size.  Returns the buffer holding the result (self-sorting
// Stockham, radices in fd.radix order).
__device__ double2 *lds_fft(double2 *a, double2 *b, int nb, const FftDesc &fd,
                            const double2 *__restrict__ tw)
{
    const int N = fd.N;
    int p = 1;
    for (int s = 0; s < fd.nst; ++s) {
        const int R = fd.radix[s];
        switch (R) {
        case 4: fft_stage<4>(a, b, nb, N, p, tw); break;
        case 2: fft_stage<2>(a, b, nb, N, p, tw); break;
        case 3: fft_stage<3>(a, b, nb, N, p, tw); break;
        default: fft_stage<5>(a, b, nb, N, p, tw); break;
        }
        __syncthreads();
        double2 *tmp = a; a = b; b = tmp;
        p *= R;
    }
    return a;
}

// planes[p][i][j] (NG x NG) = field[i mod NN][j mod NN][p] (complex) for
// p < Nmax, where (i, j) on the NG grid stand for the signed modes
// ia = i <= NG/2 ? i : i - NG (likewise ja).  Modes with |ia| or |ja| above
// kmax lie outside every ring and are set to zero.  With NG = NN this is the
// plain plane gather of Df (funcs.c:97-100).  32x32 transpose tiles through
// LDS keep both the read (along p) and the write (along j) coalesced.
__global__ __launch_bounds__(256) void k_gather(const double2 *__restrict__ field,
                                                double2 *__restrict__ planes, int NN, int NG,
                                                int kmax, int Nmax)
{
    __shared__ double2 tile[32][33];
    const int H = NN / 2 + 1;
    const int i = blockIdx.z;
    const int ia = (i > NG / 2) ? i - NG : i;
    const bool rowok = (ia <= kmax && -ia <= kmax) || NG == NN;
    const int is = ia < 0 ? ia + NN : ia;
    const int j0 = blockIdx.x * 32, p0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int jj = ty; jj < 32; jj += 8) {
        const int j = j0 + jj, p = p0 + tx;
        if (j < NG && p < Nmax) {
            const int ja = (j > NG / 2) ? j - NG : j;
            const bool ok = rowok && ((ja <= kmax && -ja <= kmax) || NG == NN);
            const int js = ja < 0 ? ja + NN : ja;
            double2 v = make_double2(0.0, 0.0);
            if (ok) v = field[((long long)is * NN + js) * H + p];
            tile[jj][tx] = v;
        }
    }
    __syncthreads();
    for (int pp = ty; pp < 32; pp += 8) {
        const int p = p0 + pp, j = j0 + tx;
        if (j < NG && p < Nmax) planes[((long long)p * NG + i) * NG + j] = tile[tx][pp];
    }
}

hipError_t launch_gather(const double2 *field, double2 *planes, int NN, int NG, int kmax,
                         int Nmax, hipStream_t s)
{
    dim3 grid((NG + 31) / 32, (Nmax + 31) / 32, NG);
    hipLaunchKernelGGL(k_gather, grid, dim3(256), 0, s, field, planes, NN, NG, kmax, Nmax);
    return hipGetLastError();
}

constexpr int kRowsPerBlock = 4;

// Row pass.  grid = (ceil(max_rows / kRowsPerBlock), ntasks).
// Output row i of task t goes to rowbuf[slot * buf_stride + i * N], with
// slot = task.fid when use_fid, else t.  The plane values come from the
// NP x NP plane grid (NP >= N): a mode (ia, ja) inside the ring satisfies
// |ia|, |ja| < N/2 and sits at (ia mod NP, ja mod NP) there.
__global__ __launch_bounds__(256) void k_rowpass(const double2 *__restrict__ planes, int NP,
                                                 double2 *__restrict__ rowbuf,
                                                 long long buf_stride, int use_fid,
                                                 const FieldTask *__restrict__ tasks, FftDesc fd,
                                                 const double2 *__restrict__ tw,
                                                 int *__restrict__ nzflag)
{
    extern __shared__ double2 lds[];
    const int N = fd.N;
    const FieldTask T = tasks[blockIdx.y];
    const int nrows = min(N, 2 * T.rowband + 1);
    const int q0 = blockIdx.x * kRowsPerBlock;
    if (q0 >= nrows) return;
    const int nb = min(kRowsPerBlock, nrows - q0);
    double2 *a = lds, *b = lds + kRowsPerBlock * N;
    int nz = 0;
    for (int idx = threadIdx.x; idx < nb * N; idx += blockDim.x) {
        const int bt = idx / N, j = idx - bt * N;
        const int q = q0 + bt;
        const int i = (q <= T.rowband) ? q : N - nrows + q;
        const int ia = (i > N / 2) ? (i - N) : i;
        const int ja = (j > N / 2) ? (j - N) : j;
        const int n = ia * ia + ja * ja;
        const bool in = (n >= T.nlo && n < T.nhi);
        double2 v = make_double2(0.0, 0.0);
        if (in) {
            if (T.plane < 0) {
                v = make_double2(1.0, 0.0);
            } else if (!(T.plane == 0 && i == 0 && j == 0)) {
                const int ip = ia < 0 ? ia + NP : ia, jp = ja < 0 ? ja + NP : ja;
                v = planes[((long long)T.plane * NP + ip) * NP + jp];
            }
        }
        nz |= (v.x != 0.0 || v.y != 0.0);
        a[bt * N + j] = v;
    }
    if (nzflag && __any(nz) && (threadIdx.x & 63) == 0) atomicOr(&nzflag[T.zid], 1);
    __syncthreads();
    double2 *res = lds_fft(a, b, nb, fd, tw);
    const long long slot = use_fid ? T.fid : blockIdx.y;
    for (int idx = threadIdx.x; idx < nb * N; idx += blockDim.x) {
        const int bt = idx / N, j = idx - bt * N;
        const int q = q0 + bt;
        const int i = (q <= T.rowband) ? q : N - nrows + q;
        rowbuf[slot * buf_stride + (long long)i * N + j] = res[bt * N + j];
    }
}

hipError_t launch_rowpass(const double2 *planes, int NP, double2 *rowbuf, long long buf_stride,
                          const FieldTask *tasks, int ntasks, int max_rows, const FftDesc &fd,
                          const double2 *tw, int *nzflag, hipStream_t s)
{
    // rowbuf == nullptr is not allowed; use_fid is encoded by the caller via
    // buf_stride sign: negative stride means slot = task index.
    int use_fid = buf_stride > 0 ? 1 : 0;
    long long stride = buf_stride > 0 ? buf_stride : -buf_stride;
    dim3 grid((max_rows + kRowsPerBlock - 1) / kRowsPerBlock, ntasks);
    size_t lds = 2 * kRowsPerBlock * fd.N * sizeof(double2);
    (void)hipFuncSetAttribute((const void *)k_rowpass, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds);
    hipLaunchKernelGGL(k_rowpass, grid, dim3(256), lds, s, planes, NP, rowbuf, stride, use_fid,
                       tasks, fd, tw, nzflag);
    return hipGetLastError();
}

constexpr int kColsPerBlock = 8;

template <typename OUT2>
__device__ __forceinline__ OUT2 to_out(double2 v);
template <>
__device__ __forceinline__ double2 to_out<double2>(double2 v) { return v; }
template <>
__device__ __forceinline__ float2 to_out<float2>(double2 v)
{
    return make_float2((float)v.x, (float)v.y);
}

// Column pass.  grid = (ceil(N / kColsPerBlock), ntasks).  Reads the row-pass
// output of task t (slot as in the row pass), transforms along i and writes
// point i*N+j of field fid at cache + fid * out_stride.  Rows outside the
// ring's band were never written by the row pass and are taken as zero.
// With `split` set (FP32 caches), the field is also written as bf16 hi/lo
// operands: for every 8 consecutive points, 32 bf16 values
// [Re hi x8][Re lo x8][Im hi x8][Im lo x8], with hi = bf16(v) and
// lo = bf16(v - hi) of the FP32 value v; out_stride is a multiple of 8, so
// no group of 8 straddles two fields (points past N*N stay as they are).
template <typename OUT2>
__global__ __launch_bounds__(256) void k_colpass(const double2 *__restrict__ rowbuf,
                                                 long long buf_stride, int use_fid,
                                                 OUT2 *__restrict__ cache,
                                                 __bf16 *__restrict__ split,
                                                 long long out_stride,
                                                 const FieldTask *__restrict__ tasks, FftDesc fd,
                                                 const double2 *__restrict__ tw)
{
    extern __shared__ double2 lds[];
    const int N = fd.N;
    const FieldTask T = tasks[blockIdx.y];
    const int nrows = min(N, 2 * T.rowband + 1);
    const int j0 = blockIdx.x * kColsPerBlock;
    const long long slot = use_fid ? T.fid : blockIdx.y;
    const double2 *src = rowbuf + slot * buf_stride;
    double2 *a = lds, *b = lds + kColsPerBlock * N;
    // layout a[c * N + i]
    for (int idx = threadIdx.x; idx < kColsPerBlock * N; idx += blockDim.x) {
        const int i = idx / kColsPerBlock, c = idx - i * kColsPerBlock;
        const int ia = (i > N / 2) ? (i - N) : i;
        const bool inband = (ia <= T.rowband && -ia <= T.rowband) || nrows == N;
        double2 v = make_double2(0.0, 0.0);
        if (inband && j0 + c < N) v = src[(long long)i * N + j0 + c];
        a[c * N + i] = v;
    }
    __syncthreads();
    double2 *res = lds_fft(a, b, kColsPerBlock, fd, tw);
    OUT2 *dst = cache + (long long)T.fid * out_stride;
    for (int idx = threadIdx.x; idx < kColsPerBlock * N; idx += blockDim.x) {
        const int i = idx / kColsPerBlock, c = idx - i * kColsPerBlock;
        if (j0 + c >= N) continue;
        const long long x = (long long)i * N + j0 + c;
        const OUT2 o = to_out<OUT2>(res[c * N + i]);
        dst[x] = o;
        if (split) {
            const float vr = (float)o.x, vi = (float)o.y;
            const __bf16 hr = (__bf16)vr, hi = (__bf16)vi;
            __bf16 *g = split + (long long)T.fid * out_stride * 4 + (x >> 3) * 32 + (x & 7);
            g[0] = hr;
            g[8] = (__bf16)(vr - (float)hr);
            g[16] = hi;
            g[24] = (__bf16)(vi - (float)hi);
        }
    }
}

template <typename OUT2>
static hipError_t launch_colpass(const double2 *rowbuf, long long buf_stride, OUT2 *cache,
                                 __bf16 *split, long long out_stride, const FieldTask *tasks,
                                 int ntasks, const FftDesc &fd, const double2 *tw, hipStream_t s)
{
    if (out_stride < (long long)fd.N * fd.N || out_stride % 8 != 0) return hipErrorInvalidValue;
    int use_fid = buf_stride > 0 ? 1 : 0;
    long long stride = buf_stride > 0 ? buf_stride : -buf_stride;
    dim3 grid((fd.N + kColsPerBlock - 1) / kColsPerBlock, ntasks);
    size_t lds = 2 * kColsPerBlock * fd.N * sizeof(double2);
    (void)hipFuncSetAttribute((const void *)k_colpass<OUT2>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k_colpass<OUT2>, grid, dim3(256), lds, s, rowbuf, stride, use_fid, cache,
                       split, out_stride, tasks, fd, tw);
    return hipGetLastError();
}

hipError_t launch_colpass_f64(const double2 *rowbuf, long long buf_stride, double2 *cache,
                              long long out_stride, const FieldTask *tasks, int ntasks,
                              const FftDesc &fd, const double2 *tw, hipStream_t s)
{
    return launch_colpass<double2>(rowbuf, buf_stride, cache, nullptr, out_stride, tasks, ntasks,
                                   fd, tw, s);
}

hipError_t launch_colpass_f32(const double2 *rowbuf, long long buf_stride, float2 *cache,
                              void *split, long long out_stride, const FieldTask *tasks,
                              int ntasks, const FftDesc &fd, const double2 *tw, hipStream_t s)
{
    return launch_colpass<float2>(rowbuf, buf_stride, cache, (__bf16 *)split, out_stride, tasks,
                                  ntasks, fd, tw, s);
}

}  // namespace bb
