// Plane gather and batched ring-masked 2-D inverse FFTs (the plane-ring
// cache).  Replaces Df / wrapid_fast (funcs.c:88-109, 537-566) and the ring
// IFFTs of calCR (funcs.c:367-382).
//
// D_{p,r}(x) = sum_{k in ring r} Delta(kx, ky, kz=p) exp(+2 pi i k.x / NN),
// unnormalised (FFTW_BACKWARD), with the DC mode of plane 0 zeroed.
//
// Two passes per field, each a batch of 1-D Stockham FFTs in LDS (FP64 for
// FP64 caches, FP32 for FP32 caches), one launch each for the fields of
// every grid level:
//  - row pass: for each row i that can hold masked modes, gather the plane
//    row, apply the ring mask (integer radius bounds), transform along j,
//    write the row into the field's cache slot;
//  - column pass: transform along i in place (rows outside the ring's band
//    are zero), write the field to the cache and, for FP32 caches, as bf16
//    hi/lo operands, each in the forms the contraction reads it in; one
//    round of workgroups loops over the column groups, loading the next
//    group during the current one's transform.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "device.hpp"
#include "fft_lds.hpp"

namespace bb {

// For each field f of a batch (blockIdx.z = f * NG + i; fields and plane sets
// back to back), planes[p][i][j] (NG x NG) = field[i mod NN][j mod NN][p]
// (complex) for
// p < Nmax, where (i, j) on the NG grid stand for the signed modes
// ia = i <= NG/2 ? i : i - NG (likewise ja).  Modes with |ia| or |ja| above
// kmax lie outside every ring and are set to zero.  With NG = NN this is the
// plain plane gather of Df (funcs.c:97-100).  32x32 transpose tiles through
// LDS keep both the read (along p) and the write (along j) coalesced.
__global__ __launch_bounds__(256) void k_gather(const double2 *__restrict__ fields,
                                                double2 *__restrict__ planes_all, int NN, int NG,
                                                int kmax, int Nmax, int *__restrict__ zero,
                                                int nzero)
{
    __shared__ double2 tile[32][33];
    if (zero) {   // the flags to clear, one thread each over the whole grid
        const long long gid =
            ((long long)(blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x) * 256 +
            threadIdx.x;
        if (gid < nzero) zero[gid] = 0;
    }
    const int H = NN / 2 + 1;
    const int f = blockIdx.z / NG, i = blockIdx.z - f * NG;   // field of the batch, row
    const double2 *field = fields + (long long)f * NN * NN * H;
    double2 *planes = planes_all + (long long)f * Nmax * NG * NG;
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

// grid.z (NG rows per field) is limited to 65535: batches of more fields
// than fit are gathered by several launches, the first of which clears the
// flags.
hipError_t launch_gather(const double2 *fields, double2 *planes, int NN, int NG, int kmax,
                         int Nmax, int nf, hipStream_t s, int *zero, int nzero)
{
    if (NG < 1 || NG > 65535 || nf < 1) return hipErrorInvalidValue;
    const int per = 65535 / NG;
    const long long fstride = (long long)NN * NN * (NN / 2 + 1), pstride = (long long)Nmax * NG * NG;
    for (int f0 = 0; f0 < nf; f0 += per) {
        const int n = std::min(per, nf - f0);
        dim3 grid((NG + 31) / 32, (Nmax + 31) / 32, NG * n);
        int *z = f0 == 0 ? zero : nullptr;
        const int nz = f0 == 0 ? nzero : 0;
        if ((long long)grid.x * grid.y * grid.z * 256 < nz) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_gather, grid, dim3(256), 0, s, fields + f0 * fstride,
                           planes + f0 * pstride, NN, NG, kmax, Nmax, z, nz);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Threads per block of both passes, and the most points a thread loads
// (rpb * N and cpb * N stay within kPassThreads * kPassPer).
constexpr int kPassThreads = 256;
constexpr int kPassPer = 10;

// The passes transform in the cache's precision: FP64 for FP64 caches, FP32
// for FP32 caches (the FP32 paths' error budget is far above FP32 rounding).
template <typename V2>
__device__ __forceinline__ V2 cvt2(double2 v);
template <>
__device__ __forceinline__ double2 cvt2<double2>(double2 v) { return v; }
template <>
__device__ __forceinline__ float2 cvt2<float2>(double2 v)
{
    return make_float2((float)v.x, (float)v.y);
}

// Fields are stored column-major: point (i, j) of an N x N field sits at
// x = j * N + i.  (The contraction sums over all points, so the order is
// free; this one makes the column pass read and write contiguous memory.)
//
// Row pass over the tasks of every level, one block per entry of `blk`:
// block {t, x} transforms rows [x * rpb, x * rpb + rpb) of task t's row band
// and writes them into the field's cache slot (cache + off + fid * stride),
// column-major, in place of the final field.  The plane values come from the
// NP x NP plane grid (NP >= N): a mode (ia, ja) inside the ring satisfies
// |ia|, |ja| < N/2 and sits at (ia mod NP, ja mod NP) there.
template <typename C2>
__global__ __launch_bounds__(kPassThreads) void k_rowpass(const double2 *__restrict__ planes, int NP,
                                                 C2 *__restrict__ cache,
                                                 const FieldTask *__restrict__ tasks,
                                                 const int2 *__restrict__ blk,
                                                 const LevelDesc *__restrict__ lv,
                                                 const double2 *__restrict__ tw_all,
                                                 int *__restrict__ nzflag, BatchStrides bs)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    C2 *lds = reinterpret_cast<C2 *>(lds_raw);
    const int2 B = blk[blockIdx.x];
    const int fb = blockIdx.y;   // field of the batch
    planes += fb * bs.planes;
    cache += fb * bs.cache;
    if (nzflag) nzflag += fb * bs.nz;
    const FieldTask T = tasks[B.x];
    const LevelDesc &L = lv[T.level];
    const int N = L.fd.N, rpb = L.rpb;
    const int nrows = min(N, 2 * T.rowband + 1);
    const int q0 = B.y * rpb;
    if (q0 >= nrows) return;
    const int nb = min(rpb, nrows - q0);
    C2 *twl = lds, *a = lds + N, *b = a + rpb * N;
    for (int t = threadIdx.x; t < N; t += blockDim.x) twl[t] = cvt2<C2>(tw_all[L.tw_off + t]);
    const unsigned long long mN = L.fd.mN;
    // All of this thread's loads are issued before any LDS store, so they
    // are in flight together (at most kPassPer per thread).
    int nz = 0;
    double2 vals[kPassPer];
#pragma unroll
    for (int e = 0; e < kPassPer; ++e) {
        const int idx = threadIdx.x + e * kPassThreads;
        double2 v = make_double2(0.0, 0.0);
        if (idx < nb * N) {
            const int bt = (int)udiv((unsigned)idx, mN), j = idx - bt * N;
            const int q = q0 + bt;
            const int i = (q <= T.rowband) ? q : N - nrows + q;
            const int ia = (i > N / 2) ? (i - N) : i;
            const int ja = (j > N / 2) ? (j - N) : j;
            const int n = ia * ia + ja * ja;
            if (n >= T.nlo && n < T.nhi) {
                if (T.plane < 0) {
                    v = make_double2(1.0, 0.0);
                } else if (!(T.plane == 0 && i == 0 && j == 0)) {
                    const int ip = ia < 0 ? ia + NP : ia, jp = ja < 0 ? ja + NP : ja;
                    v = planes[((long long)T.plane * NP + ip) * NP + jp];
                }
            }
        }
        vals[e] = v;
    }
#pragma unroll
    for (int e = 0; e < kPassPer; ++e) {
        const int idx = threadIdx.x + e * kPassThreads;
        if (idx < nb * N) {
            nz |= (vals[e].x != 0.0 || vals[e].y != 0.0);
            a[idx] = cvt2<C2>(vals[e]);   // a[bt * N + j] with idx = bt * N + j
        }
    }
    if (nzflag && __any(nz) && (threadIdx.x & 63) == 0) atomicOr(&nzflag[T.zid], 1);
    __syncthreads();
    C2 *res = lds_fft(a, b, nb, L.fd, twl);
    // Column-major store: consecutive threads take consecutive rows of one
    // column (runs of nb points).
    C2 *dst = cache + L.off + (long long)T.fid * L.stride;
    const int rsh = L.rpb_log2;
    for (int idx = threadIdx.x; idx < rpb * N; idx += blockDim.x) {
        const int bt = idx & (rpb - 1), j = idx >> rsh;
        if (bt >= nb) continue;
        const int q = q0 + bt;
        const int i = (q <= T.rowband) ? q : N - nrows + q;
        dst[(long long)j * N + i] = res[bt * N + j];
    }
}

// Two FP32 values as a bf16 pair, round to nearest even (one conversion
// instruction); `a` in the low half.
__device__ __forceinline__ unsigned pk_bf16(float a, float b)
{
    typedef float f32x2 __attribute__((ext_vector_type(2)));
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(unsigned, __builtin_convertvector((f32x2){a, b}, bf16x2));
}
// hi = bf16(v) and lo = bf16(v - hi) of two values, as bf16 pairs.  v - hi
// is exact in FP32 (hi holds the leading 8 bits of v).
__device__ __forceinline__ void split2(float a, float b, unsigned &hi, unsigned &lo)
{
    hi = pk_bf16(a, b);
    lo = pk_bf16(a - __uint_as_float(hi << 16), b - __uint_as_float(hi & 0xffff0000u));
}

// 8 consecutive points of the result to the cache (dst, unless null) and
// the bf16 operands (sp, unless null), as 16-byte stores.
__device__ __forceinline__ void store8(double2 *dst, __bf16 *, const double2 *v)
{
    if (!dst) return;
#pragma unroll
    for (int t = 0; t < 8; ++t) dst[t] = v[t];
}
__device__ __forceinline__ void store8(float2 *dst, __bf16 *sp, const float2 *v)
{
    float re[8], im[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        re[t] = v[t].x;
        im[t] = v[t].y;
    }
    if (dst) {
        float4 *d4 = reinterpret_cast<float4 *>(dst);
#pragma unroll
        for (int t = 0; t < 4; ++t)
            d4[t] = make_float4(re[2 * t], im[2 * t], re[2 * t + 1], im[2 * t + 1]);
    }
    if (sp) {
        // [Re hi x8][Re lo x8][Im hi x8][Im lo x8], hi = bf16(v), lo = bf16(v - hi)
        unsigned h[2][4], l[2][4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            split2(re[2 * q], re[2 * q + 1], h[0][q], l[0][q]);
            split2(im[2 * q], im[2 * q + 1], h[1][q], l[1][q]);
        }
        uint4 *s4 = reinterpret_cast<uint4 *>(sp);
        s4[0] = make_uint4(h[0][0], h[0][1], h[0][2], h[0][3]);
        s4[1] = make_uint4(l[0][0], l[0][1], l[0][2], l[0][3]);
        s4[2] = make_uint4(h[1][0], h[1][1], h[1][2], h[1][3]);
        s4[3] = make_uint4(l[1][0], l[1][1], l[1][2], l[1][3]);
    }
}

// Column pass over the tasks of every level, in place.  Entry {t, x} of
// `blk` transforms columns [x * cpb, x * cpb + cpb) of task t's field along
// i; those columns are the contiguous points [j0 * N, (j0 + cpb) * N) of the
// field (column-major), j0 * N and cpb * N multiples of 8.  Rows outside the
// ring's band were not written by the row pass and are taken as zero.
//
// The grid is one round of resident workgroups; workgroup g takes entries
// g, g + G, ...  While it transforms one entry in LDS and stores it, the
// loads of its next entry are in flight in registers.  Every level's
// twiddles (ntw values, the table of all levels) stay in LDS for the whole
// launch.
//
// The result is written 8 points at a time; with `split` set (FP32 caches)
// also as bf16 hi/lo operands at the same point offsets: for every 8
// consecutive points, 32 bf16 values [Re hi x8][Re lo x8][Im hi x8]
// [Im lo x8], with hi = bf16(v) and lo = bf16(v - hi) of the FP32 value v.
// The task's `forms` selects which of the two are written (the contraction
// reads a field in one form or both).  Past the field's N*N points (up to a
// multiple of 8, inside the padded stride) zeros are written.
template <typename C2>
__device__ __forceinline__ void colpass_load(const C2 *__restrict__ cache,
                                             const FieldTask *__restrict__ tasks,
                                             const int2 *__restrict__ blk, int nblk,
                                             const LevelDesc *__restrict__ lv, int bi,
                                             C2 (&vals)[kPassPer])
{
#pragma unroll
    for (int e = 0; e < kPassPer; ++e) vals[e] = mk<C2>(0, 0);
    if (bi >= nblk) return;
    const int2 B = blk[bi];
    const FieldTask T = tasks[B.x];
    const LevelDesc &L = lv[T.level];
    const int N = L.fd.N, cpb = L.cpb;
    const int j0 = B.y * cpb;
    if (j0 >= N) return;
    const int nrows = min(N, 2 * T.rowband + 1);
    const C2 *fld = cache + L.off + (long long)T.fid * L.stride + (long long)j0 * N;
    const int npt = min(cpb, N - j0) * N;
    const unsigned long long mN = L.fd.mN;
#pragma unroll
    for (int e = 0; e < kPassPer; ++e) {
        const int idx = threadIdx.x + e * kPassThreads;
        if (idx < npt) {
            const int c = (int)udiv((unsigned)idx, mN), i = idx - c * N;
            const int ia = (i > N / 2) ? (i - N) : i;
            const bool inband = (ia <= T.rowband && -ia <= T.rowband) || nrows == N;
            if (inband) vals[e] = fld[idx];
        }
    }
}

template <typename C2>
__global__ __launch_bounds__(kPassThreads) void k_colpass(C2 *__restrict__ cache,
                                                 __bf16 *__restrict__ split,
                                                 const FieldTask *__restrict__ tasks,
                                                 const int2 *__restrict__ blk, int nblk,
                                                 const LevelDesc *__restrict__ lv,
                                                 const double2 *__restrict__ tw_all, int ntw,
                                                 BatchStrides bs)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    C2 *twl_all = reinterpret_cast<C2 *>(lds_raw);
    C2 *abuf = twl_all + ((ntw + 1) & ~1);   // 16-byte aligned
    const int fb = blockIdx.y;   // field of the batch
    cache += fb * bs.cache;
    if (split) split += fb * bs.cache * 4;
    for (int t = threadIdx.x; t < ntw; t += blockDim.x) twl_all[t] = cvt2<C2>(tw_all[t]);
    C2 cur[kPassPer], nxt[kPassPer];
    int bi = blockIdx.x;
    colpass_load<C2>(cache, tasks, blk, nblk, lv, bi, cur);
    for (; bi < nblk; bi += gridDim.x) {
        colpass_load<C2>(cache, tasks, blk, nblk, lv, bi + gridDim.x, nxt);
        const int2 B = blk[bi];
        const FieldTask T = tasks[B.x];
        const LevelDesc &L = lv[T.level];
        const int N = L.fd.N, cpb = L.cpb;
        const int j0 = B.y * cpb;
        if (j0 < N) {   // uniform over the workgroup
            const long long base = L.off + (long long)T.fid * L.stride + (long long)j0 * N;
            const int npt = min(cpb, N - j0) * N;
            C2 *fld = cache + base;
            C2 *a = abuf, *b = abuf + cpb * N;
#pragma unroll
            for (int e = 0; e < kPassPer; ++e) {
                const int idx = threadIdx.x + e * kPassThreads;
                if (idx < cpb * N) a[idx] = cur[e];
            }
            __syncthreads();
            C2 *res = lds_fft(a, b, cpb, L.fd, twl_all + L.tw_off);
            const int ngroups = (npt + 7) >> 3;
            const bool own = T.forms & 1, ops = split && (T.forms & 2);
            for (int g = threadIdx.x; g < ngroups; g += blockDim.x) {
                C2 v[8];
                if (8 * g + 8 <= npt) {   // a whole group: no per-point test
#pragma unroll
                    for (int t = 0; t < 8; ++t) v[t] = res[8 * g + t];
                } else {
#pragma unroll
                    for (int t = 0; t < 8; ++t) {
                        const int x = 8 * g + t;
                        v[t] = x < npt ? res[x] : mk<C2>(0, 0);
                    }
                }
                store8(own ? fld + 8 * g : nullptr, ops ? split + (base + 8 * g) * 4 : nullptr, v);
            }
            __syncthreads();   // the next entry overwrites a and b
        }
#pragma unroll
        for (int e = 0; e < kPassPer; ++e) cur[e] = nxt[e];
    }
}

// Transforms per block: about 1024 points of work for 256 threads, within
// the 160 KiB of LDS up to N = 600 ((2 b + 1) N values of the cache's
// type).  A column
// block's points start and end on whole 8-point groups: N is even and the
// column batch at least 4.
int fft_rows_per_block(int N) { return N >= 256 ? 4 : N >= 128 ? 8 : N >= 64 ? 16 : 32; }
int fft_cols_per_block(int N) { return N >= 320 ? 4 : N >= 128 ? 8 : N >= 64 ? 16 : 32; }

template <typename C2>
static hipError_t launch_fft(const double2 *planes, int NP, C2 *cache, __bf16 *split,
                             const FieldTask *tasks, const FftBlocks &fb, const LevelDesc *d_lv,
                             const LevelDesc *h_lv, int nlev, const double2 *tw, int *nzflag,
                             const BatchStrides &bs, int nf, hipStream_t s)
{
    if (nf < 1 || nf > 65535) return hipErrorInvalidValue;
    size_t lds_r = 0, lds_c = 0;
    int ntw = 0;   // twiddles of all levels (the table is the levels' back to back)
    for (int l = 0; l < nlev; ++l) {
        const LevelDesc &L = h_lv[l];
        const int N = L.fd.N;
        if (L.stride < (long long)N * N || L.stride % 8 != 0 || L.rpb < 1 || L.cpb < 1)
            return hipErrorInvalidValue;
        if (L.cpb != (1 << L.cpb_log2) || L.rpb != (1 << L.rpb_log2) || (L.cpb * N) % 8 != 0 ||
            N >= 4096 || L.rpb * N > kPassThreads * kPassPer || L.cpb * N > kPassThreads * kPassPer)
            return hipErrorInvalidValue;
        lds_r = std::max(lds_r, (2 * (size_t)L.rpb + 1) * N * sizeof(C2));
        lds_c = std::max(lds_c, 2 * (size_t)L.cpb * N * sizeof(C2));
        ntw = std::max(ntw, L.tw_off + N);
    }
    lds_c += (size_t)((ntw + 1) & ~1) * sizeof(C2);
    if (fb.nrow > 0) {
        (void)hipFuncSetAttribute((const void *)k_rowpass<C2>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_r);
        hipLaunchKernelGGL(k_rowpass<C2>, dim3(fb.nrow, nf), dim3(kPassThreads), lds_r, s, planes, NP,
                           cache, tasks, fb.row, d_lv, tw, nzflag, bs);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (fb.ncol > 0) {
        (void)hipFuncSetAttribute((const void *)k_colpass<C2>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_c);
        // One round of resident workgroups, shared by the fields of the batch.
        int dev = 0, ncu = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_colpass<C2>, kPassThreads,
                                                         lds_c) != hipSuccess)
            return hipErrorInvalidValue;
        const long long round = (long long)std::max(1, ncu) * std::max(1, per_cu);
        const int gx = (int)std::max(1LL, std::min<long long>(fb.ncol, round / nf));
        hipLaunchKernelGGL(k_colpass<C2>, dim3(gx, nf), dim3(kPassThreads), lds_c, s, cache, split,
                           tasks, fb.col, fb.ncol, d_lv, tw, ntw, bs);
    }
    return hipGetLastError();
}

void fft_blocks(const FieldTask *tasks, int ntasks, const LevelDesc *h_lv, std::vector<int2> &row,
                std::vector<int2> &col)
{
    row.clear();
    col.clear();
    for (int t = 0; t < ntasks; ++t) {
        const LevelDesc &L = h_lv[tasks[t].level];
        const int N = L.fd.N;
        const int nrows = std::min(N, 2 * tasks[t].rowband + 1);
        for (int x = 0; x * L.rpb < nrows; ++x) row.push_back(make_int2(t, x));
        for (int x = 0; x * L.cpb < N; ++x) col.push_back(make_int2(t, x));
    }
}

hipError_t launch_fft_f64(const double2 *planes, int NP, double2 *cache, const FieldTask *tasks,
                          const FftBlocks &fb, const LevelDesc *d_lv, const LevelDesc *h_lv,
                          int nlev, const double2 *tw, int *nzflag, const BatchStrides &bs,
                          int nf, hipStream_t s)
{
    return launch_fft<double2>(planes, NP, cache, nullptr, tasks, fb, d_lv, h_lv, nlev, tw,
                               nzflag, bs, nf, s);
}

hipError_t launch_fft_f32(const double2 *planes, int NP, float2 *cache, void *split,
                          const FieldTask *tasks, const FftBlocks &fb, const LevelDesc *d_lv,
                          const LevelDesc *h_lv, int nlev, const double2 *tw, int *nzflag,
                          const BatchStrides &bs, int nf, hipStream_t s)
{
    return launch_fft<float2>(planes, NP, cache, (__bf16 *)split, tasks, fb, d_lv, h_lv, nlev, tw,
                              nzflag, bs, nf, s);
}

}  // namespace bb
