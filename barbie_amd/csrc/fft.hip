// Plane gather and the ring pass: batched ring-masked 2-D inverse FFTs (the
// plane-ring cache).  Replaces Df / wrapid_fast (funcs.c:88-109, 537-566)
// and the ring IFFTs of calCR (funcs.c:367-382).
//
// D_{p,r}(x) = sum_{k in ring r} Delta(kx, ky, kz=p) exp(+2 pi i k.x / NN),
// unnormalised (FFTW_BACKWARD), with the DC mode of plane 0 zeroed.
//
// One launch builds the fields of every grid level (FP64 for FP64 caches,
// FP32 for the FP32 and BF16x3 caches).  A workgroup takes a range of one
// field's column blocks: the transform along j is evaluated at the block's
// columns as direct sums over the ring's few modes per row (the ring's
// modes staged once in LDS as +-ja pairs), the transform along i is a batch
// of LDS Stockham FFTs, and the block is written in the forms the
// contraction reads it in (FP32 values and/or bf16 hi/lo operands).  No
// intermediate goes through memory.
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
                                                int nzero, int *__restrict__ one, int none)
{
    __shared__ double2 tile[32][33];
    if (zero) {   // the flags to clear (and set), one thread each over the whole grid
        const long long gid =
            ((long long)(blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x) * 256 +
            threadIdx.x;
        if (gid < nzero) zero[gid] = 0;
        if (one && gid < none) one[gid] = 1;
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
                         int Nmax, int nf, hipStream_t s, int *zero, int nzero, int *one)
{
    if (NG < 1 || NG > 65535 || nf < 1) return hipErrorInvalidValue;
    const int per = 65535 / NG;
    const long long fstride = (long long)NN * NN * (NN / 2 + 1), pstride = (long long)Nmax * NG * NG;
    for (int f0 = 0; f0 < nf; f0 += per) {
        const int n = std::min(per, nf - f0);
        dim3 grid((NG + 31) / 32, (Nmax + 31) / 32, NG * n);
        int *z = f0 == 0 ? zero : nullptr;
        const int nz = f0 == 0 ? nzero : 0;
        if ((long long)grid.x * grid.y * grid.z * 256 < std::max(nz, nf)) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_gather, grid, dim3(256), 0, s, fields + f0 * fstride,
                           planes + f0 * pstride, NN, NG, kmax, Nmax, z, nz, f0 == 0 ? one : nullptr,
                           nf);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// A column block of the ring pass holds at most kPassMaxPoints points
// (cpb * N).
constexpr int kPassMaxPoints = 4096;

// The ring pass transforms in the cache's precision: FP64 for FP64 caches,
// FP32 for FP32 caches (the FP32 paths' error budget is far above FP32
// rounding).
template <typename V2>
__device__ __forceinline__ V2 cvt2(double2 v);
template <>
__device__ __forceinline__ double2 cvt2<double2>(double2 v) { return v; }
template <>
__device__ __forceinline__ float2 cvt2<float2>(double2 v)
{
    return make_float2((float)v.x, (float)v.y);
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

// Fields are stored column-major: point (i, j) of an N x N field sits at
// x = j * N + i.  (The contraction sums over all points, so the order is
// free; this one makes every column block one contiguous range.)
//
// Ring pass: one entry {t, x} of `blk` builds columns [j0, j0 + cpb) of task
// t's field, j0 = x * cpb, in three steps, all in LDS:
//  1. rows: for every row i of the ring's band, the transform along j of
//     the masked plane row, evaluated at the block's columns only, as a
//     direct sum over the ring's modes on that row,
//         X[i][j] = sum_{ja : nlo <= ia^2 + ja^2 < nhi} Delta(ia, ja) w^(ja j),
//     w = exp(+2 pi i / N) (a ring holds few modes per row: the sum costs
//     about as much as the row FFT would, and nothing is written to memory
//     between the two directions); rows outside the band are zero;
//  2. columns: the cpb transforms along i (Stockham stages in LDS, radices
//     2 to 5 and their two-stage composites 6 to 16, lds_fft);
//  3. stores: the block's cpb * N contiguous points, in the cache's type
//     and, for FP32 caches with `split`, as bf16 hi/lo operands, each in the
//     forms the contraction reads the field in.
// The modes of the band rows are staged in LDS first (in groups of rows
// whose modes fit in the FFT's scratch buffer), converted to the cache's
// type; the unit fields (plane < 0) take the value 1 on the ring, and the
// DC mode of plane 0 is zero (funcs.c:101-103).  Modes (ia, ja) are taken
// for ja in (-N/2, N/2], the range a length-N row transform sees.
//
// The grid is one round of resident workgroups; workgroup g takes entries
// g, g + G, ...  Every level's twiddles stay in LDS for the launch.

// floor(sqrt(v)) for 0 <= v < 2^24.
__device__ __forceinline__ int isqrt_dev(int v)
{
    int r = (int)sqrtf((float)v);
    while (r * r > v) --r;
    while ((r + 1) * (r + 1) <= v) ++r;
    return r;
}

// The ring's modes on row ia of a length-N transform (ja in (-N/2, N/2]),
// as units of consecutive ja = ja0 .. ja0 + n - 1 >= 0: unit ja > 0 up to hn
// holds both modes +ja and -ja (they share the twiddle's cosine and sine),
// units ja = 0 and ja = N/2 hold one mode.
struct RowUnits {
    int ja0, n, hn;
};
__device__ __forceinline__ RowUnits row_units(int ia, int nlo, int nhi, int N)
{
    RowUnits u{0, 0, -1};
    const int s = ia * ia, hi2 = nhi - 1 - s;
    if (hi2 < 0) return u;
    const int jhi = isqrt_dev(hi2);
    const int lo2 = nlo - s;
    const int jlo = lo2 <= 0 ? 0 : isqrt_dev(lo2 - 1) + 1;   // least j with j^2 >= lo2
    const int hp = min(jhi, N / 2);   // positive modes: [jlo, hp]; negative: [-hn, -max(jlo, 1)]
    if (jlo > hp) return u;
    u.hn = min(jhi, N / 2 - 1);
    u.ja0 = jlo;
    u.n = hp - jlo + 1;
    return u;
}

// Sum over a row's staged units md[2u], md[2u + 1] = (A, B) of ja = ja0 + u:
// A w_r + i B w_i with w = w_N^(ja j): for a pair, A = v(+ja) + v(-ja) and
// B = v(+ja) - v(-ja), so that v(+ja) w + v(-ja) conj(w) = A Re w + i B Im w;
// for ja = 0 and ja = N/2, A = v(ja) and B = 0.
template <typename C2>
__device__ __forceinline__ C2 row_sum(const C2 *md, int n, int ja0, int j, int N,
                                      unsigned long long mN, const C2 *tw)
{
    const unsigned u = (unsigned)(ja0 * j);
    int t = (int)(u - udiv(u, mN) * (unsigned)N);
    C2 acc = mk<C2>(0, 0);
    for (int m = 0; m < n; ++m) {
        const C2 A = md[2 * m], Bv = md[2 * m + 1], w = tw[t];
        acc.x += A.x * w.x - Bv.y * w.y;
        acc.y += A.y * w.x + Bv.x * w.y;
        t += j;
        if (t >= N) t -= N;
    }
    return acc;
}

// Shape of the ring pass per cache type (measured at the NN=384 grid
// levels): FP32 workgroups of 512 threads with column blocks of up to 2048
// points and room for 1024 units (mode pairs) staged once for all of a
// workgroup's column blocks (three workgroups per CU by LDS); FP64
// workgroups of 256 threads with blocks of up to 512 points and 640 units,
// so that three FP64 workgroups also fit.  Fields with more units than that
// are staged again for every block, in groups of rows that fit in the FFT's
// scratch buffer (a row holds at most N / 2 + 1 units).
template <typename C2>
struct PassShape;
template <>
struct PassShape<float2> {
    static constexpr int threads = 512, unit_cap = 1024, block_points = 2048;
};
template <>
struct PassShape<double2> {
    static constexpr int threads = 256, unit_cap = 640, block_points = 512;
};

// Stages units [k0, k1) of a field (counted over its band rows in row
// order) into dst[2 (k - k0)], dst[2 (k - k0) + 1]; returns whether one of
// their modes is nonzero.
template <typename C2>
__device__ __forceinline__ int stage_units(const double2 *__restrict__ planes, int NP,
                                           const FieldTask &T, int N, int nrows,
                                           const int4 *rinfo, const int *roff, int g0, int g1,
                                           int k0, int k1, C2 *dst)
{
    int nz = 0;
    for (int k = k0 + (int)threadIdx.x; k < k1; k += blockDim.x) {
        int ql = g0, qh = g1 - 1;   // the row q with roff[q] <= k < roff[q + 1]
        while (ql < qh) {
            const int mid = (ql + qh + 1) >> 1;
            if (roff[mid] <= k) ql = mid;
            else qh = mid - 1;
        }
        const int q = ql;
        const int4 ri = rinfo[q];
        const int ja = ri.x + (k - roff[q]);
        const bool pair = ja > 0 && ja <= ri.z;
        const int i = q <= T.rowband ? q : N - nrows + q;
        const int ia = i > N / 2 ? i - N : i;
        double2 vp = make_double2(1.0, 0.0), vn = make_double2(pair ? 1.0 : 0.0, 0.0);
        if (T.plane >= 0) {
            const int ip = ia < 0 ? ia + NP : ia;
            const double2 *prow = planes + ((long long)T.plane * NP + ip) * NP;
            vp = (T.plane == 0 && ia == 0 && ja == 0) ? make_double2(0.0, 0.0) : prow[ja];
            vn = pair ? prow[NP - ja] : make_double2(0.0, 0.0);
        }
        nz |= (vp.x != 0.0 || vp.y != 0.0 || vn.x != 0.0 || vn.y != 0.0);
        dst[2 * (k - k0)] = cvt2<C2>(make_double2(vp.x + vn.x, vp.y + vn.y));
        dst[2 * (k - k0) + 1] = cvt2<C2>(make_double2(vp.x - vn.x, vp.y - vn.y));
    }
    return nz;
}

// Direct row sums of rows [g0, g1) at columns [j0, j0 + cpb) into a; md
// holds the units from unit k0 on.
template <typename C2>
__device__ __forceinline__ void row_sums(const int4 *rinfo, const int *roff, int g0, int g1,
                                         int k0, const C2 *md, const FieldTask &T, int N,
                                         int nrows, int j0, int cpb, int csh,
                                         unsigned long long mN, const C2 *tw, C2 *a)
{
    for (int idx = threadIdx.x; idx < (g1 - g0) * cpb; idx += blockDim.x) {
        const int q = g0 + (idx >> csh), c = idx & (cpb - 1);
        const int4 ri = rinfo[q];
        const int i = q <= T.rowband ? q : N - nrows + q;
        a[c * N + i] = row_sum(md + 2 * (roff[q] - k0), ri.y, ri.x, j0 + c, N, mN, tw);
    }
}

// One unit of work per entry of `blk`: {task, first | (end << 16)} covers
// column blocks [first, end) of the task's field (columns [x cpb, x cpb + cpb)).
template <typename C2>
__global__ __launch_bounds__(PassShape<C2>::threads) void k_ringpass(
                                                  const double2 *__restrict__ planes,
                                                  int NP, C2 *__restrict__ cache,
                                                  __bf16 *__restrict__ split,
                                                  const FieldTask *__restrict__ tasks,
                                                  const int2 *__restrict__ blk, int nblk,
                                                  const LevelDesc *__restrict__ lv,
                                                  const double2 *__restrict__ tw_all, int ntw,
                                                  int maxN, int *__restrict__ nzflag,
                                                  BatchStrides bs)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    C2 *twl_all = reinterpret_cast<C2 *>(lds_raw);
    constexpr int kUnitCap = PassShape<C2>::unit_cap;
    C2 *md = twl_all + ((ntw + 1) & ~1);                                 // kUnitCap units
    int4 *rinfo = reinterpret_cast<int4 *>(md + 2 * kUnitCap);           // per band row
    int *roff = reinterpret_cast<int *>(rinfo + maxN);                   // maxN + 1 offsets
    C2 *abuf = reinterpret_cast<C2 *>(roff + ((maxN + 4) & ~3));
    const int fb = blockIdx.y;   // field of the batch
    planes += fb * bs.planes;
    cache += fb * bs.cache;
    if (split) split += fb * bs.cache * 4;
    if (nzflag) nzflag += fb * bs.nz;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int t = tid; t < ntw; t += blockDim.x) twl_all[t] = cvt2<C2>(tw_all[t]);
    for (int bi = blockIdx.x; bi < nblk; bi += gridDim.x) {
        const int2 B = blk[bi];
        const FieldTask T = tasks[B.x];
        const LevelDesc &L = lv[T.level];
        const int N = L.fd.N, cpb = L.cpb, csh = L.cpb_log2;
        const int xb = B.y & 0xffff, xe = min(B.y >> 16, (N + cpb - 1) / cpb);
        const C2 *tw = twl_all + L.tw_off;
        const unsigned long long mN = L.fd.mN;
        const int nrows = max(0, min(N, 2 * T.rowband + 1));
        C2 *a = abuf, *b = abuf + cpb * N;   // b: FFT scratch (and unit staging of big fields)
        const int cap = cpb * N / 2;          // units b holds

        // Band rows: the ring's units per row and their offsets (wave 0).
        if (wave == 0) {
            const int per = (nrows + 63) / 64, q0 = lane * per;
            int cnt = 0;
            for (int q = q0; q < min(q0 + per, nrows); ++q) {
                const int i = q <= T.rowband ? q : N - nrows + q;
                const int ia = i > N / 2 ? i - N : i;
                const RowUnits ru = row_units(ia, T.nlo, T.nhi, N);
                rinfo[q] = make_int4(ru.ja0, ru.n, ru.hn, 0);
                cnt += ru.n;
            }
            int inc = cnt;   // inclusive prefix over the lanes
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const int v = __shfl_up(inc, d, 64);
                if (lane >= d) inc += v;
            }
            int off = inc - cnt;
            for (int q = q0; q < min(q0 + per, nrows); ++q) {
                roff[q] = off;
                off += rinfo[q].y;
            }
            if (lane == 63) roff[nrows] = inc;
        }
        __syncthreads();
        const int total = roff[nrows];
        const bool once = total <= kUnitCap;   // uniform
        int nz = 0;
        if (once) nz |= stage_units(planes, NP, T, N, nrows, rinfo, roff, 0, max(nrows, 1), 0,
                                    total, md);
        for (int x = xb; x < xe; ++x) {
            const int j0 = x * cpb;
            // Rows outside the band are zero.
            for (int idx = tid; idx < (N - nrows) * cpb; idx += blockDim.x)
                a[(idx & (cpb - 1)) * N + T.rowband + 1 + (idx >> csh)] = mk<C2>(0, 0);
            if (once) {
                __syncthreads();   // md staged (first block); b free (later blocks)
                row_sums(rinfo, roff, 0, nrows, 0, md, T, N, nrows, j0, cpb, csh, mN, tw, a);
            } else {
                // Groups of band rows whose units fit in b (a row holds at most
                // N / 2 + 1 <= cap units).
                for (int g0 = 0; g0 < nrows;) {
                    int lo = g0 + 1, hi = nrows;   // largest g1 with roff[g1] - roff[g0] <= cap
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (roff[mid] - roff[g0] <= cap) lo = mid;
                        else hi = mid - 1;
                    }
                    const int g1 = lo;
                    __syncthreads();   // b free
                    nz |= stage_units(planes, NP, T, N, nrows, rinfo, roff, g0, g1, roff[g0],
                                      roff[g1], b);
                    __syncthreads();
                    row_sums(rinfo, roff, g0, g1, roff[g0], b, T, N, nrows, j0, cpb, csh, mN, tw,
                             a);
                    g0 = g1;
                }
            }
            __syncthreads();
            C2 *res = lds_fft<C2, true>(a, b, cpb, L.fd, tw);
            const long long fbase = L.off + (long long)T.fid * L.stride + (long long)j0 * N;
            C2 *fld = cache + fbase;
            const int npt = min(cpb, N - j0) * N;
            const int ngroups = (npt + 7) >> 3;
            const bool own = T.forms & 1, ops = split && (T.forms & 2);
            for (int g = tid; g < ngroups; g += blockDim.x) {
                C2 v[8];
                if (8 * g + 8 <= npt) {   // a whole group: no per-point test
#pragma unroll
                    for (int t = 0; t < 8; ++t) v[t] = res[8 * g + t];
                } else {
#pragma unroll
                    for (int t = 0; t < 8; ++t) {
                        const int xx = 8 * g + t;
                        v[t] = xx < npt ? res[xx] : mk<C2>(0, 0);
                    }
                }
                store8(own ? fld + 8 * g : nullptr, ops ? split + (fbase + 8 * g) * 4 : nullptr,
                       v);
            }
            __syncthreads();   // the next block overwrites a and b
        }
        if (nzflag && xb == 0 && __any(nz) && lane == 0) atomicOr(&nzflag[T.zid], 1);
    }
}

// Columns per block: the largest power of two with cpb N within the block
// size of the cache type (at least 4, so that blocks of even N start on
// 8-point groups).  Larger blocks leave fewer threads idle in the FFT stages
// and cost fewer barriers per point; smaller ones fit more workgroups.
int fft_cols_per_block(int N, int elem)
{
    const int pts = elem == 8 ? PassShape<double2>::block_points : PassShape<float2>::block_points;
    int c = 4;
    while (2 * c * N <= pts) c *= 2;
    return c;
}

template <typename C2>
static hipError_t launch_fft(const double2 *planes, int NP, C2 *cache, __bf16 *split,
                             const FieldTask *tasks, const FftBlocks &fb, const LevelDesc *d_lv,
                             const LevelDesc *h_lv, int nlev, const double2 *tw, int *nzflag,
                             const BatchStrides &bs, int nf, hipStream_t s)
{
    if (nf < 1 || nf > 65535) return hipErrorInvalidValue;
    size_t lds_a = 0;
    int ntw = 0, maxN = 0;   // twiddles of all levels (the table is the levels' back to back)
    for (int l = 0; l < nlev; ++l) {
        const LevelDesc &L = h_lv[l];
        const int N = L.fd.N;
        if (L.stride < (long long)N * N || L.stride % 8 != 0 || L.cpb < 4)
            return hipErrorInvalidValue;
        // N < 1024: the twiddle index products (ja mod N) * j stay below 2^20 (udiv).
        if (L.cpb != (1 << L.cpb_log2) || (L.cpb * N) % 8 != 0 || N >= 1024 ||
            L.cpb * N > kPassMaxPoints)
            return hipErrorInvalidValue;
        lds_a = std::max(lds_a, 2 * (size_t)L.cpb * N * sizeof(C2));
        ntw = std::max(ntw, L.tw_off + N);
        maxN = std::max(maxN, N);
    }
    if (fb.n <= 0) return hipSuccess;
    const size_t lds = (size_t)(((ntw + 1) & ~1) + 2 * PassShape<C2>::unit_cap) * sizeof(C2) +
                       (size_t)maxN * sizeof(int4) +
                       (size_t)((maxN + 4) & ~3) * sizeof(int) + lds_a;
    auto k = k_ringpass<C2>;
    (void)hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds);
    // One round of resident workgroups, shared by the fields of the batch.
    int dev = 0, ncu = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, PassShape<C2>::threads, lds) !=
            hipSuccess)
        return hipErrorInvalidValue;
    per_cu = std::min(per_cu, lds_resident((long long)lds));
    const long long round = (long long)std::max(1, ncu) * std::max(1, per_cu);
    const int gx = (int)std::max(1LL, std::min<long long>(fb.n, round / nf));
    hipLaunchKernelGGL(k, dim3(gx, nf), dim3(PassShape<C2>::threads), lds, s, planes, NP, cache, split,
                       tasks, fb.units, fb.n, d_lv, tw, ntw, maxN, nzflag, bs);
    return hipGetLastError();
}

// Units of the ring pass: each task's column blocks in runs of about
// kUnitPoints points (a unit stages its field's modes once), largest first.
constexpr int kUnitPoints = 8192;

void fft_blocks(const FieldTask *tasks, int ntasks, const LevelDesc *h_lv,
                std::vector<int2> &col)
{
    col.clear();
    std::vector<std::pair<long long, int2>> units;
    for (int t = 0; t < ntasks; ++t) {
        const LevelDesc &L = h_lv[tasks[t].level];
        const int N = L.fd.N, nblocks = (N + L.cpb - 1) / L.cpb;
        const int per = std::max(1, kUnitPoints / (L.cpb * N));
        for (int x = 0; x < nblocks; x += per) {
            const int e = std::min(nblocks, x + per);
            units.push_back({(long long)(e - x) * L.cpb * N, make_int2(t, x | (e << 16))});
        }
    }
    std::stable_sort(units.begin(), units.end(),
                     [](const std::pair<long long, int2> &u, const std::pair<long long, int2> &v) {
                         return u.first > v.first;
                     });
    for (const auto &u : units) col.push_back(u.second);
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
