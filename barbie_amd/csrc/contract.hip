// Triple-product contraction on the matrix cores (MFMA).
//
// For a contraction (p1,r1,p2,r2,p3,r3) the reference computes
//     S = sum_x Re[ D_{p1,r1}(x) * D~_{p2,r2}(x) * D~_{p3,r3}(x) ]
// (barbie.c:326-337; D~ = conj(D_{|p|,r}) for a negative plane).  Here a
// plane group's contractions become 16x16 GEMM tiles with K = points:
//     S[m][n] = sum_x Pr_m(x) Cr_n(x) + Pi_m(x) (-Ci_n(x)),
// rows m = pair fields P = D1 * D~2 (formed in registers from LDS), columns
// n = 16 consecutive r3 rings C = D~3.  The real and imaginary halves are two
// MFMA chains into one accumulator, so every lane forms one pair product per
// point and nothing is computed twice.
//
// Arithmetic by precision:
//  - F64:    v_mfma_f64_16x16x4_f64, FP64 fields, FP64 accumulation;
//  - F32:    v_mfma_f32_16x16x4_f32, FP32 fields, FP32 accumulation within an
//            LDS chunk, FP64 across chunks;
//  - BF16X3: v_mfma_f32_16x16x32_bf16 on hi/lo bf16 splits of the FP32 pair
//            and column values (hi*hi + hi*lo + lo*hi), FP32 accumulation
//            within a chunk, FP64 across chunks.
//
// Lane layouts (checked on gfx950): for the K=4 forms lane l holds
// A[l%16][l/16] and B[l/16][l%16]; for 16x16x32 bf16 lane l holds
// A[l%16][8(l/16)..8(l/16)+7] and B[8(l/16)..+7][l%16].  Output lane l,
// register j is D[4j + l/16][l%16] for f64 and D[4(l/16) + j][l%16] for the
// f32-accumulating forms.
#include <hip/hip_runtime.h>

#include "device.hpp"

namespace bb {

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kThreads = 256;
constexpr int kMTW = kMaxMT / 4;   // M-tiles per wave

// LDS row pitch (elements) for a chunk: padded so that 16 lanes reading 16
// different fields at the same point hit distinct banks.
__host__ __device__ constexpr int pitch_k4_f64(int chunk) { return chunk + 1; }
__host__ __device__ constexpr int pitch_f32(int chunk) { return chunk + 2; }

// Blocks are dealt round-robin to the 8 XCDs; remap so each XCD works on a
// contiguous range of work items (neighbouring plane groups share planes, so
// their fields stay in that XCD's L2).  grid = 8 * per.
__device__ __forceinline__ int xcd_remap(int b, int per) { return (b & 7) * per + (b >> 3); }

template <typename C2>
__device__ __forceinline__ void stage_chunk(C2 *st, int pitch, const C2 *__restrict__ cache,
                                            long long npts, const WorkItem &it, int nft,
                                            long long x0, int chunk)
{
    const int nf0 = it.nf[0], nf01 = it.nf[0] + it.nf[1];
    for (int idx = threadIdx.x; idx < nft * chunk; idx += kThreads) {
        const int f = idx / chunk, xi = idx - f * chunk;
        const int fid = f < nf0 ? it.fbase[0] + f
                      : f < nf01 ? it.fbase[1] + (f - nf0) : it.fbase[2] + (f - nf01);
        C2 v;
        v.x = 0;
        v.y = 0;
        if (x0 + xi < npts) v = cache[(long long)fid * npts + x0 + xi];
        st[f * pitch + xi] = v;
    }
}

// ------------------------------------------------------------------ K = 4 --
template <typename R>
struct K4;
template <>
struct K4<double> {
    using C2 = double2;
    using Acc = f64x4;
    static __device__ __forceinline__ Acc mfma(double a, double b, Acc c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int j) { return 4 * j + (lane >> 4); }
    static constexpr int pitch(int chunk) { return pitch_k4_f64(chunk); }
};
template <>
struct K4<float> {
    using C2 = float2;
    using Acc = f32x4;
    static __device__ __forceinline__ Acc mfma(float a, float b, Acc c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int j) { return 4 * (lane >> 4) + j; }
    static constexpr int pitch(int chunk) { return pitch_f32(chunk); }
};

template <typename R>
__global__ __launch_bounds__(kThreads) void k_contract_k4(
    const typename K4<R>::C2 *__restrict__ cache, long long npts,
    const WorkItem *__restrict__ items, int nitems, const MTile *__restrict__ mtiles, int nsplit,
    int chunk, int per, double *__restrict__ out, long long split_stride, double scale)
{
    using T = K4<R>;
    using C2 = typename T::C2;
    using Acc = typename T::Acc;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    C2 *st = reinterpret_cast<C2 *>(lds_raw);

    const int logical = xcd_remap(blockIdx.x, per);
    if (logical >= nitems * nsplit) return;
    const int item_id = logical / nsplit, split = logical - item_id * nsplit;
    const WorkItem it = items[item_id];
    const int nf0 = it.nf[0], nf01 = it.nf[0] + it.nf[1];
    const int nft = nf01 + it.nf[2];
    const int pitch = T::pitch(chunk);
    const long long nchunks = (npts + chunk - 1) / chunk;
    const long long cb = nchunks * split / nsplit, ce = nchunks * (split + 1) / nsplit;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 15, kq = lane >> 4;
    const bool c2 = it.conj2 != 0, c3 = it.conj3 != 0;

    bool has[kMTW];
    int a1[kMTW], a2[kMTW], w0[kMTW][kMaxWin], wn[kMTW][kMaxWin], tw[kMTW];
#pragma unroll
    for (int t = 0; t < kMTW; ++t) {
        const int idx = wave + 4 * t;
        has[t] = idx < it.n_mt;
        a1[t] = -1;
        a2[t] = 0;
        tw[t] = 0;
#pragma unroll
        for (int v = 0; v < kMaxWin; ++v) { w0[t][v] = 0; wn[t][v] = 0; }
        if (has[t]) {
            const MTile &M = mtiles[it.mt0 + idx];
            a1[t] = M.a1[m];
            a2[t] = M.a2[m];
            tw[t] = M.tw;
#pragma unroll
            for (int v = 0; v < kMaxWin; ++v) { w0[t][v] = M.w0[v]; wn[t][v] = M.wn[v]; }
        }
    }

    double accd[kMTW][kMaxWin][4];
#pragma unroll
    for (int t = 0; t < kMTW; ++t)
#pragma unroll
        for (int v = 0; v < kMaxWin; ++v)
#pragma unroll
            for (int j = 0; j < 4; ++j) accd[t][v][j] = 0.0;

    for (long long c = cb; c < ce; ++c) {
        const long long x0 = c * chunk;
        __syncthreads();
        stage_chunk<C2>(st, pitch, cache, npts, it, nft, x0, chunk);
        __syncthreads();
#pragma unroll
        for (int t = 0; t < kMTW; ++t) {
            if (!has[t]) continue;
            Acc acc[kMaxWin];
#pragma unroll
            for (int v = 0; v < kMaxWin; ++v) acc[v] = Acc{0, 0, 0, 0};
            const C2 *r1p = st + (a1[t] < 0 ? 0 : a1[t]) * pitch;
            const C2 *r2p = st + (nf0 + a2[t]) * pitch;
            const R live = a1[t] < 0 ? R(0) : R(1);
            for (int s = 0; s < chunk; s += 4) {
                const int pt = s + kq;
                const C2 d1 = r1p[pt], d2 = r2p[pt];
                R pr, pi;
                if (c2) {
                    pr = d1.x * d2.x + d1.y * d2.y;
                    pi = d1.y * d2.x - d1.x * d2.y;
                } else {
                    pr = d1.x * d2.x - d1.y * d2.y;
                    pi = d1.x * d2.y + d1.y * d2.x;
                }
                pr *= live;
                pi *= live;
#pragma unroll
                for (int v = 0; v < kMaxWin; ++v) {
                    if (wn[t][v] == 0) continue;
                    C2 d3;
                    d3.x = 0;
                    d3.y = 0;
                    if (m < wn[t][v]) d3 = st[(nf01 + w0[t][v] + m) * pitch + pt];
                    const R nci = c3 ? d3.y : -d3.y;   // -Im(D~3)
                    acc[v] = T::mfma(pr, d3.x, acc[v]);
                    acc[v] = T::mfma(pi, nci, acc[v]);
                }
            }
#pragma unroll
            for (int v = 0; v < kMaxWin; ++v)
#pragma unroll
                for (int j = 0; j < 4; ++j) accd[t][v][j] += (double)acc[v][j];
        }
    }

    double *o = out + (long long)split * split_stride;
#pragma unroll
    for (int t = 0; t < kMTW; ++t) {
        if (!has[t]) continue;
#pragma unroll
        for (int v = 0; v < kMaxWin; ++v) {
            if (wn[t][v] == 0) continue;
            double *ot = o + (long long)(tw[t] + v) * 256;
#pragma unroll
            for (int j = 0; j < 4; ++j) ot[T::row(lane, j) * 16 + m] = accd[t][v][j] * scale;
        }
    }
}

// ------------------------------------------------------------- BF16 x 3 --
__device__ __forceinline__ void split8(const float *x, bf16x8 &hi, bf16x8 &lo)
{
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const __bf16 h = (__bf16)x[e];
        hi[e] = h;
        lo[e] = (__bf16)(x[e] - (float)h);
    }
}

__device__ __forceinline__ f32x4 mfma_bf16(const bf16x8 &a, const bf16x8 &b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__global__ __launch_bounds__(kThreads) void k_contract_bf16x3(
    const float2 *__restrict__ cache, long long npts, const WorkItem *__restrict__ items,
    int nitems, const MTile *__restrict__ mtiles, int nsplit, int chunk, int per,
    double *__restrict__ out, long long split_stride, double scale)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    float2 *st = reinterpret_cast<float2 *>(lds_raw);

    const int logical = xcd_remap(blockIdx.x, per);
    if (logical >= nitems * nsplit) return;
    const int item_id = logical / nsplit, split = logical - item_id * nsplit;
    const WorkItem it = items[item_id];
    const int nf0 = it.nf[0], nf01 = it.nf[0] + it.nf[1];
    const int nft = nf01 + it.nf[2];
    const int pitch = pitch_f32(chunk);
    const long long nchunks = (npts + chunk - 1) / chunk;
    const long long cb = nchunks * split / nsplit, ce = nchunks * (split + 1) / nsplit;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 15, kq = lane >> 4;
    const bool c2 = it.conj2 != 0, c3 = it.conj3 != 0;

    bool has[kMTW];
    int a1[kMTW], a2[kMTW], w0[kMTW][kMaxWin], wn[kMTW][kMaxWin], tw[kMTW];
#pragma unroll
    for (int t = 0; t < kMTW; ++t) {
        const int idx = wave + 4 * t;
        has[t] = idx < it.n_mt;
        a1[t] = -1;
        a2[t] = 0;
        tw[t] = 0;
#pragma unroll
        for (int v = 0; v < kMaxWin; ++v) { w0[t][v] = 0; wn[t][v] = 0; }
        if (has[t]) {
            const MTile &M = mtiles[it.mt0 + idx];
            a1[t] = M.a1[m];
            a2[t] = M.a2[m];
            tw[t] = M.tw;
#pragma unroll
            for (int v = 0; v < kMaxWin; ++v) { w0[t][v] = M.w0[v]; wn[t][v] = M.wn[v]; }
        }
    }

    double accd[kMTW][kMaxWin][4];
#pragma unroll
    for (int t = 0; t < kMTW; ++t)
#pragma unroll
        for (int v = 0; v < kMaxWin; ++v)
#pragma unroll
            for (int j = 0; j < 4; ++j) accd[t][v][j] = 0.0;

    for (long long c = cb; c < ce; ++c) {
        const long long x0 = c * chunk;
        __syncthreads();
        stage_chunk<float2>(st, pitch, cache, npts, it, nft, x0, chunk);
        __syncthreads();
#pragma unroll
        for (int t = 0; t < kMTW; ++t) {
            if (!has[t]) continue;
            f32x4 acc[kMaxWin];
#pragma unroll
            for (int v = 0; v < kMaxWin; ++v) acc[v] = f32x4{0, 0, 0, 0};
            const float2 *r1p = st + (a1[t] < 0 ? 0 : a1[t]) * pitch;
            const float2 *r2p = st + (nf0 + a2[t]) * pitch;
            const float live = a1[t] < 0 ? 0.f : 1.f;
            for (int s = 0; s < chunk; s += 32) {
                const int p0 = s + 8 * kq;   // this lane's 8 points
                float pr[8], pi[8];
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const float2 d1 = r1p[p0 + e], d2 = r2p[p0 + e];
                    if (c2) {
                        pr[e] = d1.x * d2.x + d1.y * d2.y;
                        pi[e] = d1.y * d2.x - d1.x * d2.y;
                    } else {
                        pr[e] = d1.x * d2.x - d1.y * d2.y;
                        pi[e] = d1.x * d2.y + d1.y * d2.x;
                    }
                    pr[e] *= live;
                    pi[e] *= live;
                }
                bf16x8 arh, arl, aih, ail;
                split8(pr, arh, arl);
                split8(pi, aih, ail);
#pragma unroll
                for (int v = 0; v < kMaxWin; ++v) {
                    if (wn[t][v] == 0) continue;
                    float cr[8], nci[8];
                    const bool ok = m < wn[t][v];
                    const float2 *r3p = st + (nf01 + w0[t][v] + (ok ? m : 0)) * pitch;
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        float2 d3 = r3p[p0 + e];
                        if (!ok) { d3.x = 0.f; d3.y = 0.f; }
                        cr[e] = d3.x;
                        nci[e] = c3 ? d3.y : -d3.y;
                    }
                    bf16x8 brh, brl, bih, bil;
                    split8(cr, brh, brl);
                    split8(nci, bih, bil);
                    acc[v] = mfma_bf16(arh, brh, acc[v]);
                    acc[v] = mfma_bf16(arh, brl, acc[v]);
                    acc[v] = mfma_bf16(arl, brh, acc[v]);
                    acc[v] = mfma_bf16(aih, bih, acc[v]);
                    acc[v] = mfma_bf16(aih, bil, acc[v]);
                    acc[v] = mfma_bf16(ail, bih, acc[v]);
                }
            }
#pragma unroll
            for (int v = 0; v < kMaxWin; ++v)
#pragma unroll
                for (int j = 0; j < 4; ++j) accd[t][v][j] += (double)acc[v][j];
        }
    }

    double *o = out + (long long)split * split_stride;
#pragma unroll
    for (int t = 0; t < kMTW; ++t) {
        if (!has[t]) continue;
#pragma unroll
        for (int v = 0; v < kMaxWin; ++v) {
            if (wn[t][v] == 0) continue;
            double *ot = o + (long long)(tw[t] + v) * 256;
#pragma unroll
            for (int j = 0; j < 4; ++j) ot[(4 * kq + j) * 16 + m] = accd[t][v][j] * scale;
        }
    }
}

// ----------------------------------------------------------------- host --
static size_t elem_bytes(CPrec p) { return p == CPrec::F64 ? sizeof(double2) : sizeof(float2); }

static int pitch_of(CPrec p, int chunk)
{
    return p == CPrec::F64 ? pitch_k4_f64(chunk) : pitch_f32(chunk);
}

static size_t lds_bytes(CPrec p, int chunk, int max_fields)
{
    return (size_t)max_fields * pitch_of(p, chunk) * elem_bytes(p);
}

constexpr size_t kLdsBudget = 64 * 1024;

int contract_chunk(CPrec p, int max_fields)
{
    const int step = (p == CPrec::BF16X3) ? 32 : 4;
    int chunk = (p == CPrec::F64) ? 32 : 64;
    while (chunk > step && lds_bytes(p, chunk, max_fields) > kLdsBudget) chunk /= 2;
    return chunk;
}

hipError_t launch_contract(CPrec p, const void *cache, long long npts, const WorkItem *items,
                           int nitems, const MTile *mtiles, int nsplit, int chunk, int max_fields,
                           double *out, long long split_stride, double scale, hipStream_t s)
{
    if (nitems <= 0) return hipSuccess;
    const long long total = (long long)nitems * nsplit;
    const int per = (int)((total + 7) / 8);
    dim3 grid(8 * per);
    const size_t lds = lds_bytes(p, chunk, max_fields);
    if (p == CPrec::F64) {
        (void)hipFuncSetAttribute((const void *)k_contract_k4<double>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        hipLaunchKernelGGL(k_contract_k4<double>, grid, dim3(kThreads), lds, s,
                           (const double2 *)cache, npts, items, nitems, mtiles, nsplit, chunk,
                           per, out, split_stride, scale);
    } else if (p == CPrec::F32) {
        (void)hipFuncSetAttribute((const void *)k_contract_k4<float>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        hipLaunchKernelGGL(k_contract_k4<float>, grid, dim3(kThreads), lds, s,
                           (const float2 *)cache, npts, items, nitems, mtiles, nsplit, chunk,
                           per, out, split_stride, scale);
    } else {
        (void)hipFuncSetAttribute((const void *)k_contract_bf16x3,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        hipLaunchKernelGGL(k_contract_bf16x3, grid, dim3(kThreads), lds, s,
                           (const float2 *)cache, npts, items, nitems, mtiles, nsplit, chunk, per,
                           out, split_stride, scale);
    }
    return hipGetLastError();
}

}  // namespace bb
