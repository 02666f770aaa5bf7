// Triple-product contraction on the matrix cores (MFMA).
//
// For a contraction (p1,r1,p2,r2,p3,r3) the reference computes
//     S = sum_x Re[ D_{p1,r1}(x) * D~_{p2,r2}(x) * D~_{p3,r3}(x) ]
// (barbie.c:326-337; D~ = conj(D_{|p|,r}) for a negative plane).  Here the
// contractions of a plane group become 16x16 GEMM tiles with K = points:
//     S[m][n] = sum_x Pr_m(x) Cr_n(x) + Ai_m(x) Ci_n(x),
// rows m = pair fields P = D1 * D~2 (formed in registers from LDS),
// Ai = -Im(P) (or +Im(P) when the third plane is negative, since then
// -Im(D~3) = +Ci), columns n = 16 consecutive r3 rings C = D_{|p3|}.  The real
// and imaginary halves are two MFMA chains into one accumulator, so each lane
// forms one pair product per point and nothing is computed twice.  The two
// conjugations are sign-bit flips.
//
// Arithmetic by precision:
//  - F64:    v_mfma_f64_16x16x4_f64, FP64 fields, FP64 accumulation;
//  - F32:    v_mfma_f32_16x16x4_f32, FP32 fields, FP32 accumulation over
//            kFlush chunks, FP64 across;
//  - BF16X3: v_mfma_f32_16x16x32_bf16 on hi/lo bf16 splits (hi*hi + hi*lo +
//            lo*hi) of the FP32 pair values and of the r3 fields (the latter
//            split once, by the FFT pass), FP32 accumulation over kFlush
//            chunks, FP64 across.
// Rows of a tile whose pair slot is empty and columns past the window's end
// read arbitrary staged rings in the FP64 and FP32 kernels and a ring of
// zeros in the BF16X3 kernel; MFMA keeps rows and columns independent, and
// the outputs of those slots are never read.
//
// Lane layouts (checked on gfx950): for the K=4 forms lane l holds
// A[l%16][l/16] and B[l/16][l%16]; for 16x16x32 bf16 lane l holds
// A[l%16][8(l/16)..8(l/16)+7] and B[8(l/16)..+7][l%16].  Output lane l,
// register j is D[4j + l/16][l%16] for f64 and D[4(l/16) + j][l%16] for the
// f32-accumulating forms.
//
// Staging is software-pipelined.  FP64 and FP32 kernels: each thread holds
// its share of the next chunk in registers while the waves work on the
// current one from LDS; every staged ring is 256 bytes of one chunk (plus 16
// bytes of padding, so that 16 lanes reading 16 different rings at one offset
// hit distinct banks).  BF16X3 kernel: LDS-direct loads fill a second buffer
// while the waves work on the first (DmaStager, dma_off).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device.hpp"

namespace bb {

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kTableBytes = 2 * kRowCap * 4;   // row -> cache slot, row -> split flag
constexpr int kRowBytes = 256;                 // bytes of one ring per chunk
constexpr int kPitch = kRowBytes + 16;         // LDS bytes per staged ring
constexpr int kFlush = 4;                      // chunks between FP32 -> FP64 flushes
// 8-byte staging loads per thread: 30 rings per wave, 32 units per ring.
constexpr int kPF = 15;

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Fills the row tables of an item (row -> cache slot, row -> split flag).
template <int NT>
__device__ __forceinline__ void row_tables(const WorkItem &it, int *row_slot, int *row_split)
{
    for (int r = threadIdx.x; r < it.nrows; r += NT) {
        const int v = it.rslot[r];
        row_slot[r] = v & (kRowSplit - 1);
        row_split[r] = (v & kRowSplit) ? 1 : 0;
    }
}

// Staging of one chunk: the item's rings, each kRowBytes contiguous bytes of
// its field (raw values, or the bf16 split operands for split rows), moved as
// 8-byte units.  Branch-free: every thread moves kPF units; units past the
// item's rings read ring 0's place again and land in a spare LDS ring (row
// kRowCap).  Field strides are multiples of a chunk, so a chunk never runs
// past its field.  Each unit's source, as an offset in 8-byte words from the
// item's cache base with the split flag in bit 31, is computed once per
// workgroup (init); a chunk then costs a select and a shift-add per unit.
template <int NT>
struct Stager {
    const char *cache;
    const char *split;
    int nunits;              // 8-byte units this chunk: nrows * kRowBytes / 8
    unsigned uo[kPF];
    uint2 pf[kPF];

    // field_words: 8-byte words between fields (fits 31 bits with the slots;
    // checked on the host).
    __device__ __forceinline__ void init(const int *row_slot, const int *row_split, int nrows,
                                         long long field_words)
    {
#pragma unroll
        for (int e = 0; e < kPF; ++e) {
            const int u = threadIdx.x + e * NT;
            const int r = min(u >> 5, nrows - 1), k = u & 31;   // 32 units per ring
            uo[e] = (unsigned)(row_slot[r] * field_words + k) | (row_split[r] ? 0x80000000u : 0u);
        }
    }
    __device__ __forceinline__ void load(long long chunk_off)
    {
        const char *bc = cache + chunk_off, *bs = split + chunk_off;
#pragma unroll
        for (int e = 0; e < kPF; ++e) {
            const char *b = (uo[e] & 0x80000000u) ? bs : bc;
            pf[e] = *reinterpret_cast<const uint2 *>(b + (size_t)(uo[e] & 0x7fffffffu) * 8);
        }
    }
    __device__ __forceinline__ void store(unsigned char *rows)
    {
#pragma unroll
        for (int e = 0; e < kPF; ++e) {
            const int u = threadIdx.x + e * NT;
            const int r = u < nunits ? (u >> 5) : kRowCap, k = u & 31;
            *reinterpret_cast<uint2 *>(rows + r * kPitch + k * 8) = pf[e];
        }
    }
};

// Direct-to-LDS staging (global_load_lds_dwordx4), double-buffered, for the
// BF16x3 kernel.  A chunk buffer holds the item's rings in groups of 16;
// 16-byte piece k (bytes 16k..16k+15 of the ring's 256-byte chunk) of ring r
// sits at
//     ((r / 16) * 4 + k / 4) * 1024 + (k % 4) * 256 + (r % 16) * 16,
// so one load instruction fills the 1 KiB block of a (group, quarter) pair:
// lane l brings piece 4q + l/16 of ring 16 G + l%16.  Each lane's 4 pieces of
// a quarter are 256 bytes apart (immediate offsets), and the 16 bytes at
// 16 g mod 256 lie in bank group g (4 of the 64 banks).  Rings past the
// item's last are not loaded.
//
// ds_read_b128 serves a wave in four phases of 16 lanes, and on gfx950 a
// phase holds lanes 0-3, 12-15, 20-23, 24-27 (or 4-11, 16-19, 28-31, and the
// same sets + 32; measured with the bank-conflict counter).  A tile operand
// lane (m, kq) reads quarter kq of row m's ring, so a phase holds all 16 rows
// m, half of them in one quarter and half in the next.  16 consecutive rings
// are then conflict-free, but two lanes of one phase that read the same
// ring in different quarters hit one bank group, as do two rings of one
// bank group.  Empty rows and the columns past a window's rings therefore
// read a 1 KiB block of zeros after the ring groups (cleared once, never
// loaded) at bank groups that do not collide: a column continues its
// window's sequence of bank groups, and every empty row of a tile reads one
// address, in bank group nrows mod 16, which the row colouring keeps free of
// the tile's rings.
constexpr int kDmaGroups = 2;   // 16-ring groups per wave (rowcap / 16 / waves, rounded up)

// Staged rings per item: 120 (4 waves), 240 (8 waves), 96 (4 waves sized
// for 3 workgroups per CU, kTWWOcc3 tile windows).
__host__ __device__ constexpr int dma_rowcap(int nw, int tww)
{
    return tww == kTWWOcc3 ? rowcap_occ3(nw) : rowcap_for(nw);
}
// The zero block, after the item's ring groups, and the bytes between the
// two chunk buffers.
__host__ __device__ constexpr int dma_zero_off(int nw, int tww)
{
    return (dma_rowcap(nw, tww) + 15) / 16 * 4096;
}
__host__ __device__ constexpr int dma_buf_stride(int nw, int tww)
{
    return dma_zero_off(nw, tww) + 1024;
}
__host__ __device__ constexpr int dma_lds_bytes(int nw, int tww)
{
    return 2 * dma_buf_stride(nw, tww) + kTableBytes;
}

static_assert((rowcap_for(4) + 15) / 16 <= kDmaGroups * 4 &&
                  (rowcap_for(8) + 15) / 16 <= kDmaGroups * 8,
              "every 16-ring group of an item has a wave");
static_assert(lds_resident(dma_lds_bytes(4, kTWWOcc3)) >= 3, "3 workgroups' LDS per CU");
static_assert(lds_resident(dma_lds_bytes(4, kTWW32)) >= 2 && lds_resident(dma_lds_bytes(8, kTWW32)) >= 1,
              "2 (4 waves) or 1 (8 waves) workgroups' LDS per CU");

__device__ __forceinline__ int dma_off(int r, int k)
{
    return ((r >> 4) * 4 + (k >> 2)) * 1024 + (k & 3) * 256 + (r & 15) * 16;
}

typedef __attribute__((address_space(3))) void lds_void;

template <int NW>
struct DmaStager {
    // This lane's source per group of the wave, for the chunk of the last
    // issue (advanced only after the wait for the loads that read it).
    const char *src[kDmaGroups];
    int ngroups;   // 16-ring groups of the item
    int live;      // bit e: this lane's ring of group e is one of the item's

    // Groups G = wave, wave + NW, ...: lane l serves ring 16 G + l%16 and
    // pieces 4q + l/16.
    __device__ __forceinline__ void init(const char *cache, const char *split, const int *row_slot,
                                         const int *row_split, int nrows, long long fstride,
                                         long long chunk_off, int wave, int lane)
    {
        ngroups = (nrows + 15) >> 4;
        live = 0;
#pragma unroll
        for (int e = 0; e < kDmaGroups; ++e) {
            const int G = wave + e * NW;
            if (16 * G + (lane & 15) < nrows) live |= 1 << e;
            const int r = min(16 * G + (lane & 15), nrows - 1);
            src[e] = (row_split[r] ? split : cache) + (long long)row_slot[r] * fstride * 8 +
                     chunk_off + (lane >> 4) * 16;
        }
    }
    __device__ __forceinline__ void advance()
    {
#pragma unroll
        for (int e = 0; e < kDmaGroups; ++e) src[e] += kRowBytes;
    }
    // Issues the loads of the current chunk into buffer `buf`.
    __device__ __forceinline__ void issue(unsigned char *buf, int wave)
    {
#pragma unroll
        for (int e = 0; e < kDmaGroups; ++e) {
            const int G = wave + e * NW;
            if (G < ngroups && (live >> e & 1)) {   // lanes past the item's rings stay idle
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    __builtin_amdgcn_global_load_lds((const void *)(src[e] + 64 * q),
                                                     (lds_void *)(buf + (G * 4 + q) * 1024), 16, 0,
                                                     0);
            }
        }
    }
};

// Waits for this wave's outstanding global loads (the LDS-direct ones
// included); the other counters are left alone.
__device__ __forceinline__ void wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Flips the sign of x when the mask's top bit is set.
__device__ __forceinline__ double flip(double x, unsigned mask)
{
    unsigned long long b = __double_as_longlong(x) ^ ((unsigned long long)mask << 32);
    return __longlong_as_double(b);
}
__device__ __forceinline__ float flip(float x, unsigned mask)
{
    return __uint_as_float(__float_as_uint(x) ^ mask);
}

// ------------------------------------------------------------------ K = 4 --
template <typename R>
struct K4;
template <>
struct K4<double> {
    using C2 = double2;
    using Acc = f64x4;
    static constexpr int TWW = kTWW;
    static __device__ __forceinline__ Acc mfma(double a, double b, Acc c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int j) { return 4 * j + (lane >> 4); }
    static constexpr bool kFP64Acc = true;
};
template <>
struct K4<float> {
    using C2 = float2;
    using Acc = f32x4;
    static constexpr int TWW = kTWWF32;
    static __device__ __forceinline__ Acc mfma(float a, float b, Acc c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int j) { return 4 * (lane >> 4) + j; }
    static constexpr bool kFP64Acc = false;
};

// TWW: tile windows per wave; MINB: workgroups per CU the register budget
// is sized for.
template <typename R, int NW, int TWW, int MINB>
__global__ __launch_bounds__(64 * NW, MINB) void k_contract_k4(
    const typename K4<R>::C2 *__restrict__ cache, const WorkItem *__restrict__ items,
    const WorkUnit *__restrict__ units, const TWin *__restrict__ tws, double *__restrict__ out,
    BatchStrides bs)
{
    cache += blockIdx.y * bs.cache;   // field of the batch
    out += blockIdx.y * bs.part;
    using T = K4<R>;
    using C2 = typename T::C2;
    using Acc = typename T::Acc;
    constexpr int CH = kRowBytes / (int)sizeof(C2);   // points per chunk
    extern __shared__ __align__(16) unsigned char lds[];
    int *row_slot = reinterpret_cast<int *>(lds);
    int *row_split = row_slot + kRowCap;
    unsigned char *rows = lds + kTableBytes;

    const WorkUnit U = units[blockIdx.x];
    const WorkItem &it = items[U.item];
    const int nrows = uni(it.nrows);
    row_tables<64 * NW>(it, row_slot, row_split);
    const long long cb = U.cb, ce = U.ce;
    const int lane = threadIdx.x & 63, wave = uni(threadIdx.x >> 6);
    const int m = lane & 15, kq = lane >> 4;

    // Per tile window slot (every slot is filled; unused ones are dummies
    // whose output index is -1): lane byte offsets of its A rows and B row,
    // and flags.
    const TWin *TW = tws + it.tw0 + wave * kTWW;
    int oa1[TWW], oa2[TWW], ob[TWW], fl[TWW];
#pragma unroll
    for (int j = 0; j < TWW; ++j) {
        oa1[j] = max((int)TW[j].a1[m], 0) * kPitch + kq * (int)sizeof(C2);
        oa2[j] = (int)TW[j].a2[m] * kPitch + kq * (int)sizeof(C2);
        ob[j] = min((int)TW[j].b0 + m, nrows - 1) * kPitch + kq * (int)sizeof(C2);
        fl[j] = uni(TW[j].flags);
    }

    Acc acc[TWW];
    double accd[TWW][4];
#pragma unroll
    for (int j = 0; j < TWW; ++j) {
        acc[j] = Acc{0, 0, 0, 0};
#pragma unroll
        for (int q = 0; q < 4; ++q) accd[j][q] = 0.0;
    }
    Stager<64 * NW> S;
    S.cache = reinterpret_cast<const char *>(cache + it.off);
    S.split = S.cache;
    S.nunits = nrows * (kRowBytes / 8);
    __syncthreads();
    S.init(row_slot, row_split, nrows, it.fstride * (long long)sizeof(C2) / 8);

    if (cb < ce) S.load(cb * kRowBytes);
    int since = 0;
    for (long long c = cb; c < ce; ++c) {
        __syncthreads();
        S.store(rows);
        __syncthreads();
        if (c + 1 < ce) S.load((c + 1) * kRowBytes);
#pragma unroll
        for (int s = 0; s < CH / 4; ++s) {   // 4 points per MFMA pair
            const int so = s * 4 * (int)sizeof(C2);
            R pr = 0, ai = 0;
            C2 bcur = *reinterpret_cast<const C2 *>(rows + ob[0] + so);
#pragma unroll
            for (int j = 0; j < TWW; ++j) {
                const bool newa = (fl[j] & TW_NEWA) != 0;
                C2 d1, d2;
                d1.x = d1.y = d2.x = d2.y = 0;
                if (newa) {
                    d1 = *reinterpret_cast<const C2 *>(rows + oa1[j] + so);
                    d2 = *reinterpret_cast<const C2 *>(rows + oa2[j] + so);
                }
                C2 bnext = bcur;
                if (j + 1 < TWW) bnext = *reinterpret_cast<const C2 *>(rows + ob[j + 1] + so);
                if (newa) {
                    const unsigned m2 = (fl[j] & TW_CONJ2) ? 0x80000000u : 0u;
                    const unsigned m3 = (fl[j] & TW_CONJ3) ? 0u : 0x80000000u;
                    const R e2 = flip(d2.y, m2);
                    pr = d1.x * d2.x - d1.y * e2;
                    ai = flip(d1.x * e2 + d1.y * d2.x, m3);
                }
                acc[j] = T::mfma(pr, bcur.x, acc[j]);
                acc[j] = T::mfma(ai, bcur.y, acc[j]);
                bcur = bnext;
            }
        }
        if (!T::kFP64Acc && (++since == kFlush || c + 1 == ce)) {
            since = 0;
#pragma unroll
            for (int j = 0; j < TWW; ++j) {
#pragma unroll
                for (int q = 0; q < 4; ++q) accd[j][q] += (double)acc[j][q];
                acc[j] = Acc{0, 0, 0, 0};
            }
        }
    }

    const double scale = it.scale;
#pragma unroll
    for (int j = 0; j < TWW; ++j) {
        const int ot_idx = TW[j].out;
        if (ot_idx >= 0) {
            double *ot = out + ((long long)ot_idx + U.split) * 256;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const double v = T::kFP64Acc ? (double)acc[j][q] : accd[j][q];
                ot[T::row(lane, q) * 16 + m] = v * scale;
            }
        }
    }
}

// ------------------------------------------------------------- BF16 x 3 --
__device__ __forceinline__ f32x4 mfma_bf16(const bf16x8 &a, const bf16x8 &b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// The bf16 pairs (-1, 0) and (0, -1) in registers.  They are written by
// v_mov_b32 with 32-bit literals: this compiler folds the constant (-1, 0)
// into the inline constant -1.0, which v_dot2c_f32_bf16 reads as the FP32
// bit pattern 0xbf800000, i.e. the pair (0, -1).
struct NegUnit {
    bf16x2 m0, m1;
    __device__ __forceinline__ NegUnit()
    {
        unsigned u0, u1;
        asm volatile("v_mov_b32 %0, 0xbf80" : "=v"(u0));
        asm volatile("v_mov_b32 %0, 0xbf800000" : "=v"(u1));
        m0 = __builtin_bit_cast(bf16x2, u0);
        m1 = __builtin_bit_cast(bf16x2, u1);
    }
};

// hi = bf16(x) (round to nearest even), lo = bf16(x - hi), eight values.
// Each pair is converted once (v_cvt_pk_bf16_f32).  The residual x - hi is
// one bf16 dot product per value, dot2(hi pair, (-1, 0)) + x (or (0, -1)):
// the product of a bf16 with -1 is exact and x - hi is representable in FP32
// (hi is within a factor 2 of x), so the sum is exact.
__device__ __forceinline__ void split8(const f32x8 &x, const NegUnit &nu, bf16x8 &hi, bf16x8 &lo)
{
    const bf16x2 m0 = nu.m0, m1 = nu.m1;
    u32x4 h, l;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const f32x2 v = {x[2 * q], x[2 * q + 1]};
        const bf16x2 hb = __builtin_convertvector(v, bf16x2);
        const f32x2 r = {__builtin_amdgcn_fdot2_f32_bf16(hb, m0, v.x, false),
                         __builtin_amdgcn_fdot2_f32_bf16(hb, m1, v.y, false)};
        h[q] = __builtin_bit_cast(unsigned, hb);
        l[q] = __builtin_bit_cast(unsigned, __builtin_convertvector(r, bf16x2));
    }
    hi = __builtin_bit_cast(bf16x8, h);
    lo = __builtin_bit_cast(bf16x8, l);
}

// Pair products of one tile's rows over the chunk's 8 points of this lane,
// split into bf16 hi/lo A operands: Pr = Re(D1 D~2), Ai = -Im(D1 D~2), or
// +Im when the third plane is negative.  C2 / C3: conjugations as template
// parameters, so that the signs become instruction modifiers.
template <bool C2, bool C3>
__device__ __forceinline__ void pair_operands(const unsigned char *rows, int oa1, int oa2,
                                              const NegUnit &nu, bf16x8 &arh, bf16x8 &arl,
                                              bf16x8 &aih, bf16x8 &ail)
{
    f32x8 pr, ai;
#pragma unroll
    for (int h = 0; h < 4; ++h) {   // two points per float4; pieces 256 bytes apart
        const float4 a = *reinterpret_cast<const float4 *>(rows + oa1 + 256 * h);
        const float4 b = *reinterpret_cast<const float4 *>(rows + oa2 + 256 * h);
        const float e0 = C2 ? -b.y : b.y, e1 = C2 ? -b.w : b.w;
        pr[2 * h] = a.x * b.x - a.y * e0;
        pr[2 * h + 1] = a.z * b.z - a.w * e1;
        const float i0 = a.x * e0 + a.y * b.x, i1 = a.z * e1 + a.w * b.z;
        ai[2 * h] = C3 ? i0 : -i0;
        ai[2 * h + 1] = C3 ? i1 : -i1;
    }
    split8(pr, nu, arh, arl);
    split8(ai, nu, aih, ail);
}

// A staged ring of 32 points: raw rows hold float2 per point; split rows
// hold, per group of 8 points, [Re hi x8][Re lo x8][Im hi x8][Im lo x8] bf16.
// One work unit (the workgroup's share of an item) of the BF16x3 kernels:
// NW waves, TWW tile windows per wave.
template <int NW, int TWW>
__device__ __forceinline__ void bf16x3_unit(const float2 *__restrict__ cache,
                                            const __bf16 *__restrict__ splitc,
                                            const WorkItem &it, const WorkUnit U,
                                            const TWin *__restrict__ tws,
                                            double *__restrict__ out, unsigned char *lds)
{
    constexpr int BUF = dma_buf_stride(NW, TWW), ZERO = dma_zero_off(NW, TWW);
    int *row_slot = reinterpret_cast<int *>(lds + 2 * BUF);
    int *row_split = row_slot + kRowCap;
    const int nrows = uni(it.nrows);
    row_tables<64 * NW>(it, row_slot, row_split);
    const long long cb = U.cb, ce = U.ce;
    const int lane = threadIdx.x & 63, wave = uni(threadIdx.x >> 6);
    const int m = lane & 15, kq = lane >> 4;

    const TWin *TW = tws + it.tw0 + wave * kTWW;
    const int cnt = uni(it.cnt[wave]);
    int oa1[TWW], oa2[TWW], ob[TWW], fl[TWW];
    const int zrow = ZERO + 16 * (nrows & 15);   // empty rows: one address of the zero block
#pragma unroll
    for (int j = 0; j < TWW; ++j) {
        // Empty rows and the columns past the window's rings read zeros:
        // their products are zeros, which cost the same instructions but
        // less power than arbitrary data.  (Reading distinct staged rings
        // there instead ran slower: the kernel is partly power-bound.)
        const bool row_ok = TW[j].a1[m] >= 0, col_ok = m < TW[j].nb;
        const int b = TW[j].b0 + m;
        oa1[j] = row_ok ? dma_off(TW[j].a1[m], 4 * kq) : zrow;
        oa2[j] = row_ok ? dma_off(TW[j].a2[m], 4 * kq) : zrow;
        ob[j] = col_ok ? dma_off(b, 4 * kq) : ZERO + 16 * (b & 15);
        fl[j] = uni(TW[j].flags);
    }

    const NegUnit nu;
    f32x4 acc[TWW];
    double accd[TWW][4];
#pragma unroll
    for (int j = 0; j < TWW; ++j) {
        acc[j] = f32x4{0, 0, 0, 0};
#pragma unroll
        for (int q = 0; q < 4; ++q) accd[j][q] = 0.0;
    }
    // The zero blocks of both buffers (no load writes them).
    if (threadIdx.x < 128)
        *reinterpret_cast<float4 *>(lds + (threadIdx.x >> 6) * BUF + ZERO + (threadIdx.x & 63) * 16) =
            make_float4(0.f, 0.f, 0.f, 0.f);
    DmaStager<NW> S;
    __syncthreads();
    S.init(reinterpret_cast<const char *>(cache + it.off),
           reinterpret_cast<const char *>(splitc + it.off * 4), row_slot, row_split, nrows,
           it.fstride, cb * kRowBytes, wave, lane);

    if (cb < ce) S.issue(lds, wave);
    int since = 0;
    for (long long c = cb; c < ce; ++c) {
        // This wave's loads of chunk c are in LDS; after the barrier, every
        // wave's are, and every wave is done with chunk c - 1, whose buffer
        // the loads of chunk c + 1 then overwrite.
        wait_vmem();
        __syncthreads();
        const int cur = (int)((c - cb) & 1);
        // rows (read) and nxt (written by the loads of chunk c + 1) are the
        // two buffers: restrict tells the compiler, so that its waits for
        // LDS-direct loads do not cover reads of the other buffer.
        auto step = [&](const unsigned char *__restrict__ rows, unsigned char *__restrict__ nxt) {
            if (c + 1 < ce) {
                S.advance();
                S.issue(nxt, wave);
            }
            bf16x8 arh = {}, arl = {}, aih = {}, ail = {};
#pragma unroll
            for (int j = 0; j < TWW; ++j) {
                if (j >= cnt) break;   // uniform: the wave's remaining slots are empty
                // B first, so its latency overlaps the pair products.
                const unsigned char *bp = rows + ob[j];
                const bf16x8 brh = *reinterpret_cast<const bf16x8 *>(bp);
                const bf16x8 brl = *reinterpret_cast<const bf16x8 *>(bp + 256);
                const bf16x8 bih = *reinterpret_cast<const bf16x8 *>(bp + 512);
                const bf16x8 bil = *reinterpret_cast<const bf16x8 *>(bp + 768);
                if (fl[j] & TW_NEWA) {
                    switch (fl[j] & (TW_CONJ2 | TW_CONJ3)) {
                    case 0: pair_operands<false, false>(rows, oa1[j], oa2[j], nu, arh, arl, aih, ail); break;
                    case TW_CONJ2: pair_operands<true, false>(rows, oa1[j], oa2[j], nu, arh, arl, aih, ail); break;
                    case TW_CONJ3: pair_operands<false, true>(rows, oa1[j], oa2[j], nu, arh, arl, aih, ail); break;
                    default: pair_operands<true, true>(rows, oa1[j], oa2[j], nu, arh, arl, aih, ail); break;
                    }
                }
                acc[j] = mfma_bf16(arh, brh, acc[j]);
                acc[j] = mfma_bf16(aih, bih, acc[j]);
                acc[j] = mfma_bf16(arh, brl, acc[j]);
                acc[j] = mfma_bf16(aih, bil, acc[j]);
                acc[j] = mfma_bf16(arl, brh, acc[j]);
                acc[j] = mfma_bf16(ail, bih, acc[j]);
            }
        };
        step(lds + cur * BUF, lds + (cur ^ 1) * BUF);
        if (++since == kFlush || c + 1 == ce) {
            since = 0;
#pragma unroll
            for (int j = 0; j < TWW; ++j) {
#pragma unroll
                for (int q = 0; q < 4; ++q) accd[j][q] += (double)acc[j][q];
                acc[j] = f32x4{0, 0, 0, 0};
            }
        }
    }

    const double scale = it.scale;
#pragma unroll
    for (int j = 0; j < TWW; ++j) {
        const int ot_idx = TW[j].out;
        if (ot_idx >= 0) {
            double *ot = out + ((long long)ot_idx + U.split) * 256;
#pragma unroll
            for (int q = 0; q < 4; ++q) ot[(4 * kq + q) * 16 + m] = accd[j][q] * scale;
        }
    }
}

// NW waves, TWW tile windows per wave, MINB workgroups per CU; one
// workgroup per unit, blockIdx.y the field of the batch.
template <int NW, int TWW, int MINB>
__global__ __launch_bounds__(64 * NW, MINB) void k_contract_bf16x3(
    const float2 *__restrict__ cache, const __bf16 *__restrict__ splitc,
    const WorkItem *__restrict__ items, const WorkUnit *__restrict__ units,
    const TWin *__restrict__ tws, double *__restrict__ out, BatchStrides bs)
{
    extern __shared__ __align__(16) unsigned char lds[];
    const WorkUnit U = units[blockIdx.x];
    bf16x3_unit<NW, TWW>(cache + blockIdx.y * bs.cache, splitc + blockIdx.y * bs.cache * 4,
                         items[U.item], U, tws, out + blockIdx.y * bs.part, lds);
}

// Persistent form over the units of several estimates (fields of different
// geometries and batches): a round of resident workgroups takes units, in
// the order given, from a queue counter until none is left;
// each unit carries its own field, work lists and output.
template <int NW, int TWW, int MINB>
__global__ __launch_bounds__(64 * NW, MINB) void k_contract_bf16x3_mixed(
    const MixUnit *__restrict__ mu, int nunits, int *__restrict__ next)
{
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ int take;
    for (;;) {
        __syncthreads();   // the previous unit is done with the LDS
        if (threadIdx.x == 0) take = atomicAdd(next, 1);
        __syncthreads();
        const int u = uni(take);   // uniform: the unit's pointers are scalar loads
        if (u >= nunits) break;
        const MixUnit M = mu[u];
        bf16x3_unit<NW, TWW>(M.cache, M.split, *M.item, M.unit, M.tws, M.out, lds);
    }
}

__global__ void k_zero_int(int *p) { *p = 0; }

// ----------------------------------------------------------------- host --
int contract_chunk(CPrec p) { return p == CPrec::F64 ? kRowBytes / 16 : kRowBytes / 8; }

template <int NW>
static hipError_t launch_nw(CPrec p, int tww, const void *cache, const void *split,
                            const WorkItem *items, const WorkUnit *units, int nunits,
                            const TWin *tws, double *out, const BatchStrides &bs, int nf,
                            hipStream_t s)
{
    dim3 grid((unsigned)nunits, (unsigned)nf);
    const size_t lds = kTableBytes + (size_t)(kRowCap + 1) * kPitch;   // + spare ring
    if (p == CPrec::F64) {
        auto k = k_contract_k4<double, NW, kTWW, 8 / NW>;
        (void)hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds);
        hipLaunchKernelGGL(k, grid, dim3(64 * NW), lds, s, (const double2 *)cache, items, units,
                           tws, out, bs);
    } else if (p == CPrec::F32) {
        auto k = k_contract_k4<float, NW, kTWWF32, 8 / NW>;
        (void)hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds);
        hipLaunchKernelGGL(k, grid, dim3(64 * NW), lds, s, (const float2 *)cache, items, units,
                           tws, out, bs);
    } else if (tww == kTWWOcc3) {
        if (NW != 4) return hipErrorInvalidValue;
        auto k = k_contract_bf16x3<4, kTWWOcc3, 3>;
        const size_t lds_d = dma_lds_bytes(4, kTWWOcc3);
        (void)hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds_d);
        hipLaunchKernelGGL(k, grid, dim3(256), lds_d, s, (const float2 *)cache,
                           (const __bf16 *)split, items, units, tws, out, bs);
    } else {
        auto k = k_contract_bf16x3<NW, kTWW32, 8 / NW>;
        const size_t lds_d = dma_lds_bytes(NW, kTWW32);
        (void)hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds_d);
        hipLaunchKernelGGL(k, grid, dim3(64 * NW), lds_d, s, (const float2 *)cache,
                           (const __bf16 *)split, items, units, tws, out, bs);
    }
    return hipGetLastError();
}

hipError_t launch_contract_mixed(int tww, const MixUnit *units, int nunits, int *counter,
                                 hipStream_t s)
{
    if (nunits <= 0) return hipSuccess;
    int ncu = 0, dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, s, counter);
    if (tww == kTWWOcc3) {
        auto k = k_contract_bf16x3_mixed<4, kTWWOcc3, 3>;
        const size_t lds_d = dma_lds_bytes(4, kTWWOcc3);
        (void)hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds_d);
        const int grid = std::min(nunits, 3 * ncu);
        hipLaunchKernelGGL(k, dim3(grid), dim3(256), lds_d, s, units, nunits, counter);
    } else {
        auto k = k_contract_bf16x3_mixed<4, kTWW32, 2>;
        const size_t lds_d = dma_lds_bytes(4, kTWW32);
        (void)hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds_d);
        const int grid = std::min(nunits, 2 * ncu);
        hipLaunchKernelGGL(k, dim3(grid), dim3(256), lds_d, s, units, nunits, counter);
    }
    return hipGetLastError();
}

hipError_t launch_contract(CPrec p, int waves, int tww, const void *cache, const void *split,
                           const WorkItem *items, const WorkUnit *units, int nunits,
                           const TWin *tws, double *out, const BatchStrides &bs, int nf,
                           hipStream_t s)
{
    if (nunits <= 0) return hipSuccess;
    if (nf < 1 || nf > 65535) return hipErrorInvalidValue;
    if (waves == 4)
        return launch_nw<4>(p, tww, cache, split, items, units, nunits, tws, out, bs, nf, s);
    return launch_nw<8>(p, tww, cache, split, items, units, nunits, tws, out, bs, nf, s);
}

}  // namespace bb
