// Device-side data structures and kernel launchers shared by the estimator
// (internal to libbarbie_amd).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>

#include "barbie.h"

namespace bb {

constexpr int kMaxStages = 16;

// Factorisation of an FFT length into radices 4, 2, 3, 5 (Stockham order).
// Exact division for 0 <= n < 2^20 and 1 <= d < 2^12 (the index ranges of
// the FFT passes): n / d = (n * ceil(2^40 / d)) >> 40.
inline unsigned long long udiv_magic(unsigned d) { return ((1ULL << 40) + d - 1) / d; }
__device__ __forceinline__ unsigned udiv(unsigned n, unsigned long long m)
{
    return (unsigned)(((unsigned long long)n * m) >> 40);
}

struct FftDesc {
    int N;
    int nst;
    int radix[kMaxStages];
    unsigned long long mN;                  // udiv_magic(N)
    unsigned long long mNR[kMaxStages];     // udiv_magic(N / radix[s])
    unsigned long long mP[kMaxStages];      // udiv_magic(radix[0] * .. * radix[s-1])
};

// One plane-ring field to build: D_{plane,ring}(x) = IFFT2(mask_ring * plane).
struct FieldTask {
    int32_t plane;    // kz plane (>= 0); -1 for the unit field (mask only)
    int32_t ring;
    int32_t fid;      // destination field slot in the level's cache
    int32_t rowband;  // rows |ia| <= rowband can hold masked modes
    int32_t nlo, nhi; // ring holds ia^2 + ja^2 in [nlo, nhi)
    int32_t zid;      // nonzero flag of (plane, ring): plane * Nr + ring
    int32_t level;    // grid level (index into the LevelDesc table)
    int32_t forms;    // results the ring pass writes: bit 0 the cache's own values, bit 1
                      // the bf16 hi/lo operands (when the launch has a split buffer)
};

// FFT geometry of one grid level.
struct LevelDesc {
    FftDesc fd;
    int tw_off;          // first twiddle of the level in the twiddle table
    int cpb, cpb_log2;   // columns per block of the ring pass (2^k)
    int max_rows;        // widest row band of the level's fields
    long long off;       // first point of the level in the cache
    long long stride;    // points between fields (multiple of 8)
};
// Columns per block of the ring pass for length N and values of `elem`
// bytes (8: FP64 caches, 4: FP32 caches).
int fft_cols_per_block(int N, int elem);

// ---------------------------------------------------------------------------
// Grid levels.  The sum over x of a triple product of fields band-limited to
// |k_x|, |k_y| <= K1, K2, K3 is the same (up to the factor (NN/NG)^2) on every
// grid of NG > K1 + K2 + K3 points per side: no sum of three modes can wrap
// around.  Each contraction runs on the smallest of a few such grids
// ("levels"); every level holds its own copy of the fields it needs.
constexpr int kMaxLevels = 16;

// ---------------------------------------------------------------------------
// Contraction work lists.
//
// A plane group (p1, p2, p3 = -p1-p2) contributes contractions
//   S = sum_x Re[ D_{p1,r1}(x) D~_{p2,r2}(x) D~_{p3,r3}(x) ].
// Its (r1, r2) pairs are packed 16 to a pair block; the r3 rings a pair block
// needs are covered by windows of 16 consecutive rings.  One (pair block,
// window) is one 16x16 output tile ("tile window"), a GEMM with K = points:
// rows P_m(x) = D1 D~2, columns the window's r3 fields.
//
// A work item gathers the tile windows of several plane groups that share
// their planes (groups with the same set {p1, |p2|, |p3|}).  A workgroup of
// kWaves waves stages the item's rings through LDS one chunk of points at a
// time; each wave owns up to kTWW tile windows and keeps their sums in
// registers across all chunks of its x-range.
constexpr int kWaves = 8;       // most waves of a contraction workgroup (4 or 8 are used)
constexpr int kTWW = 12;        // tile windows per wave (slots reserved per wave)
constexpr int kTWW32 = 8;       // tile windows per wave, BF16x3 form
constexpr int kTWWF32 = 6;      // tile windows per wave, FP32 form
constexpr int kMaxBlk = 48;     // ring blocks per item
constexpr int kRowCap = 240;    // staged rings per item (8 waves; 30 per wave)
__host__ __device__ constexpr int rowcap_for(int waves) { return 30 * waves; }
// Staged rings per item of the BF16x3 kernel sized for 3 workgroups per CU
// (six groups of 16 in each LDS chunk buffer).
__host__ __device__ constexpr int rowcap_occ3(int waves) { return 24 * waves; }
constexpr int kTWWOcc3 = 6;   // tile windows per wave, BF16x3 at 3 workgroups per CU

// Workgroups of `lds` bytes of LDS that fit one CU (160 KiB).  The LDS is
// allocated in granules of 1280 bytes: workgroups of 53 920 bytes ran 2 per
// CU although the occupancy query reports 3 (up to 54 613 bytes), and 53 536
// bytes ran 3.
constexpr int kLdsPerCU = 160 * 1024, kLdsGranule = 1280;
__host__ __device__ constexpr int lds_resident(long long lds)
{
    return lds <= 0 ? 1 << 20 : (int)(kLdsPerCU / ((lds + kLdsGranule - 1) / kLdsGranule * kLdsGranule));
}

enum : int16_t { TW_CONJ2 = 1, TW_CONJ3 = 2, TW_NEWA = 4 };

struct TWin {
    int16_t a1[16];   // row m: LDS row of D1 (-1: empty row)
    int16_t a2[16];   // row m: LDS row of D2
    int16_t b0;       // LDS row of the window's first r3 ring
    int16_t nb;       // valid columns
    int16_t flags;    // TW_*
    int16_t pad;
    int32_t out;      // first partial tile; x-split s writes tile out + s (-1: unused slot)
};

struct WorkItem {
    long long off;               // first point of the item's level in the cache
    long long npts;              // points per field of the level (NG^2)
    long long fstride;           // points between fields (npts rounded up to 32; zero padding)
    double scale;                // (NN / NG)^2 of the level
    int32_t blk_slot[kMaxBlk];   // field slot (in the level) of the block's first ring
    int16_t blk_rows[kMaxBlk];   // rings in the block (consecutive slots)
    int16_t blk_split[kMaxBlk];  // 1: staged as bf16 hi/lo B operands
    int16_t nblk, nrows;
    int32_t tw0;                 // first of kWaves * kTWW reserved TWins
    int8_t cnt[kWaves];          // TWins per wave
    // Staged row r: field slot (in the level) of its ring, with kRowSplit set
    // when the row is staged as bf16 hi/lo B operands.  Rows need not follow
    // the blocks' order: the BF16x3 layout places pair rings so that the 16
    // rings of a tile's A rows fall into distinct LDS bank groups.
    int32_t rslot[kRowCap];
};
constexpr int32_t kRowSplit = 1 << 30;

// One workgroup of the contraction launch: chunks [cb, ce) of an item's
// points, written to partial tile `split` of each of its windows.
struct WorkUnit {
    int32_t item, split, cb, ce;
};

// Contraction record used by the binning kernel (bin-sorted order).
// One contraction for the binning (20 bytes).  The nonzero flag of a field
// (the zero-field rule) is flag |p| * Nr + r.
struct BinRec {
    int16_t p1, p2, p3;
    int16_t nsp;             // partial sums (x-splits), 256 doubles apart
    uint8_t r1, r2, r3, pad; // rings (Nr <= 255)
    int32_t rp;              // ring triple (index into Num)
    int32_t pos;             // position of the contraction's first partial sum
};
static_assert(sizeof(BinRec) == 20, "binning record layout");

struct BinParams {
    int nbins;
    int n_mtbin;
    int NN;
    int Nr;        // rings per plane (nonzero flag of (p, r): |p| * Nr + r)
    double Cunits;
    double ff, beta;
};

// Launchers (fft.hip, contract.hip, binning.hip).
// nf fields of NN * NN * (NN/2+1) complex values each, back to back.
// zero (may be null): nzero ints the gather also sets to 0 (the fields'
// nonzero flags, which the ring pass then raises); one (may be null): nf
// ints it sets to 1 (the fields' flag-match results, which binning clears).
hipError_t launch_gather(const double2 *fields, double2 *planes, int NN, int NG, int kmax,
                         int Nmax, int nf, hipStream_t s, int *zero = nullptr, int nzero = 0,
                         int *one = nullptr);
// Per-field strides of a batched launch (blockIdx.y, or blockIdx.z for the
// gather, selects the field): elements between consecutive fields' buffers.
struct BatchStrides {
    long long planes = 0;   // double2: Nmax * NG * NG
    long long cache = 0;    // cache elements (double2 or float2; bf16 operands x4)
    long long nz = 0;       // nonzero flags
    long long part = 0;     // partial sums (doubles)
    long long bins = 0;     // bin outputs (doubles)
    long long binpart = 0;  // per-segment bin sums (doubles)
    long long binpart_s = 0;  // per-segment S sums (doubles)
};

// Work units of the ring pass: {task, first | (end << 16)} column-block
// ranges in device memory, built by fft_blocks on the host.
struct FftBlocks {
    const int2 *units = nullptr;
    int n = 0;
};
void fft_blocks(const FieldTask *tasks, int ntasks, const LevelDesc *h_lv,
                std::vector<int2> &units);
// The ring pass over every task (fields of all levels), writing the FP64
// cache, or the FP32 cache plus (split != null) the bf16 operands.
// planes: NP x NP per plane (signed modes wrapped), NP >= every level's N;
// d_lv / h_lv: the same LevelDesc table in device and host memory.
hipError_t launch_fft_f64(const double2 *planes, int NP, double2 *cache, const FieldTask *tasks,
                          const FftBlocks &fb, const LevelDesc *d_lv, const LevelDesc *h_lv,
                          int nlev, const double2 *tw, int *nzflag, const BatchStrides &bs,
                          int nf, hipStream_t s);
hipError_t launch_fft_f32(const double2 *planes, int NP, float2 *cache, void *split,
                          const FieldTask *tasks, const FftBlocks &fb, const LevelDesc *d_lv,
                          const LevelDesc *h_lv, int nlev, const double2 *tw, int *nzflag,
                          const BatchStrides &bs, int nf, hipStream_t s);

// Contraction arithmetic.
enum class CPrec { F64 = 0, F32 = 1, BF16X3 = 2 };
// Points per LDS chunk of each precision.
int contract_chunk(CPrec p);
// Tile windows per wave of each precision.
inline int contract_tww(CPrec p)
{
    return p == CPrec::F64 ? kTWW : p == CPrec::F32 ? kTWWF32 : kTWW32;
}
// cache: field values (FP64 or FP32) of all levels; split: bf16 hi/lo
// operands of the same fields at the same point offsets (BF16x3 only, rows
// flagged blk_split are read from it).  One workgroup per unit.
// waves: 4 or 8 waves per workgroup; tww: tile windows per wave (the work
// lists must be packed for both).
// One unit of a mixed BF16x3 launch: an estimate's field and work lists.
struct MixUnit {
    const float2 *cache;
    const __bf16 *split;
    const WorkItem *item;
    const TWin *tws;
    double *out;
    WorkUnit unit;
};
// All units of several BF16x3 estimates with 4-wave layouts of one shape
// (tww kTWW32 or kTWWOcc3) in one persistent launch; counter: one int of
// device memory (the unit queue).
hipError_t launch_contract_mixed(int tww, const MixUnit *units, int nunits, int *counter,
                                 hipStream_t s);
hipError_t launch_contract(CPrec p, int waves, int tww, const void *cache, const void *split,
                           const WorkItem *items, const WorkUnit *units, int nunits,
                           const TWin *tws, double *out, const BatchStrides &bs, int nf,
                           hipStream_t s);

// Records (sorted by bin) are split into segments of at most kBinSeg records
// within one bin: sg[s]..sg[s+1] are segment s's records, bin_sg[b]..
// bin_sg[b+1] the segments of bin b.  binpart holds 48 sums per segment.
constexpr int kBinSeg = 2048;
// Field-independent binning: the norm and analytic sums of each bin depend
// on the field only through its nonzero flags (the zero rule).  `sums` holds
// them ([3][nbins * 16], bins as launch_binning writes them; the first third
// unused) for the reference flags `ref`; launch_binning then accumulates
// the 16 S sums per bin alone (binpart_s, [segment][16] per field) and runs
// the full 48-sum pass only for fields whose flags differ (match[f] = 0,
// decided on the device: launch_gather sets match[f] = 1 when it clears
// the flags, the S pass clears it on a difference).  The bits are the same
// either way.
struct BinStatic {
    const int *ref = nullptr;
    int nflags = 0;
    const double *sums = nullptr;
    int *match = nullptr;          // [batch]
    double *binpart_s = nullptr;   // null: no field-independent sums
};
hipError_t launch_binning(const int32_t *sg, int nseg, const int32_t *bin_sg, const BinRec *recs,
                          const double *S, const double *num, const int *nzflag, const double *avg,
                          const BinParams &bp, double *binpart, double *bins_bs,
                          double *bins_norm, double *bins_biana, const BatchStrides &bs, int nf,
                          hipStream_t s, const double *W = nullptr, int nrec = 0,
                          const BinStatic &bst = BinStatic{});
// The records' field-independent binning coefficients, [33][nrec] doubles
// (k_binning's own computation): with W, launch_binning reads them instead
// of computing them for every field, with the same bits.
hipError_t launch_bin_coeffs(const BinRec *recs, int nrec, const double *avg, const BinParams &bp,
                             double *W, hipStream_t s);
// Largest coefficient table an estimator keeps (batches of more than one field).
constexpr size_t kBinTableBytes = 256ull << 20;

// v of the lane that the DPP control CTRL selects (within a row of 16
// lanes), for a double: both halves moved.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v)
{
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// Sum over the wave, the same bits in every lane: pairwise sums within
// pairs, quads (quad permutes), half rows and rows (mirrors) by DPP, then
// across the four rows by two lane exchanges.  Each step adds a lane's value
// and its partner's, so both end with the same sum.
__device__ __forceinline__ double wave_sum(double v)
{
    v += dpp_f64<0xB1>(v);    // quad_perm [1, 0, 3, 2]
    v += dpp_f64<0x4E>(v);    // quad_perm [2, 3, 0, 1]
    v += dpp_f64<0x141>(v);   // row_half_mirror
    v += dpp_f64<0x140>(v);   // row_mirror
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    return v;
}

// bb_tune with a switch for the mixed BF16x3 + FP32 candidates (allow_mix
// false: BF16x3 candidates use the BF16x3 kernel only).
int tune_layout(const bb_plan *p, const bb_options *base, const double *d_field, int repeats,
                bool allow_mix, bb_options *best, double *best_ms);

// Radices and division constants of a length-N transform (false: N is not
// 2,3,5-smooth or too long), and its twiddles exp(+2 pi i t / N).
bool fft_desc(int N, FftDesc &fd);
std::vector<double2> fft_twiddles(int N);

}  // namespace bb
