// Device-side data structures and kernel launchers shared by the estimator
// (internal to libbarbie_amd).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace bb {

constexpr int kMaxStages = 16;

// Factorisation of an FFT length into radices 4, 2, 3, 5 (Stockham order).
struct FftDesc {
    int N;
    int nst;
    int radix[kMaxStages];
};

// One plane-ring field to build: D_{plane,ring}(x) = IFFT2(mask_ring * plane).
struct FieldTask {
    int32_t plane;    // kz plane (>= 0); -1 for the unit field (mask only)
    int32_t ring;
    int32_t fid;      // destination field slot in the cache
    int32_t rowband;  // rows |ia| <= rowband can hold masked modes
    int32_t nlo, nhi; // ring holds ia^2 + ja^2 in [nlo, nhi)
};

// ---------------------------------------------------------------------------
// Contraction work lists.
//
// A plane group (p1, p2, p3 = -p1-p2) contributes contractions
//   S = sum_x Re[ D_{p1,r1}(x) D~_{p2,r2}(x) D~_{p3,r3}(x) ].
// Its (r1, r2) pairs are packed 16 to an M-tile; the r3 rings an M-tile needs
// are covered by up to kMaxWin windows of 16 consecutive rings.  One
// (M-tile, window) is one 16x16 output tile, evaluated as a GEMM with
// K = points: rows P_m(x) = D1 D~2, columns D~3.
//
// A work item is one group's run of up to kMaxMT M-tiles; a workgroup stages
// the item's fields (three contiguous ring blocks, one per plane) through LDS
// one x-chunk at a time.
constexpr int kMaxWin = 3;
constexpr int kMaxMT = 8;       // 4 waves x 2 M-tiles

struct MTile {
    int16_t a1[16];      // row m: local ring index in block 0 (-1: empty row)
    int16_t a2[16];      // row m: local ring index in block 1
    int16_t w0[kMaxWin]; // window v: first local ring index in block 2
    int16_t wn[kMaxWin]; // window v: valid columns (0: no window)
    int32_t tw;          // output tile index of window 0 (windows follow)
};

struct WorkItem {
    int32_t fbase[3];    // cache slot of the first ring of each block
    int16_t nf[3];       // rings per block
    int16_t conj2, conj3;
    int16_t n_mt;        // M-tiles in this item (<= kMaxMT)
    int32_t mt0;         // first M-tile
};

// Contraction record used by the binning kernel (bin-sorted order).
struct BinRec {
    int16_t p1, p2, p3, r1, r2, r3;
    int32_t rp;      // ring triple (index into Num)
    int32_t pos;     // position of the contraction's sum in the tile outputs
    int32_t f1, f2, f3;  // field slots (zero-field rule)
};

struct BinParams {
    int nbins;
    int n_mtbin;
    int NN;
    int nsplit;              // partial sums per contraction (x-splits)
    long long split_stride;  // doubles between splits
    double Cunits;
    double ff, beta;
};

// Launchers (fft.hip, contract.hip, binning.hip).
hipError_t launch_gather(const double2 *field, double2 *planes, int NN, int NG, int kmax,
                         int Nmax, hipStream_t s);
hipError_t launch_rowpass(const double2 *planes, double2 *rowbuf, long long buf_stride,
                          const FieldTask *tasks, int ntasks, int max_rows, const FftDesc &fd,
                          const double2 *tw, int *nzflag, hipStream_t s);
hipError_t launch_colpass_f64(const double2 *rowbuf, long long buf_stride, double2 *cache,
                              const FieldTask *tasks, int ntasks, const FftDesc &fd,
                              const double2 *tw, hipStream_t s);
hipError_t launch_colpass_f32(const double2 *rowbuf, long long buf_stride, float2 *cache,
                              const FieldTask *tasks, int ntasks, const FftDesc &fd,
                              const double2 *tw, hipStream_t s);

// Contraction arithmetic.
enum class CPrec { F64 = 0, F32 = 1, BF16X3 = 2 };
// Points per LDS chunk for a given precision and maximum staged fields.
int contract_chunk(CPrec p, int max_fields);
hipError_t launch_contract(CPrec p, const void *cache, long long npts, const WorkItem *items,
                           int nitems, const MTile *mtiles, int nsplit, int chunk, int max_fields,
                           double *out, long long split_stride, double scale, hipStream_t s);

hipError_t launch_binning(const int32_t *seg, const BinRec *recs, const double *S,
                          const double *num, const int *nzflag, const double *avg,
                          const BinParams &bp, double *bins_bs, double *bins_norm,
                          double *bins_biana, hipStream_t s);

}  // namespace bb
