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

// A contraction tile: up to MP pair fields P_i = D_{f1[i]} * D~_{f2[i]} and a
// window of nt consecutive r3 fields starting at f3, all in one plane group.
// acc[i][j] = sum_x Re(P_i(x) * D~_{f3+j}(x)); out[i*NT + j] is the
// contraction index the result belongs to, or -1.
constexpr int kTileMP = 4;
constexpr int kTileNT = 16;
struct Tile {
    int32_t f1[kTileMP];
    int32_t f2[kTileMP];
    int32_t f3;
    int32_t nt;
    int32_t conj2;    // 1: second operand is conjugated (negative plane)
    int32_t conj3;    // 1: third operand is conjugated
    int32_t out;      // offset into the out map (kTileMP*kTileNT entries)
    int32_t pad[3];
};

// Contraction record used by the binning kernel (bin-sorted order).
struct BinRec {
    int16_t p1, p2, p3, r1, r2, r3;
    int32_t rp;      // ring triple (index into Num)
    int32_t cidx;    // contraction index (reference enumeration order)
    int32_t f1, f2, f3;  // field slots (zero-field rule)
};

struct BinParams {
    int nbins;
    int n_mtbin;
    int NN;
    double LL_k1, BW_k1;
    double Cunits;
    double ff, beta;
};

// Launchers (fft.hip, contract.hip, binning.hip).
hipError_t launch_gather(const double2 *field, double2 *planes, int NN, int Nmax, hipStream_t s);
hipError_t launch_rowpass(const double2 *planes, double2 *rowbuf, long long buf_stride,
                          const FieldTask *tasks, int ntasks, int max_rows, const FftDesc &fd,
                          const double2 *tw, int *nzflag, hipStream_t s);
hipError_t launch_colpass_f64(const double2 *rowbuf, long long buf_stride, double2 *cache,
                              const FieldTask *tasks, int ntasks, const FftDesc &fd,
                              const double2 *tw, hipStream_t s);
hipError_t launch_colpass_f32(const double2 *rowbuf, long long buf_stride, float2 *cache,
                              const FieldTask *tasks, int ntasks, const FftDesc &fd,
                              const double2 *tw, hipStream_t s);
hipError_t launch_contract_f64(const double2 *cache, long long npts, const Tile *tiles,
                               int ntiles, const int32_t *outmap, double *S, hipStream_t s);
hipError_t launch_contract_f32(const float2 *cache, long long npts, const Tile *tiles,
                               int ntiles, const int32_t *outmap, double *S, hipStream_t s);
hipError_t launch_binning(const int32_t *seg, const BinRec *recs, const double *S,
                          const double *num, const int *nzflag, const double *avg,
                          const BinParams &bp, double *bins_bs, double *bins_norm,
                          double *bins_biana, hipStream_t s);

}  // namespace bb
