// Triple-product contraction on VALU (FP64 and FP32 paths).
//
// For a contraction (p1,r1,p2,r2,p3,r3) the reference computes
//     S = sum_x Re[ D_{p1,r1}(x) * D~_{p2,r2}(x) * D~_{p3,r3}(x) ]
// (barbie.c:326-337; D~ = conj(D_{|p|,r}) for a negative plane).  Here the
// sum is recast as S = sum_x Re[P(x) * C(x)] with the pair field
// P = D_{p1,r1} * D~_{p2,r2} formed once per point and reused across a window
// of r3 rings.  One workgroup evaluates one tile (kTileMP pairs x kTileNT r3)
// over all NN^2 points; each lane keeps the tile's partial sums in registers
// and the workgroup reduces them in a fixed order (deterministic).
#include <hip/hip_runtime.h>

#include "device.hpp"

namespace bb {

template <typename R>
struct Cplx;
template <>
struct Cplx<double> { using T = double2; };
template <>
struct Cplx<float> { using T = float2; };

// Blocks are dealt round-robin to the 8 XCDs; remap so that each XCD works
// on one contiguous range of tiles (tiles of a plane group are adjacent, so
// their fields stay in that XCD's L2).  grid = 8 * per blocks.
__device__ __forceinline__ int xcd_tile(int b, int per) { return (b & 7) * per + (b >> 3); }

template <typename R, typename A>
__device__ __forceinline__ A wave_sum(A v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <typename R>
__global__ __launch_bounds__(256) void k_contract(const typename Cplx<R>::T *__restrict__ cache,
                                                  long long npts, const Tile *__restrict__ tiles,
                                                  int ntiles, int per,
                                                  const int32_t *__restrict__ outmap,
                                                  double *__restrict__ S)
{
    using C2 = typename Cplx<R>::T;
    constexpr int MP = kTileMP, NT = kTileNT;
    const int t = xcd_tile(blockIdx.x, per);
    if (t >= ntiles) return;
    const Tile tl = tiles[t];
    const R s2 = tl.conj2 ? R(-1) : R(1);
    const R s3 = tl.conj3 ? R(-1) : R(1);
    int np = 0;
#pragma unroll
    for (int i = 0; i < MP; ++i) np += (tl.f1[i] >= 0);
    const int nt = tl.nt;

    const C2 *f1p[MP], *f2p[MP];
#pragma unroll
    for (int i = 0; i < MP; ++i) {
        f1p[i] = cache + (long long)max(tl.f1[i], 0) * npts;
        f2p[i] = cache + (long long)max(tl.f2[i], 0) * npts;
    }
    const C2 *f3p = cache + (long long)tl.f3 * npts;

    R acc[MP][NT];
#pragma unroll
    for (int i = 0; i < MP; ++i)
#pragma unroll
        for (int j = 0; j < NT; ++j) acc[i][j] = R(0);

    for (long long x = threadIdx.x; x < npts; x += blockDim.x) {
        R pr[MP], pi[MP];
#pragma unroll
        for (int i = 0; i < MP; ++i) {
            pr[i] = R(0);
            pi[i] = R(0);
            if (i < np) {
                C2 a = f1p[i][x], b = f2p[i][x];
                R bi = s2 * b.y;
                pr[i] = a.x * b.x - a.y * bi;
                pi[i] = a.x * bi + a.y * b.x;
            }
        }
#pragma unroll
        for (int j = 0; j < NT; ++j) {
            if (j < nt) {
                C2 c = f3p[(long long)j * npts + x];
                R ci = s3 * c.y;
#pragma unroll
                for (int i = 0; i < MP; ++i) acc[i][j] += pr[i] * c.x - pi[i] * ci;
            }
        }
    }

    // Reduce each accumulator: wave (64 lanes) then the 4 waves, in FP64.
    __shared__ double red[4][MP * NT];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < MP; ++i)
#pragma unroll
        for (int j = 0; j < NT; ++j) {
            double v = wave_sum<R, double>((double)acc[i][j]);
            if (lane == 0) red[wave][i * NT + j] = v;
        }
    __syncthreads();
    if (threadIdx.x < MP * NT) {
        const int k = threadIdx.x;
        const int dst = outmap[tl.out + k];
        if (dst >= 0) S[dst] = ((red[0][k] + red[1][k]) + red[2][k]) + red[3][k];
    }
}

static hipError_t launch(const void *cache, bool f64, long long npts, const Tile *tiles,
                         int ntiles, const int32_t *outmap, double *S, hipStream_t s)
{
    if (ntiles <= 0) return hipSuccess;
    int per = (ntiles + 7) / 8;
    dim3 grid(8 * per);
    if (f64)
        hipLaunchKernelGGL(k_contract<double>, grid, dim3(256), 0, s,
                           (const double2 *)cache, npts, tiles, ntiles, per, outmap, S);
    else
        hipLaunchKernelGGL(k_contract<float>, grid, dim3(256), 0, s,
                           (const float2 *)cache, npts, tiles, ntiles, per, outmap, S);
    return hipGetLastError();
}

hipError_t launch_contract_f64(const double2 *cache, long long npts, const Tile *tiles,
                               int ntiles, const int32_t *outmap, double *S, hipStream_t s)
{
    return launch(cache, true, npts, tiles, ntiles, outmap, S, s);
}

hipError_t launch_contract_f32(const float2 *cache, long long npts, const Tile *tiles,
                               int ntiles, const int32_t *outmap, double *S, hipStream_t s)
{
    return launch(cache, false, npts, tiles, ntiles, outmap, S, s);
}

}  // namespace bb
