// Power-spectrum multipoles of a field: the program calpow_mom
// (calpow_mom.c:22-117, funcs.c:860-1097) on the GPU.
//
// Every mode the reference's three loops visit (half lines :926-958, half
// planes :962-1001, half cube :1005-1049) is read once, by consecutive
// threads in the r2c buffer's memory order; its log bin, Legendre weights and
// |Delta|^2 are accumulated per thread for all bins (predicated adds, so
// that the accumulators stay in registers), reduced per workgroup in a fixed
// tree order and summed over workgroups in index order (deterministic).
//
// The per-mode arithmetic keeps the reference's types: m, mu and the P_l
// are `float`, formed from float products and double sqrt/division exactly
// as C evaluates the reference's expressions (contraction to FMA is off).
// The log bin floorf(log10(m N) / scale) is a non-decreasing step function
// of the float m N; its steps are found on the host with the C library's
// log10 (the reference's), so a mode's bin is decided by float compares and
// matches the reference's bit for bit.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "barbie.h"
#include "plan.hpp"

namespace {

constexpr int kPsMaxBins = 10;   // the reference's NB (calpow_mom.c:85)
constexpr int kPsThreads = 256;
constexpr int kPsBlocks = 1024;
constexpr int kPsAcc = 9;   // power, P2, P4, P6 weighted, kmode, no, no2, no4, no6

struct PsParams {
    int N, Nbin;
    float fac;
    float thr[kPsMaxBins + 1];   // bin d holds m N in [thr[d], thr[d+1])
};

// Mode (i, j, k) of the r2c buffer: 1 and (a, b) when a reference loop
// visits it (a, b: the folded indices that loop uses), else 0.
__device__ __forceinline__ bool visited(int N, int i, int j, int k, int &a, int &b)
{
    const int h = N / 2;
    if (k > 0 && k < h) {                                  // half cube
        a = i > h ? N - i : i;
        b = j > h ? N - j : j;
        return true;
    }
    if (j >= 1 && j < h) {                                 // half planes (k = 0, N/2)
        a = i > h ? N - i : i;
        b = j;
        return true;
    }
    if ((j == 0 || j == h) && i >= 1 && i <= h) {          // half lines
        a = i;
        b = j;
        return true;
    }
    return false;
}

__global__ __launch_bounds__(kPsThreads) void k_power(const double2 *__restrict__ field,
                                                      PsParams P, long long nmodes,
                                                      double *__restrict__ part)
{
#pragma clang fp contract(off)
    const int N = P.N, H = N / 2 + 1;
    double acc[kPsMaxBins][kPsAcc];
#pragma unroll
    for (int d = 0; d < kPsMaxBins; ++d)
#pragma unroll
        for (int q = 0; q < kPsAcc; ++q) acc[d][q] = 0.0;
    for (long long x = (long long)blockIdx.x * kPsThreads + threadIdx.x; x < nmodes;
         x += (long long)gridDim.x * kPsThreads) {
        const int k = (int)(x % H);
        const long long ij = x / H;
        const int j = (int)(ij % N), i = (int)(ij / N);
        int a, b;
        if (!visited(N, i, j, k, a, b)) continue;
        const float fac = P.fac;
        const float m = (float)sqrt((double)(fac * (float)a * (float)a + fac * (float)b * (float)b +
                                             fac * (float)k * (float)k));
        const float mu = (float)((1.0 * k) / (double)((float)N * m));
        const float P2 = (float)(0.5 * (3. * (double)mu * (double)mu - 1.));
        const float mu4 = powf(mu, 4.0f), mu6 = powf(mu, 6.0f);
        const float P4 = (float)(0.125 * (35.0 * (double)mu4 - 30.0 * (double)mu * (double)mu + 3));
        const float P6 = (float)(0.0625 * (231.0 * (double)mu6 - 315.0 * (double)mu4 +
                                           105.0 * (double)mu * (double)mu - 5));
        const float v = m * (float)N;
        int d = -1;
#pragma unroll
        for (int s = 0; s <= kPsMaxBins; ++s)
            if (s <= P.Nbin && v >= P.thr[s]) d = s;
        if (d < 0 || d >= P.Nbin) continue;
        const double2 z = field[x];
        const double p = z.x * z.x + z.y * z.y;
        const double w[kPsAcc] = {p,
                                  p * (double)P2,
                                  p * (double)P4,
                                  p * (double)P6,
                                  (double)m,
                                  1.0,
                                  (double)(P2 * P2),
                                  (double)(P4 * P4),
                                  (double)(P6 * P6)};
#pragma unroll
        for (int s = 0; s < kPsMaxBins; ++s)
            if (s == d)
#pragma unroll
                for (int q = 0; q < kPsAcc; ++q) acc[s][q] += w[q];
    }
    // Workgroup sums, a bin at a time: a fixed tree over the threads.
    __shared__ double red[kPsThreads][kPsAcc];
    for (int s = 0; s < P.Nbin; ++s) {
#pragma unroll
        for (int q = 0; q < kPsAcc; ++q) {
            double v = 0.0;
#pragma unroll
            for (int t = 0; t < kPsMaxBins; ++t)
                if (t == s) v = acc[t][q];
            red[threadIdx.x][q] = v;
        }
        __syncthreads();
        for (int w = kPsThreads / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w)
#pragma unroll
                for (int q = 0; q < kPsAcc; ++q) red[threadIdx.x][q] += red[threadIdx.x + w][q];
            __syncthreads();
        }
        if (threadIdx.x < kPsAcc)
            part[((long long)blockIdx.x * P.Nbin + s) * kPsAcc + threadIdx.x] = red[0][threadIdx.x];
        __syncthreads();
    }
}

// Sum over workgroups in index order: one thread per (bin, accumulator).
__global__ void k_power_sum(const double *__restrict__ part, int nblocks, int n,
                            double *__restrict__ out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double v = 0.0;
    for (int g = 0; g < nblocks; ++g) v += part[(long long)g * n + t];
    out[t] = v;
}

// Smallest positive float v with floorf(log10(v) / scale) >= d (the
// reference's bin rule, funcs.c:939), by bisection over the ordered bit
// patterns of positive floats.
float bin_threshold(float scale, int d)
{
    auto bin = [&](float v) { return (int)floorf((float)(log10((double)v) / (double)scale)); };
    uint32_t lo = 1, hi = 0x7f7fffffu;   // smallest subnormal .. largest finite
    float vhi;
    std::memcpy(&vhi, &hi, 4);
    if (bin(vhi) < d) return INFINITY;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        float v;
        std::memcpy(&v, &mid, 4);
        if (bin(v) >= d) hi = mid;
        else lo = mid + 1;
    }
    float v;
    std::memcpy(&v, &lo, 4);
    return v;
}

}  // namespace

extern "C" int bb_power_multipoles(const double *field_r2c, int on_device, int NN, int Nbin,
                                   double LL, double *out)
{
    using bb::set_error;
    if (!field_r2c || !out) return set_error(BB_E_INVALID, "null argument");
    if (NN < 4 || (NN & 1) || NN > 4096) return set_error(BB_E_INVALID, "NN must be even, 4..4096");
    if (Nbin < 1 || Nbin > kPsMaxBins)
        return set_error(BB_E_INVALID, "Nbin must be in [1, " + std::to_string(kPsMaxBins) + "]");
    if (!(LL > 0)) return set_error(BB_E_INVALID, "LL must be positive");
    const long long N = NN, H = N / 2 + 1, nmodes = N * N * H;
    PsParams P;
    std::memset(&P, 0, sizeof P);
    P.N = NN;
    P.Nbin = Nbin;
    P.fac = 1. / (1. * N * N);                        // funcs.c:917
    const float scale = (float)(log10(0.5 * N) / Nbin);   // funcs.c:924
    for (int d = 0; d <= kPsMaxBins; ++d) P.thr[d] = d <= Nbin ? bin_threshold(scale, d) : INFINITY;
    int dev_count = 0;
    if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count < 1) {
        (void)hipGetLastError();
        return set_error(BB_E_DEVICE, "no HIP device");
    }
    hipStream_t s = nullptr;
    double2 *d_field = nullptr;
    double *d_part = nullptr, *d_sum = nullptr;
    const int n = Nbin * kPsAcc;
    std::vector<double> acc((size_t)n);
    int rc = BB_OK;
    auto fail = [&](int code, const std::string &msg) {
        (void)hipGetLastError();
        if (!rc) rc = set_error(code, msg);
    };
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) fail(BB_E_DEVICE, "stream");
    const double2 *src = reinterpret_cast<const double2 *>(field_r2c);
    if (!rc && !on_device) {
        if (hipMalloc(&d_field, sizeof(double2) * nmodes) != hipSuccess)
            fail(BB_E_DEVMEM, "field buffer");
        else if (hipMemcpyAsync(d_field, field_r2c, sizeof(double2) * nmodes,
                                hipMemcpyHostToDevice, s) != hipSuccess)
            fail(BB_E_DEVICE, "field copy");
        src = d_field;
    }
    if (!rc && (hipMalloc(&d_part, sizeof(double) * n * kPsBlocks) != hipSuccess ||
                hipMalloc(&d_sum, sizeof(double) * n) != hipSuccess))
        fail(BB_E_DEVMEM, "partial sums");
    if (!rc) {
        hipLaunchKernelGGL(k_power, dim3(kPsBlocks), dim3(kPsThreads), 0, s, src, P, nmodes,
                           d_part);
        hipLaunchKernelGGL(k_power_sum, dim3((n + 127) / 128), dim3(128), 0, s, d_part, kPsBlocks,
                           n, d_sum);
        if (hipGetLastError() != hipSuccess) fail(BB_E_DEVICE, "power kernels");
    }
    if (!rc && (hipMemcpyAsync(acc.data(), d_sum, sizeof(double) * n, hipMemcpyDeviceToHost, s) !=
                    hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess))
        fail(BB_E_DEVICE, "power kernels failed");
    if (d_field) (void)hipFree(d_field);
    if (d_part) (void)hipFree(d_part);
    if (d_sum) (void)hipFree(d_sum);
    if (s) (void)hipStreamDestroy(s);
    if (rc) return rc;
    // Normalisation, funcs.c:1055-1071; columns of calpow_mom.c:108-111.
    const double norml = 1. / (1. * N * N * N), pi = M_PI;
    for (int i = 0; i < Nbin; ++i) {
        const double *A = acc.data() + (size_t)i * kPsAcc;
        double *o = out + 6 * i;
        o[0] = 0.0;
        o[1] = 0.0;
        o[2] = A[1];
        o[3] = A[2];
        o[4] = A[3];
        o[5] = A[5];
        if (A[5] > 0.0) {
            o[1] = norml * A[0] / (1.0 * A[5] * pow(LL, 3.));
            o[0] = 2. * pi * A[4] / (A[5] * LL);
        }
        if (A[6] > 0.0) o[2] = norml * A[1] / (1.0 * A[6] * pow(LL, 3.));
        if (A[7] > 0.0) o[3] = norml * A[2] / (1.0 * A[7] * pow(LL, 3.));
        if (A[8] > 0.0) o[4] = norml * A[3] / (1.0 * A[8] * pow(LL, 3.));
    }
    return BB_OK;
}
