// Power-spectrum multipoles of a field: the program calpow_mom
// (calpow_mom.c:22-117, funcs.c:860-1097) on the GPU.
//
// The reference's three loops (half lines :926-958, half planes :962-1001,
// half cube :1005-1049) visit each mode (i, j, k) of the r2c buffer with the
// folded indices a = min(i, N - i), b = min(j, N - j) (lines and planes fold
// fewer), and a mode's bin and Legendre weights depend on (a, b, k) only.
// The kernel therefore walks the (N/2 + 1)^3 folded triples: a thread
// computes the weights of a triple once and loads the field at its two to
// four mirror modes (i in {a, N - a}, j in {b, N - b}, those the reference
// visits), so that every summed mode is still read exactly once, by
// consecutive threads along k.  Triples whose integer a^2 + b^2 + k^2 lies
// beyond the last bin's edge (about half of the octant) are skipped before
// any floating-point work.  Each thread keeps (a, b, k) by carries instead
// of divisions and accumulates per bin in registers (branches per bin);
// workgroups reduce by wave sums (DPP) added in wave order, and a second
// kernel adds the workgroups the same way (deterministic).
//
// The per-mode arithmetic keeps the reference's types and roundings: m, mu
// and the P_l are `float`, formed as C evaluates the reference's expressions
// (contraction to FMA is off); the powf powers of mu are the correctly
// rounded ones.
// The log bin floorf(log10(m N) / scale) is a non-decreasing step function
// of the float m N; its steps are found on the host with the C library's
// log10 (the reference's), so a mode's bin is decided by float compares and
// matches the reference's bit for bit.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "barbie.h"
#include "device.hpp"
#include "plan.hpp"

namespace {

constexpr int kPsMaxBins = 10;   // the reference's NB (calpow_mom.c:85)
constexpr int kPsThreads = 256;
// Workgroups per pass: one round of the chip at the pass's occupancy (3 and
// 4 workgroups of 4 waves per CU at 140 and 124 VGPRs, 256 CUs).
constexpr int kPsBlocksField = 768;
constexpr int kPsBlocksGeom = 1024;
constexpr int kPsBlocks = kPsBlocksGeom;   // partial-sum rows allocated
// Sums per bin: of the field (power and its P2, P4, P6 weighted forms) and
// of the geometry alone (kmode, no, no2, no4, no6 of funcs.c:940-945), which
// depend on N and the bins only and are computed once per (N, Nbin).
constexpr int kPsField = 4;
constexpr int kPsGeom = 5;
constexpr int kPsAcc = kPsField + kPsGeom;
constexpr int kPsOut = kPsMaxBins * kPsGeom;   // the larger of the two passes

// A folded triple of the walk: whether it is summed, which of its mirror
// modes 1..3 are (mode 0 is whenever the triple is), and their values.
struct Triple {
    double2 z[4];
    int a, b, k;
    bool ok, v1, v2, v3;
};

struct PsParams {
    int N, Nbin;
    float fac;
    int r2_cut;                  // a^2 + b^2 + k^2 above this: beyond the last bin
    int sA, sB, sK;              // the grid stride as (a, b, k) steps
    float thr[kPsMaxBins + 1];   // bin d holds m N in [thr[d], thr[d+1])
};

// Log bin of the folded triple (a, b, k) and its weights (funcs.c:932-945):
// -1 when the triple lies outside the bins.
__device__ __forceinline__ int triple_terms(const PsParams &P, int a, int b, int ck, float &m,
                                            float &P2, float &P4, float &P6)
{
#pragma clang fp contract(off)
    const int N = P.N;
    const float fac = P.fac;
    // The reference takes the double sqrt of a float and divides the double
    // k by a float, rounding both to float.  Double has more than 2 x 24 + 2
    // bits, so these equal the correctly rounded float sqrt and quotient
    // (double rounding is innocuous for sqrt and division at that width),
    // which sqrtf and / are here (-fhip-fp32-correctly-rounded-divide-sqrt).
    m = sqrtf(fac * (float)a * (float)a + fac * (float)b * (float)b + fac * (float)ck * (float)ck);
    const float mu = (float)ck / ((float)N * m);
    P2 = (float)(0.5 * (3. * (double)mu * (double)mu - 1.));
    // powf(mu, 4) and powf(mu, 6) of the reference: mu^2 is exact in
    // double, and one rounding to float of the double products gives the
    // correctly rounded powers.  The C library's powf is within an ulp of
    // them and equal for all but 0.07% of float arguments in [2^-20, 1].
    const double mu2 = (double)mu * (double)mu;
    const float mu4 = (float)(mu2 * mu2), mu6 = (float)(mu2 * mu2 * mu2);
    P4 = (float)(0.125 * (35.0 * (double)mu4 - 30.0 * (double)mu * (double)mu + 3));
    P6 = (float)(0.0625 * (231.0 * (double)mu6 - 315.0 * (double)mu4 +
                           105.0 * (double)mu * (double)mu - 5));
    const float v = m * (float)N;
    int d = -1;
#pragma unroll
    for (int s = 0; s <= kPsMaxBins; ++s)
        if (s <= P.Nbin && v >= P.thr[s]) d = s;
    return d < P.Nbin ? d : -1;
}

// kGeom: the geometry sums (no field access); else the field sums.
template <bool kGeom>
__global__ __launch_bounds__(kPsThreads) void k_power(const double2 *__restrict__ field,
                                                      PsParams P, long long ntriples,
                                                      double *__restrict__ part)
{
#pragma clang fp contract(off)
    constexpr int NA = kGeom ? kPsGeom : kPsField;
    const int N = P.N, h = N / 2, H = h + 1;
    double acc[kPsMaxBins][NA];
#pragma unroll
    for (int d = 0; d < kPsMaxBins; ++d)
#pragma unroll
        for (int q = 0; q < NA; ++q) acc[d][q] = 0.0;
    const long long stride = (long long)gridDim.x * kPsThreads;
    long long t = (long long)blockIdx.x * kPsThreads + threadIdx.x;
    // (a, b, k): the next triple to probe, kept by carries.
    int k = (int)(t % H), b, a;
    {
        const long long ab = t / H;
        b = (int)(ab % H);
        a = (int)(ab / H);
    }
    // Probe of triple tp: whether it is summed, which mirror modes the
    // reference visits, and the loads of their values; then one stride on.
    // Mirror modes: (a, b), (N-a, b), (a, N-b), (N-a, N-b) for 0 < k < N/2
    // (the second of a pair only when the index is neither 0 nor N/2); for
    // k = 0 and N/2 the half planes (1 <= b < N/2: i in {a, N-a}, j = b) and
    // half lines (b in {0, N/2}, 1 <= a <= N/2: (a, b) only).  All four
    // loads are always issued (a mirror that is not summed reads element 0,
    // one line shared by all waves), so that the count of loads in flight
    // is fixed and a use waits for its own loads only.
    auto probe = [&](long long tp, Triple &q) {
        const bool a2 = a > 0 && a < h, b2 = b > 0 && b < h;
        const bool inner = k > 0 && k < h;
        q.ok = tp < ntriples && a * a + b * b + k * k <= P.r2_cut && (inner || b2 || a > 0);
        q.v1 = a2 && (inner || b2);
        q.v2 = inner && b2;
        q.v3 = inner && a2 && b2;
        q.a = a;
        q.b = b;
        q.k = k;
        const long long r0 = (long long)a * N, r1 = (long long)(N - a) * N;
        const long long x0 = (r0 + b) * H + k, x1 = (r1 + b) * H + k;
        const long long x2 = (r0 + (N - b)) * H + k, x3 = (r1 + (N - b)) * H + k;
        if constexpr (!kGeom) {
            q.z[0] = field[q.ok ? x0 : 0];
            q.z[1] = field[q.ok && q.v1 ? x1 : 0];
            q.z[2] = field[q.ok && q.v2 ? x2 : 0];
            q.z[3] = field[q.ok && q.v3 ? x3 : 0];
        }
        k += P.sK;
        if (k >= H) {
            k -= H;
            ++b;
        }
        b += P.sB;
        if (b >= H) {
            b -= H;
            ++a;
        }
        a += P.sA;
    };
    // Two slots: a slot's triple is summed, then the slot is refilled two
    // strides on, while the other slot's loads are in flight.
    Triple q0, q1;
    probe(t, q0);
    probe(t + stride, q1);
    for (; t < ntriples; t += 2 * stride) {
#pragma unroll
        for (int slot = 0; slot < 2; ++slot) {
            Triple &q = slot ? q1 : q0;
            if (q.ok) {
                float m, P2, P4, P6;
                const int d = triple_terms(P, q.a, q.b, q.k, m, P2, P4, P6);
                if (d >= 0) {
                    double w[NA];
                    if constexpr (kGeom) {
                        const double cnt = 1.0 + (double)q.v1 + (double)q.v2 + (double)q.v3;
                        w[0] = cnt * (double)m;
                        w[1] = cnt;
                        w[2] = cnt * (double)(P2 * P2);
                        w[3] = cnt * (double)(P4 * P4);
                        w[4] = cnt * (double)(P6 * P6);
                    } else {
                        double p = q.z[0].x * q.z[0].x + q.z[0].y * q.z[0].y;
                        if (q.v1) p += q.z[1].x * q.z[1].x + q.z[1].y * q.z[1].y;
                        if (q.v2) p += q.z[2].x * q.z[2].x + q.z[2].y * q.z[2].y;
                        if (q.v3) p += q.z[3].x * q.z[3].x + q.z[3].y * q.z[3].y;
                        w[0] = p;
                        w[1] = p * (double)P2;
                        w[2] = p * (double)P4;
                        w[3] = p * (double)P6;
                    }
#pragma unroll
                    for (int s = 0; s < kPsMaxBins; ++s)
                        if (s == d)
#pragma unroll
                            for (int u = 0; u < NA; ++u) acc[s][u] += w[u];
                }
            }
            probe(t + (2 + slot) * stride, q);
        }
    }
    // Workgroup sums: each wave's (DPP within rows, lane exchanges across
    // rows; the same bits in every lane), then the waves' in wave order.
    __shared__ double red[kPsThreads / 64][kPsMaxBins * NA];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int s = 0; s < kPsMaxBins; ++s)
#pragma unroll
        for (int q = 0; q < NA; ++q) {
            const double v = bb::wave_sum(acc[s][q]);
            if (lane == 0) red[wave][s * NA + q] = v;
        }
    __syncthreads();
    for (int t = threadIdx.x; t < P.Nbin * NA; t += kPsThreads) {
        double v = red[0][t];
        for (int w = 1; w < kPsThreads / 64; ++w) v += red[w][t];
        part[(long long)blockIdx.x * P.Nbin * NA + t] = v;
    }
}

// Sum over workgroups: one workgroup per (bin, accumulator); thread t adds
// the partials t, t + 256, ... in order, then wave sums and the waves' in
// order.
__global__ __launch_bounds__(kPsThreads) void k_power_sum(const double *__restrict__ part,
                                                          int nblocks, int n,
                                                          double *__restrict__ out)
{
    __shared__ double red[kPsThreads / 64];
    const int o = blockIdx.x, t = threadIdx.x;
    double v = 0.0;
    for (int g = t; g < nblocks; g += kPsThreads) v += part[(long long)g * n + o];
    v = bb::wave_sum(v);
    if ((t & 63) == 0) red[t >> 6] = v;
    __syncthreads();
    if (t == 0) {
        double sum = red[0];
        for (int w = 1; w < kPsThreads / 64; ++w) sum += red[w];
        out[o] = sum;
    }
}

// Per-device partial-sum buffers and stream, made on a device's first call
// and kept until the process ends; calls are serialised by ps_mu.
struct PsScratch {
    double *part = nullptr, *sum = nullptr;
    hipStream_t stream = nullptr;
};
std::mutex ps_mu;
std::vector<PsScratch> ps_scratch;
// Geometry sums per (N, Nbin): [bin][kPsGeom], the same on every device.
std::map<std::pair<int, int>, std::vector<double>> ps_geom;

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

// Validates the arguments and fills the kernel parameters (bins, the
// thresholds of the log bins, the triple cut).
int power_params(int NN, int Nbin, double LL, PsParams &P)
{
    using bb::set_error;
    if (NN < 4 || (NN & 1) || NN > 4096) return set_error(BB_E_INVALID, "NN must be even, 4..4096");
    if (Nbin < 1 || Nbin > kPsMaxBins)
        return set_error(BB_E_INVALID, "Nbin must be in [1, " + std::to_string(kPsMaxBins) + "]");
    if (!(LL > 0)) return set_error(BB_E_INVALID, "LL must be positive");
    const long long N = NN;
    std::memset(&P, 0, sizeof P);
    P.N = NN;
    P.Nbin = Nbin;
    P.fac = 1. / (1. * N * N);                        // funcs.c:917
    const float scale = (float)(log10(0.5 * N) / Nbin);   // funcs.c:924
    for (int d = 0; d <= kPsMaxBins; ++d) P.thr[d] = d <= Nbin ? bin_threshold(scale, d) : INFINITY;
    // The float m N of a mode is within 1e-6 of sqrt(a^2 + b^2 + k^2), so
    // folded triples with the integer sum above (thr[Nbin] (1 + 1e-4))^2
    // have a bin of at least Nbin and are not summed.
    const double edge = (double)P.thr[Nbin] * (1.0 + 1e-4);
    P.r2_cut = std::isfinite(edge) && edge * edge < (double)INT_MAX ? (int)(edge * edge) : INT_MAX;
    return BB_OK;
}

// Launches one pass (geometry or field sums) and its sum over workgroups:
// partials in part, the Nbin * (kPsGeom or kPsField) sums in sum.
hipError_t power_pass(bool geom, const double2 *src, PsParams P, double *part, double *sum,
                      hipStream_t s)
{
    const long long H = P.N / 2 + 1, ntriples = H * H * H;
    const int n = P.Nbin * (geom ? kPsGeom : kPsField);
    const int nb = geom ? kPsBlocksGeom : kPsBlocksField;
    const long long stride = (long long)nb * kPsThreads;
    P.sK = (int)(stride % H);
    P.sB = (int)((stride / H) % H);
    P.sA = (int)(stride / H / H);
    if (geom)
        hipLaunchKernelGGL(k_power<true>, dim3(nb), dim3(kPsThreads), 0, s, src, P, ntriples, part);
    else
        hipLaunchKernelGGL(k_power<false>, dim3(nb), dim3(kPsThreads), 0, s, src, P, ntriples,
                           part);
    hipLaunchKernelGGL(k_power_sum, dim3(n), dim3(kPsThreads), 0, s, part, nb, n, sum);
    return hipGetLastError();
}

// The rows {kmean, P0, P2, P4, P6, num} from the field sums fsum
// [bin][kPsField] and the geometry sums geom [bin][kPsGeom]
// (normalisation of funcs.c:1055-1071, columns of calpow_mom.c:108-111).
void power_rows(int NN, int Nbin, double LL, const double *fsum, const double *geom, double *out)
{
    const long long N = NN;
    double acc[kPsMaxBins * kPsAcc];
    for (int i = 0; i < Nbin; ++i)
        for (int q = 0; q < kPsAcc; ++q)
            acc[i * kPsAcc + q] = q < kPsField ? fsum[i * kPsField + q]
                                               : geom[(size_t)i * kPsGeom + q - kPsField];
    const double norml = 1. / (1. * N * N * N), pi = M_PI;
    for (int i = 0; i < Nbin; ++i) {
        const double *A = acc + (size_t)i * kPsAcc;
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
}

}  // namespace

extern "C" int bb_power_multipoles(const double *field_r2c, int on_device, int NN, int Nbin,
                                   double LL, double *out)
{
    using bb::set_error;
    if (!field_r2c || !out) return set_error(BB_E_INVALID, "null argument");
    PsParams P;
    int rc = power_params(NN, Nbin, LL, P);
    if (rc) return rc;
    const long long N = NN, H = N / 2 + 1, nmodes = N * N * H;
    int dev = 0, dev_count = 0;
    if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count < 1 ||
        hipGetDevice(&dev) != hipSuccess) {
        (void)hipGetLastError();
        return set_error(BB_E_DEVICE, "no HIP device");
    }
    std::lock_guard<std::mutex> lock(ps_mu);
    if ((int)ps_scratch.size() < dev_count) ps_scratch.resize(dev_count);
    PsScratch &sc = ps_scratch[dev];
    if (!sc.part) {
        if (hipMalloc(&sc.part, sizeof(double) * kPsOut * kPsBlocks) != hipSuccess ||
            hipMalloc(&sc.sum, sizeof(double) * kPsOut) != hipSuccess) {
            (void)hipGetLastError();
            if (sc.part) (void)hipFree(sc.part);
            sc.part = nullptr;
            return set_error(BB_E_DEVMEM, "partial sums");
        }
        if (hipStreamCreateWithFlags(&sc.stream, hipStreamNonBlocking) != hipSuccess) {
            (void)hipGetLastError();
            (void)hipFree(sc.part);
            (void)hipFree(sc.sum);
            sc = PsScratch{};
            return set_error(BB_E_DEVICE, "stream");
        }
    }
    hipStream_t s = sc.stream;
    double2 *d_field = nullptr;
    auto fail = [&](int code, const std::string &msg) {
        (void)hipGetLastError();
        if (!rc) rc = set_error(code, msg);
    };
    // One pass and its sum over workgroups into host memory.
    auto run = [&](bool geom, const double2 *src, double *host) {
        const int n = Nbin * (geom ? kPsGeom : kPsField);
        if (power_pass(geom, src, P, sc.part, sc.sum, s) != hipSuccess)
            return fail(BB_E_DEVICE, "power kernels");
        if (hipMemcpyAsync(host, sc.sum, sizeof(double) * n, hipMemcpyDeviceToHost, s) !=
                hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            fail(BB_E_DEVICE, "power kernels failed");
    };
    std::vector<double> &geom = ps_geom[{NN, Nbin}];
    if (geom.empty()) {
        std::vector<double> g((size_t)Nbin * kPsGeom);
        run(true, nullptr, g.data());
        if (rc) {
            ps_geom.erase({NN, Nbin});
            return rc;
        }
        geom = std::move(g);
    }
    const double2 *src = reinterpret_cast<const double2 *>(field_r2c);
    if (!on_device) {
        if (hipMalloc(&d_field, sizeof(double2) * nmodes) != hipSuccess)
            fail(BB_E_DEVMEM, "field buffer");
        else if (hipMemcpyAsync(d_field, field_r2c, sizeof(double2) * nmodes,
                                hipMemcpyHostToDevice, s) != hipSuccess)
            fail(BB_E_DEVICE, "field copy");
        src = d_field;
    }
    double fsum[kPsMaxBins * kPsField];
    if (!rc) run(false, src, fsum);
    if (d_field) {
        (void)hipStreamSynchronize(s);
        (void)hipFree(d_field);
    }
    if (rc) return rc;
    power_rows(NN, Nbin, LL, fsum, geom.data(), out);
    return BB_OK;
}

// ------------------------------------------------- the no-alloc variant --
namespace {

// Device scratch of bb_power_multipoles_noalloc: the workgroup partials of
// one pass (kPsBlocks rows of kPsOut sums), then the geometry sums and the
// field sums (kPsOut each).  Every pass overwrites the partials it reads, so
// nothing needs clearing between calls.  Each device holds its own copy.
constexpr int kNaPart = kPsOut * kPsBlocks;
constexpr int kNaDoubles = kNaPart + 2 * kPsOut;
__device__ __align__(256) double scratch[kNaDoubles];

// Device address of `scratch` per device, looked up once per device with that
// device current (the host-side name is not a device address).
constexpr int kNaMaxDevices = 64;
double *na_ptr[kNaMaxDevices] = {};
std::mutex na_mu;

}  // namespace

extern "C" int bb_power_multipoles_noalloc(const double *d_field_r2c, int NN, int Nbin,
                                           double LL, void *stream, double *out)
{
    using bb::set_error;
    if (!d_field_r2c || !out) return set_error(BB_E_INVALID, "null argument");
    PsParams P;
    int rc = power_params(NN, Nbin, LL, P);
    if (rc) return rc;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        (void)hipGetLastError();
        return set_error(BB_E_DEVICE, "no HIP device");
    }
    if (dev < 0 || dev >= kNaMaxDevices)
        return set_error(BB_E_DEVICE, "device index beyond the supported 64");
    std::lock_guard<std::mutex> lock(na_mu);
    if (!na_ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            (void)hipGetLastError();
            return set_error(BB_E_DEVICE, "scratch symbol lookup failed");
        }
        na_ptr[dev] = static_cast<double *>(p);
    }
    double *part = na_ptr[dev], *gsum = part + kNaPart, *fsum_d = gsum + kPsOut;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const double2 *src = reinterpret_cast<const double2 *>(d_field_r2c);
    double geom[kPsMaxBins * kPsGeom], fsum[kPsMaxBins * kPsField];
    if (power_pass(true, nullptr, P, part, gsum, s) != hipSuccess ||
        power_pass(false, src, P, part, fsum_d, s) != hipSuccess) {
        (void)hipGetLastError();
        return set_error(BB_E_DEVICE, "power kernels");
    }
    if (hipMemcpyAsync(geom, gsum, sizeof(double) * Nbin * kPsGeom, hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipMemcpyAsync(fsum, fsum_d, sizeof(double) * Nbin * kPsField, hipMemcpyDeviceToHost,
                       s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        (void)hipGetLastError();
        return set_error(BB_E_DEVICE, "power kernels failed");
    }
    power_rows(NN, Nbin, LL, fsum, geom, out);
    return BB_OK;
}
