// Batched 1-D Stockham FFTs in LDS (radices 4, 2, 3, 5), shared by the
// plane-ring transforms (fft.hip) and the field generator (field.hip).
// Internal to libbarbie_amd.
#pragma once
#include <hip/hip_runtime.h>

#include "device.hpp"

namespace bb {

// Complex arithmetic on double2 (FP64 caches) or float2 (FP32 caches).
template <typename V2>
__device__ __forceinline__ V2 mk(decltype(V2::x) x, decltype(V2::x) y)
{
    V2 r;
    r.x = x;
    r.y = y;
    return r;
}
template <typename V2>
__device__ __forceinline__ V2 cmul(V2 a, V2 b)
{
    return mk<V2>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename V2>
__device__ __forceinline__ V2 cadd(V2 a, V2 b) { return mk<V2>(a.x + b.x, a.y + b.y); }
template <typename V2>
__device__ __forceinline__ V2 csub(V2 a, V2 b) { return mk<V2>(a.x - b.x, a.y - b.y); }
// multiply by +i
template <typename V2>
__device__ __forceinline__ V2 cmuli(V2 a) { return mk<V2>(-a.y, a.x); }

// Length-R DFT with exp(+2 pi i q s / R), in place, for R = 2, 3, 4, 5.
template <int R, typename V2>
struct Butterfly;
template <typename V2>
struct Butterfly<4, V2> {
    static __device__ __forceinline__ void run(V2 *v)
    {
        V2 a = v[0], b = v[1], c = v[2], d = v[3];
        V2 s0 = cadd(a, c), d0 = csub(a, c), s1 = cadd(b, d), d1 = cmuli(csub(b, d));
        v[0] = cadd(s0, s1);
        v[1] = cadd(d0, d1);
        v[2] = csub(s0, s1);
        v[3] = csub(d0, d1);
    }
};
template <typename V2>
struct Butterfly<2, V2> {
    static __device__ __forceinline__ void run(V2 *v)
    {
        V2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};
template <typename V2>
struct Butterfly<3, V2> {
    static __device__ __forceinline__ void run(V2 *v)
    {
        using R = decltype(V2::x);
        const R h = (R)0.86602540378443864676;  // sqrt(3)/2
        V2 a = v[0], t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
        V2 m = mk<V2>(a.x - (R)0.5 * t1.x, a.y - (R)0.5 * t1.y);
        V2 j = mk<V2>(-h * t2.y, h * t2.x);  // i*h*t2
        v[0] = cadd(a, t1);
        v[1] = cadd(m, j);
        v[2] = csub(m, j);
    }
};
template <typename V2>
struct Butterfly<5, V2> {
    static __device__ __forceinline__ void run(V2 *v)
    {
        using R = decltype(V2::x);
        const R c1 = (R)0.30901699437494742410, c2 = (R)-0.80901699437494742410;
        const R s1 = (R)0.95105651629515357212, s2 = (R)0.58778525229247312917;
        V2 a = v[0];
        V2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
        V2 u1 = csub(v[1], v[4]), u2 = csub(v[2], v[3]);
        V2 m1 = mk<V2>(a.x + c1 * t1.x + c2 * t2.x, a.y + c1 * t1.y + c2 * t2.y);
        V2 m2 = mk<V2>(a.x + c2 * t1.x + c1 * t2.x, a.y + c2 * t1.y + c1 * t2.y);
        // i*(s1*u1 + s2*u2) and i*(s2*u1 - s1*u2)
        V2 n1 = mk<V2>(-(s1 * u1.y + s2 * u2.y), s1 * u1.x + s2 * u2.x);
        V2 n2 = mk<V2>(-(s2 * u1.y - s1 * u2.y), s2 * u1.x - s1 * u2.x);
        v[0] = cadd(a, cadd(t1, t2));
        v[1] = cadd(m1, n1);
        v[4] = csub(m1, n1);
        v[2] = cadd(m2, n2);
        v[3] = csub(m2, n2);
    }
};

// One Stockham stage of radix R over nb transforms of length N: reads
// x[i + t N/R], twiddles by W_{pR}^{t k} (k = i mod p), writes
// y[(i-k) R + k + t p].
template <int R, typename V2>
__device__ __forceinline__ void fft_stage(const V2 *a, V2 *b, int nb, int N, int p,
                                          unsigned long long mNR, unsigned long long mP,
                                          const V2 *tw)
{
    const unsigned NR = (unsigned)(N / R);
    const unsigned tstep = (unsigned)(N / (p * R));
    const unsigned total = (unsigned)nb * NR;
    for (unsigned idx = threadIdx.x; idx < total; idx += blockDim.x) {
        const unsigned bt = udiv(idx, mNR);
        const unsigned i = idx - bt * NR;
        const unsigned k = i - udiv(i, mP) * (unsigned)p;
        const V2 *x = a + bt * (unsigned)N;
        V2 *y = b + bt * (unsigned)N;
        V2 v[R];
#pragma unroll
        for (int t = 0; t < R; ++t) v[t] = x[i + t * NR];
        if (k > 0) {
#pragma unroll
            for (int t = 1; t < R; ++t) v[t] = cmul(v[t], tw[t * k * tstep]);
        }
        Butterfly<R, V2>::run(v);
        const unsigned j = (i - k) * R + k;
#pragma unroll
        for (int t = 0; t < R; ++t) y[j + t * (unsigned)p] = v[t];
    }
}

// nb transforms of length N stored back to back in a (LDS); b is scratch of
// the same size, tw the N twiddles (LDS).  Returns the buffer holding the
// result (self-sorting Stockham, radices in fd.radix order).
template <typename V2>
__device__ V2 *lds_fft(V2 *a, V2 *b, int nb, const FftDesc &fd, const V2 *tw)
{
    const int N = fd.N;
    int p = 1;
    for (int s = 0; s < fd.nst; ++s) {
        const int R = fd.radix[s];
        const unsigned long long mNR = fd.mNR[s], mP = fd.mP[s];
        switch (R) {
        case 4: fft_stage<4>(a, b, nb, N, p, mNR, mP, tw); break;
        case 2: fft_stage<2>(a, b, nb, N, p, mNR, mP, tw); break;
        case 3: fft_stage<3>(a, b, nb, N, p, mNR, mP, tw); break;
        default: fft_stage<5>(a, b, nb, N, p, mNR, mP, tw); break;
        }
        __syncthreads();
        V2 *tmp = a; a = b; b = tmp;
        p *= R;
    }
    return a;
}

}  // namespace bb
