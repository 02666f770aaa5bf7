// Batched 1-D Stockham FFTs in LDS (radices 4, 2, 3, 5 and composites of two
// of them), shared by the
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

// Length-R DFT with exp(+2 pi i q s / R), in place, for R = 2, 3, 4, 5 (and
// the composites below).
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

// cos and sin of 2 pi m / R, m < R (correctly rounded FP64), for the
// composite butterflies.
template <int R>
struct TwConst;
template <>
struct TwConst<6> {
    static constexpr double c[6] = {1.0, 0.5, -0.5, -1.0, -0.5, 0.5};
    static constexpr double s[6] = {
        0.0, 0.8660254037844386, 0.8660254037844386, 0.0, -0.8660254037844386,
        -0.8660254037844386};
};
template <>
struct TwConst<8> {
    static constexpr double c[8] = {
        1.0, 0.7071067811865476, 0.0, -0.7071067811865476, -1.0, -0.7071067811865476, 0.0,
        0.7071067811865476};
    static constexpr double s[8] = {
        0.0, 0.7071067811865476, 1.0, 0.7071067811865476, 0.0, -0.7071067811865476, -1.0,
        -0.7071067811865476};
};
template <>
struct TwConst<9> {
    static constexpr double c[9] = {
        1.0, 0.766044443118978, 0.17364817766693036, -0.5, -0.9396926207859084,
        -0.9396926207859084, -0.5, 0.17364817766693036, 0.766044443118978};
    static constexpr double s[9] = {
        0.0, 0.6427876096865394, 0.984807753012208, 0.8660254037844386, 0.3420201433256687,
        -0.3420201433256687, -0.8660254037844386, -0.984807753012208, -0.6427876096865394};
};
template <>
struct TwConst<10> {
    static constexpr double c[10] = {
        1.0, 0.8090169943749475, 0.30901699437494745, -0.30901699437494745, -0.8090169943749475,
        -1.0, -0.8090169943749475, -0.30901699437494745, 0.30901699437494745, 0.8090169943749475};
    static constexpr double s[10] = {
        0.0, 0.5877852522924731, 0.9510565162951535, 0.9510565162951535, 0.5877852522924731,
        0.0, -0.5877852522924731, -0.9510565162951535, -0.9510565162951535, -0.5877852522924731};
};
template <>
struct TwConst<12> {
    static constexpr double c[12] = {
        1.0, 0.8660254037844386, 0.5, 0.0, -0.5, -0.8660254037844386, -1.0, -0.8660254037844386,
        -0.5, 0.0, 0.5, 0.8660254037844386};
    static constexpr double s[12] = {
        0.0, 0.5, 0.8660254037844386, 1.0, 0.8660254037844386, 0.5, 0.0, -0.5,
        -0.8660254037844386, -1.0, -0.8660254037844386, -0.5};
};
template <>
struct TwConst<15> {
    static constexpr double c[15] = {
        1.0, 0.9135454576426009, 0.6691306063588582, 0.30901699437494745, -0.10452846326765347,
        -0.5, -0.8090169943749475, -0.9781476007338057, -0.9781476007338057,
        -0.8090169943749475, -0.5, -0.10452846326765347, 0.30901699437494745,
        0.6691306063588582, 0.9135454576426009};
    static constexpr double s[15] = {
        0.0, 0.4067366430758002, 0.7431448254773942, 0.9510565162951535, 0.9945218953682733,
        0.8660254037844386, 0.5877852522924731, 0.20791169081775934, -0.20791169081775934,
        -0.5877852522924731, -0.8660254037844386, -0.9945218953682733, -0.9510565162951535,
        -0.7431448254773942, -0.4067366430758002};
};
template <>
struct TwConst<16> {
    static constexpr double c[16] = {
        1.0, 0.9238795325112867, 0.7071067811865476, 0.3826834323650898, 0.0,
        -0.3826834323650898, -0.7071067811865476, -0.9238795325112867, -1.0,
        -0.9238795325112867, -0.7071067811865476, -0.3826834323650898, 0.0, 0.3826834323650898,
        0.7071067811865476, 0.9238795325112867};
    static constexpr double s[16] = {
        0.0, 0.3826834323650898, 0.7071067811865476, 0.9238795325112867, 1.0,
        0.9238795325112867, 0.7071067811865476, 0.3826834323650898, 0.0, -0.3826834323650898,
        -0.7071067811865476, -0.9238795325112867, -1.0, -0.9238795325112867,
        -0.7071067811865476, -0.3826834323650898};
};

// x * exp(+2 pi i m / R) for a compile-time m: exact for the multiples of
// a quarter turn, a complex product with the rounded constants otherwise.
template <int R, int M, typename V2>
__device__ __forceinline__ V2 twc(V2 x)
{
    using T = decltype(V2::x);
    constexpr int m = M % R;
    if constexpr (m == 0) return x;
    else if constexpr (4 * m == R) return cmuli(x);
    else if constexpr (2 * m == R) return mk<V2>(-x.x, -x.y);
    else if constexpr (4 * m == 3 * R) return mk<V2>(x.y, -x.x);
    else return cmul(x, mk<V2>((T)TwConst<R>::c[m], (T)TwConst<R>::s[m]));
}

// Length R1 R2 DFT in registers (Cooley-Tukey): R2 DFTs of length R1 over
// v[R2 n1 + n2], twiddles exp(+2 pi i n2 k1 / R), R1 DFTs of length R2;
// X[k1 + R1 k2] lands in v[k1 + R1 k2].
template <int R1, int R2, typename V2>
struct Composite {
    static constexpr int R = R1 * R2;
    template <int N2, int K1>
    static __device__ __forceinline__ void tw_row(V2 *t, V2 *y)
    {
        if constexpr (K1 < R1) {
            y[N2 * R1 + K1] = twc<R, N2 * K1>(t[K1]);
            tw_row<N2, K1 + 1>(t, y);
        }
    }
    template <int N2>
    static __device__ __forceinline__ void first(const V2 *v, V2 *y)
    {
        if constexpr (N2 < R2) {
            V2 t[R1];
#pragma unroll
            for (int n1 = 0; n1 < R1; ++n1) t[n1] = v[R2 * n1 + N2];
            Butterfly<R1, V2>::run(t);
            tw_row<N2, 0>(t, y);
            first<N2 + 1>(v, y);
        }
    }
    static __device__ __forceinline__ void run(V2 *v)
    {
        V2 y[R];
        first<0>(v, y);
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) {
            V2 t[R2];
#pragma unroll
            for (int n2 = 0; n2 < R2; ++n2) t[n2] = y[n2 * R1 + k1];
            Butterfly<R2, V2>::run(t);
#pragma unroll
            for (int k2 = 0; k2 < R2; ++k2) v[k1 + R1 * k2] = t[k2];
        }
    }
};
template <typename V2> struct Butterfly<6, V2> : Composite<2, 3, V2> {};
template <typename V2> struct Butterfly<8, V2> : Composite<4, 2, V2> {};
template <typename V2> struct Butterfly<9, V2> : Composite<3, 3, V2> {};
template <typename V2> struct Butterfly<10, V2> : Composite<2, 5, V2> {};
template <typename V2> struct Butterfly<12, V2> : Composite<4, 3, V2> {};
template <typename V2> struct Butterfly<15, V2> : Composite<3, 5, V2> {};
template <typename V2> struct Butterfly<16, V2> : Composite<4, 4, V2> {};

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
// result (self-sorting Stockham, radices in fd.radix order).  With
// kComposite, stages may also have the composite radices 6 to 16 (two
// radix-2..5 stages merged in registers: half the LDS round trips and
// barriers, more registers); without, only 2, 3, 4 and 5 are compiled.
template <typename V2, bool kComposite = false>
__device__ V2 *lds_fft(V2 *a, V2 *b, int nb, const FftDesc &fd, const V2 *tw)
{
    const int N = fd.N;
    int p = 1;
    for (int s = 0; s < fd.nst; ++s) {
        const int R = fd.radix[s];
        const unsigned long long mNR = fd.mNR[s], mP = fd.mP[s];
        if constexpr (kComposite) {
            switch (R) {
            case 4: fft_stage<4>(a, b, nb, N, p, mNR, mP, tw); break;
            case 2: fft_stage<2>(a, b, nb, N, p, mNR, mP, tw); break;
            case 3: fft_stage<3>(a, b, nb, N, p, mNR, mP, tw); break;
            case 5: fft_stage<5>(a, b, nb, N, p, mNR, mP, tw); break;
            case 6: fft_stage<6>(a, b, nb, N, p, mNR, mP, tw); break;
            case 8: fft_stage<8>(a, b, nb, N, p, mNR, mP, tw); break;
            case 9: fft_stage<9>(a, b, nb, N, p, mNR, mP, tw); break;
            case 10: fft_stage<10>(a, b, nb, N, p, mNR, mP, tw); break;
            case 12: fft_stage<12>(a, b, nb, N, p, mNR, mP, tw); break;
            case 15: fft_stage<15>(a, b, nb, N, p, mNR, mP, tw); break;
            default: fft_stage<16>(a, b, nb, N, p, mNR, mP, tw); break;
            }
        } else {
            switch (R) {
            case 4: fft_stage<4>(a, b, nb, N, p, mNR, mP, tw); break;
            case 2: fft_stage<2>(a, b, nb, N, p, mNR, mP, tw); break;
            case 3: fft_stage<3>(a, b, nb, N, p, mNR, mP, tw); break;
            default: fft_stage<5>(a, b, nb, N, p, mNR, mP, tw); break;
            }
        }
        __syncthreads();
        V2 *tmp = a; a = b; b = tmp;
        p *= R;
    }
    return a;
}

}  // namespace bb
