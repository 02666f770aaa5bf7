/*
 * TEST INFRASTRUCTURE ONLY -- part of the CPU oracle.  Only tests/ and the
 * smoke check may load this code, and only as the checker.
 *
 * CPU restatement of the seeded field generator (the GPU's
 * bb_generate_field), which follows the reference's generator
 * (generate_field.c:22-107; Fill_Modes, Make_Image, Non_Gauss, Get_Modes,
 * RSD of ngrf_2d_funcs.c:341-577, 734-765) with a counter-based random
 * stream instead of GSL's mt19937 + ziggurat (absent here).
 *
 * Pins:
 *  - orc_philox4x64: Philox4x64-10 (Salmon, Moraes, Dror and Shaw, SC 2011)
 *    equals numpy.random.Philox's raw output bit for bit at many counters
 *    and keys (tests/test_field_oracle.py);
 *  - orc_gen_log / orc_gen_sincos_2pi: within a few ulp of the C library's
 *    log, sin and cos over many arguments (the same test file); the normal
 *    pairs are standard normal (moments, a two-sided KS test);
 *  - orc_generate_field's transforms use the oracle's FFT (oracle_fft.c,
 *    pinned to the DFT definition); the field's power follows V P(k) and
 *    delta is real (tests).
 * Floating-point steps use explicit fma() and no contraction
 * (-ffp-contract=off), so that the GPU's identical sequence gives the same
 * bits for the Gaussian modes.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

int orc_fft1d(int n, int sign, double *data);

typedef unsigned long long u64;

void orc_philox4x64(const u64 ctr[4], const u64 key[2], u64 out[4])
{
    const u64 M0 = 0xD2E7470EE14C6C93ULL, M1 = 0xCA5A826395121157ULL;
    const u64 W0 = 0x9E3779B97F4A7C15ULL, W1 = 0xBB67AE8584CAA73BULL;
    u64 c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        const unsigned __int128 p0 = (unsigned __int128)M0 * c0;
        const unsigned __int128 p1 = (unsigned __int128)M1 * c2;
        const u64 hi0 = (u64)(p0 >> 64), lo0 = (u64)p0, hi1 = (u64)(p1 >> 64), lo1 = (u64)p1;
        const u64 n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += W0;
        k1 += W1;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* ln x, x in (0, 1]: 2 atanh((m - 1)/(m + 1)) series, m in [sqrt(1/2), sqrt(2)). */
double orc_gen_log(double x)
{
    int e;
    double m = frexp(x, &e);
    if (m < 0.70710678118654752440) {
        m = m * 2.0;
        e = e - 1;
    }
    const double f = m - 1.0;
    const double s = f / (m + 1.0);
    const double z = s * s;
    double p = 1.0 / 23.0;
    p = fma(p, z, 1.0 / 21.0);
    p = fma(p, z, 1.0 / 19.0);
    p = fma(p, z, 1.0 / 17.0);
    p = fma(p, z, 1.0 / 15.0);
    p = fma(p, z, 1.0 / 13.0);
    p = fma(p, z, 1.0 / 11.0);
    p = fma(p, z, 1.0 / 9.0);
    p = fma(p, z, 1.0 / 7.0);
    p = fma(p, z, 1.0 / 5.0);
    p = fma(p, z, 1.0 / 3.0);
    const double lm = fma(2.0 * s * z, p, 2.0 * s);
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    return fma((double)e, ln2_hi, fma((double)e, ln2_lo, lm));
}

static void sincos_small(double a, double *s, double *c)
{
    const double z = a * a;
    double ps = -1.0 / 51090942171709440000.0;
    ps = fma(ps, z, 1.0 / 121645100408832000.0);
    ps = fma(ps, z, -1.0 / 355687428096000.0);
    ps = fma(ps, z, 1.0 / 1307674368000.0);
    ps = fma(ps, z, -1.0 / 6227020800.0);
    ps = fma(ps, z, 1.0 / 39916800.0);
    ps = fma(ps, z, -1.0 / 362880.0);
    ps = fma(ps, z, 1.0 / 5040.0);
    ps = fma(ps, z, -1.0 / 120.0);
    ps = fma(ps, z, 1.0 / 6.0);
    *s = fma(-a * z, ps, a);
    double pc = 1.0 / 2432902008176640000.0;
    pc = fma(pc, z, -1.0 / 6402373705728000.0);
    pc = fma(pc, z, 1.0 / 20922789888000.0);
    pc = fma(pc, z, -1.0 / 87178291200.0);
    pc = fma(pc, z, 1.0 / 479001600.0);
    pc = fma(pc, z, -1.0 / 3628800.0);
    pc = fma(pc, z, 1.0 / 40320.0);
    pc = fma(pc, z, -1.0 / 720.0);
    pc = fma(pc, z, 1.0 / 24.0);
    pc = fma(pc, z, -0.5);
    *c = fma(z, pc, 1.0);
}

/* sin and cos of 2 pi u, u in [0, 1). */
void orc_gen_sincos_2pi(double u, double *s, double *c)
{
    const double t = 4.0 * u;
    const double qd = floor(t);
    const int q = (int)qd;
    const double f = t - qd;
    const double pio2_hi = 1.57079632679489655800e+00;
    const double pio2_lo = 6.12323399573676603587e-17;
    double ss, cc;
    if (f > 0.5) {
        const double g = 1.0 - f;
        sincos_small(fma(g, pio2_hi, g * pio2_lo), &cc, &ss);
    } else {
        sincos_small(fma(f, pio2_hi, f * pio2_lo), &ss, &cc);
    }
    switch (q & 3) {
    case 0: *s = ss; *c = cc; break;
    case 1: *s = cc; *c = -ss; break;
    case 2: *s = -ss; *c = -cc; break;
    default: *s = -cc; *c = ss; break;
    }
}

/* The standard normal pair of counter c and key (seed, 0). */
void orc_gen_normal_pair(u64 c, u64 seed, double *g)
{
    const u64 ctr[4] = {c, 0, 0, 0}, key[2] = {seed, 0};
    u64 w[4];
    orc_philox4x64(ctr, key, w);
    const double u1 = (double)(w[0] >> 11) * (1.0 / 9007199254740992.0);
    const double u2 = (double)(w[1] >> 11) * (1.0 / 9007199254740992.0);
    const double r = sqrt(-2.0 * orc_gen_log(1.0 - u1));
    double s, co;
    orc_gen_sincos_2pi(u2, &s, &co);
    g[0] = r * co;
    g[1] = r * s;
}

typedef struct {
    double amp, index, kunit, kcut, fnl, beta, LL;
    int stage;
} orc_field_params;

/* The Gaussian modes (Fill_Modes, ngrf_2d_funcs.c:341-505) into the r2c
 * buffer A ([NN][NN][NN/2+1] complex), each divided by vol. */
static void fill_modes(int NN, u64 seed, const orc_field_params *P, double vol, double *A)
{
    const int H = NN / 2 + 1;
    const double V = pow(NN * P->LL, 3.0);
    const long nk2 = 3L * (NN / 2) * (NN / 2) + 1;
    double *amp = (double *)calloc((size_t)nk2, sizeof(double));
    for (long k2 = 1; k2 < nk2; ++k2) {
        const double k = P->kunit * sqrt((double)k2);
        if (P->kcut > 0 && !(k < P->kcut)) continue;
        amp[k2] = sqrt(V / 2. * (P->amp * pow(k, P->index)));
    }
    const long n = (long)NN * NN * H;
    for (long c = 0; c < n; ++c) {
        const int k = (int)(c % H);
        const long ij = c / H;
        const int j = (int)(ij % NN), i = (int)(ij / NN);
        const long ia = i <= NN / 2 ? i : i - NN, ja = j <= NN / 2 ? j : j - NN;
        const double a = amp[ia * ia + ja * ja + (long)k * k];
        double g[2], re, im;
        if (k == 0 || k == NN / 2) {
            const int im_ = (NN - i) % NN, jm = (NN - j) % NN;
            const long cm = ((long)im_ * NN + jm) * H + k;
            if (cm == c) {
                orc_gen_normal_pair((u64)c, seed, g);
                re = a * g[0];
                im = 0.0;
            } else if (c < cm) {
                orc_gen_normal_pair((u64)c, seed, g);
                re = a * g[0];
                im = a * g[1];
            } else {
                orc_gen_normal_pair((u64)cm, seed, g);
                re = a * g[0];
                im = -(a * g[1]);
            }
        } else {
            orc_gen_normal_pair((u64)c, seed, g);
            re = a * g[0];
            im = a * g[1];
        }
        A[2 * c] = re / vol;
        A[2 * c + 1] = im / vol;
    }
    free(amp);
}

/* 1-D transforms along axis `ax` (0: i, 1: j) of the complex [NN][NN][H]
 * array, sign +1 (backward) or -1 (forward). */
static void axis_fft(int NN, double *A, int ax, int sign)
{
    const int H = NN / 2 + 1;
    double *line = (double *)malloc(sizeof(double) * 2 * (size_t)NN);
    for (int o = 0; o < NN; ++o)
        for (int k = 0; k < H; ++k) {
            for (int e = 0; e < NN; ++e) {
                const long idx = ax == 0 ? ((long)e * NN + o) * H + k : ((long)o * NN + e) * H + k;
                line[2 * e] = A[2 * idx];
                line[2 * e + 1] = A[2 * idx + 1];
            }
            orc_fft1d(NN, sign, line);
            for (int e = 0; e < NN; ++e) {
                const long idx = ax == 0 ? ((long)e * NN + o) * H + k : ((long)o * NN + e) * H + k;
                A[2 * idx] = line[2 * e];
                A[2 * idx + 1] = line[2 * e + 1];
            }
        }
    free(line);
}

/* The whole generator into out ([NN][NN][NN+2] doubles, FFTW's r2c layout);
 * stage 1: the Gaussian modes only (not divided by V). */
int orc_generate_field(int NN, u64 seed, const orc_field_params *P, double *out)
{
    const int H = NN / 2 + 1;
    const long nrows = (long)NN * NN;
    const double V = pow(NN * P->LL, 3.0);
    fill_modes(NN, seed, P, P->stage == 1 ? 1.0 : V, out);
    if (P->stage == 1) return 0;
    axis_fft(NN, out, 0, +1);
    axis_fft(NN, out, 1, +1);
    double *row = (double *)malloc(sizeof(double) * 2 * (size_t)NN);
    for (long r = 0; r < nrows; ++r) {   /* c2r along k: Hermitian extension, real part */
        double *x = out + r * (NN + 2);
        for (int k = 0; k < NN; ++k) {
            const int kk = k < H ? k : NN - k;
            row[2 * k] = x[2 * kk];
            row[2 * k + 1] = k < H ? x[2 * kk + 1] : -x[2 * kk + 1];
        }
        orc_fft1d(NN, +1, row);
        for (int k = 0; k < NN; ++k) x[k] = row[2 * k];
    }
    /* Non_Gauss, ngrf_2d_funcs.c:553-577 */
    double var = 0.0;
    for (long r = 0; r < nrows; ++r)
        for (int k = 0; k < NN; ++k) var += out[r * (NN + 2) + k] * out[r * (NN + 2) + k];
    var /= ((double)NN * NN * NN);
    for (long r = 0; r < nrows; ++r)
        for (int k = 0; k < NN; ++k) {
            const double v = out[r * (NN + 2) + k];
            out[r * (NN + 2) + k] = v + P->fnl * (v * v - var);
        }
    /* Get_Modes: times LL^3, r2c along k, then j and i */
    const double L3 = pow(P->LL, 3.);
    for (long r = 0; r < nrows; ++r) {
        double *x = out + r * (NN + 2);
        for (int k = 0; k < NN; ++k) {
            row[2 * k] = x[k] * L3;
            row[2 * k + 1] = 0.0;
        }
        orc_fft1d(NN, -1, row);
        for (int k = 0; k < H; ++k) {
            x[2 * k] = row[2 * k];
            x[2 * k + 1] = row[2 * k + 1];
        }
    }
    free(row);
    axis_fft(NN, out, 1, -1);
    axis_fft(NN, out, 0, -1);
    /* RSD, ngrf_2d_funcs.c:749-760 */
    for (long c = 1; c < nrows * H; ++c) {
        const int k = (int)(c % H);
        const long ij = c / H;
        const int j = (int)(ij % NN), i = (int)(ij / NN);
        const double ia = i > NN / 2 ? NN - i : i, ja = j > NN / 2 ? NN - j : j;
        const double kk = sqrt(1. * (ia * ia + ja * ja + (double)k * k));
        const double myu = k / kk;
        const double f = 1. + P->beta * myu * myu;
        out[2 * c] *= f;
        out[2 * c + 1] *= f;
    }
    return 0;
}
