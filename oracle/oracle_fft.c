/*
 * TEST INFRASTRUCTURE ONLY -- part of the CPU oracle.  Only tests/, the smoke
 * check and the CPU-baseline leg of bench.py may load this code, and only as
 * the checker (or as the timed CPU baseline), never as the product path.
 *
 * Mixed-radix (2,3,4,5) complex FFT in FP64, written from the DFT definition:
 *     X[k] = sum_n x[n] * exp(sign * 2*pi*i * n*k / N)
 * sign = +1 is FFTW_BACKWARD (the transform the reference calls at
 * funcs.c:108, :349, :375, :546; fftw_plan_dft_2d(..., FFTW_BACKWARD, ...)),
 * unnormalised.  FFTW itself is absent from this image, so the oracle pins its
 * FFT to the DFT definition (brute-force DFT at small N, numpy.fft.ifft2 at
 * larger N) rather than to FFTW's rounding.
 *
 * Algorithm: recursive decimation in time.  For N = r*m, the r sub-sequences
 * x[q + r*t] are transformed (length m) into out[q*m + k]; then for each k the
 * r values out[q*m+k] are twiddled by W_N^{q k} and combined by a length-r DFT
 * into X[k + m*s].  Both sets of positions are {k + m*j}, so the combine is
 * done in place.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct { double re, im; } ocplx;

static int factorize(int n, int *f)
{
    int k = 0;
    while (n % 4 == 0) { f[k++] = 4; n /= 4; }
    while (n % 2 == 0) { f[k++] = 2; n /= 2; }
    while (n % 3 == 0) { f[k++] = 3; n /= 3; }
    while (n % 5 == 0) { f[k++] = 5; n /= 5; }
    if (n != 1) return -1;
    return k;
}

/* tw[t] = exp(sign*2*pi*i*t/N), t in [0,N) */
static void make_twiddles(int N, int sign, ocplx *tw)
{
    for (int t = 0; t < N; ++t) {
        /* reduce the angle to the first octant-free form t/N in [0,1) */
        double a = 2.0 * M_PI * (double)t / (double)N;
        tw[t].re = cos(a);
        tw[t].im = sign * sin(a);
    }
}

static void fft_rec(int n, const ocplx *in, long stride, ocplx *out,
                    const int *fac, const ocplx *tw, int N)
{
    if (n == 1) { out[0] = in[0]; return; }
    int r = fac[0], m = n / r;
    for (int q = 0; q < r; ++q)
        fft_rec(m, in + q * stride, stride * r, out + (long)q * m, fac + 1, tw, N);
    int step = N / n;          /* W_n^{e} = tw[e*step] */
    int rstep = N / r;         /* W_r^{e} = tw[e*rstep] */
    ocplx t[5], y[5];
    for (int k = 0; k < m; ++k) {
        for (int q = 0; q < r; ++q) {
            ocplx v = out[(long)q * m + k];
            ocplx w = tw[((long)q * k * step) % N];
            t[q].re = v.re * w.re - v.im * w.im;
            t[q].im = v.re * w.im + v.im * w.re;
        }
        for (int s = 0; s < r; ++s) {
            double re = 0.0, im = 0.0;
            for (int q = 0; q < r; ++q) {
                ocplx w = tw[((long)q * s * rstep) % N];
                re += t[q].re * w.re - t[q].im * w.im;
                im += t[q].re * w.im + t[q].im * w.re;
            }
            y[s].re = re; y[s].im = im;
        }
        for (int s = 0; s < r; ++s) out[k + (long)m * s] = y[s];
    }
}

/* In-place 1-D transform of n complex values (interleaved re,im). */
int orc_fft1d(int n, int sign, double *data)
{
    int fac[64];
    if (n <= 0 || factorize(n, fac) < 0) return -1;
    ocplx *tw = (ocplx *)malloc(sizeof(ocplx) * (size_t)n);
    ocplx *tmp = (ocplx *)malloc(sizeof(ocplx) * (size_t)n);
    if (!tw || !tmp) { free(tw); free(tmp); return -2; }
    make_twiddles(n, sign, tw);
    fft_rec(n, (const ocplx *)data, 1, tmp, fac, tw, n);
    memcpy(data, tmp, sizeof(ocplx) * (size_t)n);
    free(tw); free(tmp);
    return 0;
}

/*
 * In-place 2-D backward (sign +1) unnormalised transform of an NN x NN
 * row-major complex array (interleaved), the transform the reference gets
 * from fftw_plan_dft_2d(NN, NN, D, D, FFTW_BACKWARD, FFTW_ESTIMATE).
 */
int orc_ifft2(int NN, double *data)
{
    int fac[64];
    if (NN <= 0 || factorize(NN, fac) < 0) return -1;
    ocplx *tw = (ocplx *)malloc(sizeof(ocplx) * (size_t)NN);
    ocplx *col = (ocplx *)malloc(sizeof(ocplx) * (size_t)NN);
    ocplx *tmp = (ocplx *)malloc(sizeof(ocplx) * (size_t)NN);
    if (!tw || !col || !tmp) { free(tw); free(col); free(tmp); return -2; }
    make_twiddles(NN, +1, tw);
    ocplx *a = (ocplx *)data;
    for (int i = 0; i < NN; ++i) {
        fft_rec(NN, a + (long)i * NN, 1, tmp, fac, tw, NN);
        memcpy(a + (long)i * NN, tmp, sizeof(ocplx) * (size_t)NN);
    }
    for (int j = 0; j < NN; ++j) {
        for (int i = 0; i < NN; ++i) col[i] = a[(long)i * NN + j];
        fft_rec(NN, col, 1, tmp, fac, tw, NN);
        for (int i = 0; i < NN; ++i) a[(long)i * NN + j] = tmp[i];
    }
    free(tw); free(col); free(tmp);
    return 0;
}
