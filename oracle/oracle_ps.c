/*
 * TEST INFRASTRUCTURE ONLY -- CPU oracle of the power-spectrum multipoles
 * program `calpow_mom` (calpow_mom.c:22-117, calpow_mom in funcs.c:860-1097).
 * Only tests/ and the CPU-baseline leg of bench.py may load it, as the
 * checker.
 *
 * A restatement that keeps the reference's arithmetic types: the mode
 * modulus m, mu, the Legendre weights P2/P4/P6 and the log-bin scale are
 * `float` (funcs.c:868), the accumulators `double`, and the expressions are
 * evaluated in the reference's order (float * long products, the double
 * sqrt and log10 of float arguments, floorf of the double quotient).  The
 * three loops visit the r2c modes as the reference does: half lines
 * (funcs.c:926-958), half planes (:962-1001), half cube (:1005-1049).
 *
 * out: 6 * Nbin doubles, per bin {kmean, P0, P2, P4, P6, num} after the
 * reference's normalisation (funcs.c:1055-1071), i.e. the columns of
 * psoutN<NN>_R<NRi>.dat (calpow_mom.c:104-113).
 */
#include <math.h>
#include <stdlib.h>

static void ps_add(double *acc, double re, double im, float m, float P2, float P4, float P6)
{
    const double p = (double)((re * re) + (im * im));
    acc[0] += p;
    acc[1] += (double)(((re * re) + (im * im)) * P2);
    acc[2] += (double)(((re * re) + (im * im)) * P4);
    acc[3] += (double)(((re * re) + (im * im)) * P6);
    acc[4] += (double)(m);
    acc[5] += 1.0;
    acc[6] += (double)(P2 * P2);
    acc[7] += (double)(P4 * P4);
    acc[8] += (double)(P6 * P6);
}

/* Mode (a, b, c) of the r2c buffer at `index`: weights and log bin as in
 * funcs.c:932-945. */
static void ps_mode(const double *field, long N, float fac, float scale, int Nbin, long a, long b,
                    long c, long index, double *acc)
{
    float m, mu, P2, P4, P6;
    m = sqrt(fac * a * a + fac * b * b + fac * c * c);
    mu = (1.0 * c) / (N * m);
    P2 = 0.5 * (3. * mu * mu - 1.);
    P4 = 0.125 * (35.0 * powf(mu, 4.0) - 30.0 * mu * mu + 3);
    P6 = 0.0625 * (231.0 * powf(mu, 6.0) - 315.0 * powf(mu, 4.0) + 105.0 * mu * mu - 5);
    const int d = (int)floorf(log10(m * N) / scale);
    if (d >= 0 && d < Nbin)
        ps_add(acc + 9 * d, field[2 * index], field[2 * index + 1], m, P2, P4, P6);
}

int orc_calpow_mom(int NN, const double *field, int Nbin, double LL, double *out)
{
    const long N = NN;
    double *acc = calloc((size_t)Nbin * 9, sizeof(double));
    if (!acc) return -1;
    const float fac = 1. / (1. * N * N);
    const float scale = log10(0.5 * N) / Nbin;
    const double norml = 1. / (1. * N * N * N);
    const long H = N / 2 + 1;
    /* half lines */
    for (long i = 1; i <= N / 2; i++)
        for (long j = 0; j <= N / 2; j = j + N / 2)
            for (long k = 0; k <= N / 2; k = k + N / 2)
                ps_mode(field, N, fac, scale, Nbin, i, j, k, i * N * H + j * H + k, acc);
    /* half planes */
    for (long i = 0; i < N; i++) {
        const long a = (i > N / 2) ? N - i : i;
        for (long j = 1; j < N / 2; j++)
            for (long k = 0; k <= N / 2; k = k + N / 2)
                ps_mode(field, N, fac, scale, Nbin, a, j, k, i * N * H + j * H + k, acc);
    }
    /* half cube */
    for (long i = 0; i < N; i++) {
        const long a = (i > N / 2) ? N - i : i;
        for (long j = 0; j < N; j++) {
            const long b = (j > N / 2) ? N - j : j;
            for (long k = 1; k < N / 2; k++)
                ps_mode(field, N, fac, scale, Nbin, a, b, k, i * N * H + j * H + k, acc);
        }
    }
    const double pi = M_PI;
    for (int i = 0; i < Nbin; i++) {
        double *A = acc + 9 * i, *o = out + 6 * i;
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
    free(acc);
    return 0;
}
