/*
 * TEST INFRASTRUCTURE ONLY -- the CPU oracle of the bispectrum-multipole
 * estimator.  Only tests/, the smoke check and the CPU-baseline leg of
 * bench.py may load this library, and only as the checker (or as the timed
 * CPU baseline), never as the product path.
 *
 * A plain FP64 restatement of the reference program `barbie` (barbie.c) and
 * of the support routines it calls (funcs.c).  Each function cites the
 * reference lines it follows; loop order, enumeration order, floating-point
 * expression order and the exact-zero rule are kept as in the reference so
 * that, given the same plane-ring fields, the per-contraction sums and the
 * bin accumulations reproduce the reference's arithmetic.
 *
 * Differences from the reference, all deliberate:
 *  - FFTW is replaced by orc_ifft2 (oracle_fft.c), pinned to the DFT
 *    definition; FFTW's own rounding is not reproduced.
 *  - The plane-ring cache is held in memory (lazily built per field) instead
 *    of being written to and re-read from ResFiles_N<NN>/data_files/.
 *  - bin_Biana is zero-initialised (the reference leaves it uninitialised,
 *    docs/REVIEW.md:7), so output column 5 is defined.
 *  - For Nr not a multiple of ten the piecewise ring rule is truncated at Nr
 *    (the reference loop at barbie.c:111-114 writes past the array).
 *
 * Parity status of each function (how it is pinned) is stated next to it.
 * orc_input_ps, orc_sort and orc_cal_coeff are also compared bit for bit
 * with the reference's own functions, compiled from its funcs.c into
 * oracle/_ref by `make ref` (the rest of the reference needs FFTW and GSL).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int orc_ifft2(int NN, double *data);

#define NPOLES 16

typedef struct {
    int NN, Nmax, Nr, n_mtbin;
    int UL_k1, LL_k1, BW_k1;
    double LL;      /* cell size, barbie.c:67 */
    double ff;      /* analytic-comparison amplitude, barbie.c:79 */
    double beta;    /* Kaiser factor for the analytic comparison, barbie.c:80 */
} orc_params;

/* ---------------------------------------------------------------------- */
/* Ring edges.  barbie.c:110-130.  Pinned by closed-form values in tests.   */
int orc_edges_reference(int Nr, int Nmax, double *low, double *high, double *avg)
{
    if (Nr <= 0) return -1;
    int current = 0;
    for (int i = 0; i < Nr; i += 10)
        for (int j = i; j < i + 10 && j < Nr; ++j) {
            low[j] = current;
            current += (i / 10) + 1;
        }
    for (int i = 1; i < Nr; ++i) {
        low[i] -= .1;
        high[i - 1] = low[i];
    }
    high[Nr - 1] = 1. * Nmax;
    for (int i = 0; i < Nr; ++i) avg[i] = (high[i] + low[i]) / 2.;
    return 0;
}

/* Uniform-width variant of the same rule (block width fixed to `width`).
 * last_to_nmax = 1 closes the last ring at Nmax as barbie.c:120 does;
 * 0 closes it at its own width. */
int orc_edges_uniform(int Nr, double width, int Nmax, int last_to_nmax,
                      double *low, double *high, double *avg)
{
    if (Nr <= 0 || !(width > 0)) return -1;
    double current = 0.;
    for (int i = 0; i < Nr; ++i) { low[i] = current; current += width; }
    for (int i = 1; i < Nr; ++i) {
        low[i] -= .1;
        high[i - 1] = low[i];
    }
    high[Nr - 1] = last_to_nmax ? 1. * Nmax : current - .1;
    for (int i = 0; i < Nr; ++i) avg[i] = (high[i] + low[i]) / 2.;
    return 0;
}

/* Ring predicate, barbie.c:141-146 / funcs.c:248-255. */
static int in_ring(int NN, int i, int j, double lo, double hi)
{
    int ia = (i > NN / 2) ? (i - NN) : i;
    int ja = (j > NN / 2) ? (j - NN) : j;
    double kk = sqrt(1. * (ia * ia + ja * ja));
    return (kk >= lo && kk < hi);
}

/* ---------------------------------------------------------------------- */
/* calCR, funcs.c:328-424: ring triples and Num = sum_x I1 I2 I3.
 * Pinned in tests: Num/NN^2 equals a brute-force lattice-triangle count,
 * and rcount equals the enumeration counts of the reference geometry.
 * Returns rcount (or -1 if more than cap triples). rdata: rcount*4 doubles. */
int orc_calcr(int NN, int Nr, const double *low, const double *high,
              const double *avg, double *rdata, int cap)
{
    long NN2 = (long)NN * NN;
    double *II = (double *)malloc(sizeof(double) * (size_t)Nr * NN2);
    double *AA = (double *)malloc(sizeof(double) * 2 * (size_t)NN2);
    if (!II || !AA) { free(II); free(AA); return -2; }
    for (int r = 0; r < Nr; ++r)
        for (int i = 0; i < NN; ++i)
            for (int j = 0; j < NN; ++j)
                II[r * NN2 + (long)i * NN + j] = in_ring(NN, i, j, low[r], high[r]) ? 1. : 0.;
    for (int r = 0; r < Nr; ++r) {
        for (long x = 0; x < NN2; ++x) { AA[2 * x] = II[r * NN2 + x]; AA[2 * x + 1] = 0.; }
        orc_ifft2(NN, AA);
        for (long x = 0; x < NN2; ++x) II[r * NN2 + x] = AA[2 * x];
    }
    int rcount = 0;
    for (int r1 = 0; r1 < Nr; ++r1)
        for (int r2 = 0; r2 <= r1; ++r2) {
            double tt = avg[r2] / avg[r1];
            if (tt >= 0.5 && tt <= 1.) {
                for (int r3 = 0; r3 <= r2; ++r3) {
                    double mu = 0.5 * (tt + 1. / tt - (avg[r3] / avg[r1]) * (avg[r3] / avg[r2]));
                    if (mu >= 0.5 / tt && mu <= 1.) {
                        double Num = 0.;
                        const double *a = II + r1 * NN2, *b = II + r2 * NN2, *c = II + r3 * NN2;
                        for (long x = 0; x < NN2; ++x) Num += a[x] * b[x] * c[x];
                        if (rcount >= cap) { free(II); free(AA); return -1; }
                        rdata[4 * rcount + 0] = r1;
                        rdata[4 * rcount + 1] = r2;
                        rdata[4 * rcount + 2] = r3;
                        rdata[4 * rcount + 3] = Num;
                        rcount++;
                    }
                }
            }
        }
    free(II); free(AA);
    return rcount;
}

/* ---------------------------------------------------------------------- */
/* Df, funcs.c:88-109 with sign = +1 as called from wrapid_fast
 * (funcs.c:553): gather plane kz=plane of the r2c buffer, apply the ring
 * mask, zero the DC mode of plane 0, and apply the 2-D backward FFT.
 * field: FFTW in-place r2c buffer [NN][NN][NN+2] doubles.
 * out: NN*NN complex (interleaved).  Pinned in tests against numpy ifft2 and
 * a brute-force DFT. */
void orc_df(int NN, const double *field, int plane, double lo, double hi, double *out)
{
    long H = NN / 2 + 1;
    for (int i = 0; i < NN; ++i)
        for (int j = 0; j < NN; ++j) {
            long index_1 = (long)i * NN * H + (long)j * H + plane;
            long index_2 = (long)i * NN + j;
            double m = in_ring(NN, i, j, lo, hi) ? 1. : 0.;
            out[2 * index_2] = field[2 * index_1] * m;
            out[2 * index_2 + 1] = field[2 * index_1 + 1] * m;
        }
    if (plane == 0) { out[0] = 0.; out[1] = 0.; }
    orc_ifft2(NN, out);
}

/* sort, funcs.c:113-126: descending bubble sort, strict '<' (ties keep
 * input order). */
void orc_sort(double K[3][2])
{
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2 - i; j++)
            if (K[j][0] < K[j + 1][0]) {
                double s1 = K[j][0], s2 = K[j][1];
                K[j][0] = K[j + 1][0]; K[j][1] = K[j + 1][1];
                K[j + 1][0] = s1; K[j + 1][1] = s2;
            }
}

/* cal_coeff, funcs.c:708-745.  Pinned in tests: the m=0 entries against
 * Legendre polynomials, the m!=0 entries against associated Legendre
 * functions times cos(m*phi), norm[] against (P_l^m)^2, each up to a
 * constant fixed at special points. */
void orc_cal_coeff(double px, double pz, double *bs, double *norm)
{
    bs[0] = 1.;
    bs[1] = 2. * (3. * pz * pz - 1.);
    bs[2] = -sqrt(2. / 3.) * px * pz;
    bs[3] = 2. * sqrt(0.6666666666666666) * (-1. + 2. * pow(px, 2) + pow(pz, 2));
    bs[4] = 8. * (3. - 30. * pow(pz, 2) + 35. * pow(pz, 4));
    bs[5] = (-4. * px * pz * (-3. + 7. * pow(pz, 2))) / sqrt(5.);
    bs[6] = 4. * sqrt(0.4) * (-1 + 2. * pow(px, 2) + pow(pz, 2)) * (-1. + 7. * pow(pz, 2));
    bs[7] = (-4. * px * pz * (-3. + 4. * pow(px, 2) + 3. * pow(pz, 2))) / sqrt(35.);
    bs[8] = 8. * sqrt(0.05714285714285714) * (8. * pow(px, 4) + 8. * pow(px, 2) * (-1. + pow(pz, 2)) + pow(-1. + pow(pz, 2), 2));
    bs[9] = 16. * (-5. + 105. * pow(pz, 2) - 315. * pow(pz, 4) + 231. * pow(pz, 6));
    bs[10] = -8. * sqrt(0.09523809523809523) * px * pz * (5. - 30. * pow(pz, 2) + 33. * pow(pz, 4));
    bs[11] = (32. * (-1. + 2. * pow(px, 2) + pow(pz, 2)) * (1. - 18. * pow(pz, 2) + 33. * pow(pz, 4))) / sqrt(105.0);
    bs[12] = (-16. * px * pz * (-3. + 4. * pow(px, 2) + 3. * pow(pz, 2)) * (-3. + 11. * pow(pz, 2))) / sqrt(105.0);
    bs[13] = (16. * sqrt(0.2857142857142857) * (-1. + 11. * pow(pz, 2)) * (8. * pow(px, 4) + 8. * pow(px, 2) * (-1. + pow(pz, 2)) + pow(-1. + pow(pz, 2), 2))) / 3.;
    bs[14] = (-16 * pz * (16 * pow(px, 5) + 20 * pow(px, 3) * (-1 + pow(pz, 2)) + 5 * px * pow(-1 + pow(pz, 2), 2))) / (3. * sqrt(77.0));
    bs[15] = (32. * (32. * pow(px, 6) + 48. * pow(px, 4) * (-1. + pow(pz, 2)) + 18. * pow(px, 2) * pow(-1. + pow(pz, 2), 2) + pow(-1. + pow(pz, 2), 3.))) / sqrt(231.);

    norm[0] = 1.;
    norm[1] = ((3. * pz * pz - 1.)) * ((3. * pz * pz - 1.));
    norm[2] = pz * pz * (1. - pz * pz);
    norm[3] = pow(-1. + pow(pz, 2), 2);
    norm[4] = pow(3. - 30. * pow(pz, 2) + 35. * pow(pz, 4), 2);
    norm[5] = (1. - pow(pz, 2)) * pow(-3. * pz + 7. * pow(pz, 3), 2);
    norm[6] = pow(-1. + pow(pz, 2), 2) * pow(-1. + 7. * pow(pz, 2), 2);
    norm[7] = pow(pz, 2) * pow(1. - pow(pz, 2), 3);
    norm[8] = pow(-1. + pow(pz, 2), 4);
    norm[9] = pow(-5. + 105. * pow(pz, 2) - 315. * pow(pz, 4) + 231. * pow(pz, 6), 2);
    norm[10] = (1. - pow(pz, 2)) * pow(5. * pz - 30. * pow(pz, 3) + 33. * pow(pz, 5), 2);
    norm[11] = pow(-1. + pow(pz, 2), 2) * pow(1. - 18. * pow(pz, 2) + 33. * pow(pz, 4), 2);
    norm[12] = pow(1. - pow(pz, 2), 3) * pow(-3. * pz + 11. * pow(pz, 3), 2);
    norm[13] = pow(-1 + pow(pz, 2), 4) * pow(-1 + 11 * pow(pz, 2), 2);
    norm[14] = pow(pz, 2) * pow(1 - pow(pz, 2), 5);
    norm[15] = pow(-1. + pow(pz, 2), 6);
}

/* input_ps, funcs.c:73-81.  Pinned bit for bit against the reference's own
 * function (tests/test_golden.py). */
double orc_input_ps(double kk) { return 1.0 * pow(kk, -2.0); }
#define input_ps orc_input_ps

/* Triangle parameters of a sorted K[3][2] = {(|k|, k_z)}, barbie.c:293-297:
 * out = {K1, T, MU, pz, px}.  Pinned by closed 3-D triangles: MU is minus
 * the cosine between the two largest sides, pz the z cosine of the largest
 * and px the z component of the unit vector in the triangle's plane
 * perpendicular to the largest side (tests/test_golden.py). */
void orc_triangle_params(const double K[3][2], double *out)
{
    double T = K[1][0] / K[0][0];
    double MU = 0.5 * (T + 1. / T - (K[2][0] / K[0][0]) * (K[2][0] / K[1][0]));
    double pz = K[0][1] / K[0][0];
    out[0] = K[0][0];
    out[1] = T;
    out[2] = MU;
    out[3] = pz;
    out[4] = ((K[1][1] / K[1][0]) + MU * pz) / (sqrt(1. - (MU * MU)));
}

/* ---------------------------------------------------------------------- */
/* The estimator loop, barbie.c:203-373.
 *
 * Bins are [n_k1bin][n_mtbin][n_mtbin][16] doubles, accumulated (not
 * zeroed) by this call so that p1 ranges can be split over calls.
 * p1 in [p1_begin, p1_end).  nthreads > 1 parallelises the rp loop of each
 * p1 with per-thread bins merged in thread order (this changes the
 * summation order; nthreads = 1 keeps the reference's order).
 * contr_out (optional): per contraction, in the reference's enumeration
 * order, 8 doubles {p1, r1, p2, r2, p3, r3, Bispec_real, flat bin index}
 * (threaded runs place each ring triple's records at offsets counted by the
 * serial pre-pass, so the order and the values are those of a serial run).
 * Returns the number of contractions (K1 inside (LL_k1, UL_k1)) visited, or
 * a negative error. */
typedef struct {
    int NN, Nr;
    const double *field, *low, *high;
    double **cache;     /* [Nmax*Nr] lazily built fields, 2*NN*NN doubles */
    const char *dir;    /* on-disk mode: dataPlaneRing_* files are read from here */
    int NRi;
    double *buf[3];     /* on-disk mode: buffer1..3 of barbie.c:274-324 */
} orc_cache;

static const double *cache_get(orc_cache *c, int p, int r)
{
    double **slot = &c->cache[(long)p * c->Nr + r];
    if (!*slot) {
        double *d = (double *)malloc(sizeof(double) * 2 * (size_t)c->NN * c->NN);
        if (!d) return NULL;
        orc_df(c->NN, c->field, p, c->low[r], c->high[r], d);
        *slot = d;
    }
    return *slot;
}

/* A field of the contraction.  In-memory mode: the lazily built cache.
 * On-disk mode: the reference's access pattern (barbie.c:274-277 for
 * buffer1 once per ring triple, :315-324 for buffer2 and buffer3 once per
 * contraction): open, read 2 NN^2 doubles, close, into buffer `slot`. */
static const double *field_at(orc_cache *c, int p, int r, int slot)
{
    if (!c->dir) return cache_get(c, p, r);
    char name[1024];
    snprintf(name, sizeof name, "%s/dataPlaneRing_%d_%d_%d_R%d", c->dir, c->NN, p, r, c->NRi);
    FILE *f = fopen(name, "rb");
    if (!f) return NULL;
    const size_t n = 2 * (size_t)c->NN * c->NN;
    const size_t got = fread(c->buf[slot], sizeof(double), n, f);
    fclose(f);
    return got == n ? c->buf[slot] : NULL;
}

/* wrapid_fast (funcs.c:537-566): D_{p,r} of every plane p < Nmax and ring
 * r < Nr (those with want[p*Nr+r] != 0 when want is not NULL) written to
 * <dir>/dataPlaneRing_<NN>_<p>_<r>_R<NRi> as 2 NN^2 doubles, row-major
 * interleaved complex.  The transforms run on nthreads threads (the
 * reference runs FFTW with 4, barbie.c:68,151-153).  Returns the number of
 * files written, or a negative error. */
long orc_write_plane_cache(int NN, const double *field, int Nmax, int Nr, const double *low,
                           const double *high, const char *dir, int NRi, int nthreads,
                           const unsigned char *want)
{
    long written = 0, bad = 0;
    if (nthreads < 1) nthreads = 1;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1) reduction(+:written,bad)
#endif
    for (long k = 0; k < (long)Nmax * Nr; ++k) {
        if (want && !want[k]) continue;
        const int p = (int)(k / Nr), r = (int)(k % Nr);
        double *d = (double *)malloc(sizeof(double) * 2 * (size_t)NN * NN);
        if (!d) { bad++; continue; }
        orc_df(NN, field, p, low[r], high[r], d);
        char name[1024];
        snprintf(name, sizeof name, "%s/dataPlaneRing_%d_%d_%d_R%d", dir, NN, p, r, NRi);
        FILE *f = fopen(name, "wb");
        if (!f || fwrite(d, sizeof(double), 2 * (size_t)NN * NN, f) != 2 * (size_t)NN * NN) bad++;
        else written++;
        if (f) fclose(f);
        free(d);
    }
    return bad ? -1 : written;
}

static long estimate_impl(const orc_params *P, const double *low, const double *high,
                          const double *avg, int rcount, const double *rdata,
                          const double *field, const char *dir, int NRi, int p1_begin,
                          int p1_end, int nthreads, double *bin_bs, double *bin_norm,
                          double *bin_biana, double *contr_out, long contr_cap);

long orc_estimate(const orc_params *P, const double *low, const double *high,
                  const double *avg, int rcount, const double *rdata,
                  const double *field, int p1_begin, int p1_end, int nthreads,
                  double *bin_bs, double *bin_norm, double *bin_biana,
                  double *contr_out, long contr_cap)
{
    return estimate_impl(P, low, high, avg, rcount, rdata, field, NULL, 0, p1_begin, p1_end,
                         nthreads, bin_bs, bin_norm, bin_biana, contr_out, contr_cap);
}

/* The same loop reading its fields from the reference's on-disk cache
 * (orc_write_plane_cache) exactly as barbie.c:274-324 does: one thread, one
 * file read per field use.  Used to time the reference's own access pattern;
 * the sums are those of orc_estimate. */
long orc_estimate_disk(const orc_params *P, const double *low, const double *high,
                       const double *avg, int rcount, const double *rdata, const char *dir,
                       int NRi, int p1_begin, int p1_end, double *bin_bs, double *bin_norm,
                       double *bin_biana, double *contr_out, long contr_cap)
{
    if (!dir) return -1;
    return estimate_impl(P, low, high, avg, rcount, rdata, NULL, dir, NRi, p1_begin, p1_end, 1,
                         bin_bs, bin_norm, bin_biana, contr_out, contr_cap);
}

static long estimate_impl(const orc_params *P, const double *low, const double *high,
                          const double *avg, int rcount, const double *rdata,
                          const double *field, const char *dir, int NRi, int p1_begin,
                          int p1_end, int nthreads, double *bin_bs, double *bin_norm,
                          double *bin_biana, double *contr_out, long contr_cap)
{
    const int NN = P->NN, Nmax = P->Nmax, Nr = P->Nr, n_mtbin = P->n_mtbin;
    const int UL_k1 = P->UL_k1, LL_k1 = P->LL_k1, BW_k1 = P->BW_k1;
    const int n_k1bin = (int)((UL_k1 - LL_k1) / (1. * BW_k1));
    const double vol = pow(NN * P->LL, 3);
    const double Cunits = 2. * M_PI / (NN * P->LL);
    (void)vol;
    const long NN2 = (long)NN * NN;
    const long nbin = (long)n_k1bin * n_mtbin * n_mtbin;
    if (nthreads < 1) nthreads = 1;

    orc_cache C;
    memset(&C, 0, sizeof C);
    C.NN = NN; C.Nr = Nr; C.field = field; C.low = low; C.high = high;
    C.dir = dir; C.NRi = NRi;
    C.cache = (double **)calloc((size_t)Nmax * Nr, sizeof(double *));
    if (!C.cache) return -2;
    if (dir) {
        nthreads = 1;
        for (int k = 0; k < 3; ++k) {
            C.buf[k] = (double *)malloc(sizeof(double) * 2 * (size_t)NN * NN);
            if (!C.buf[k]) {
                for (int q = 0; q < k; ++q) free(C.buf[q]);
                free(C.cache);
                return -2;
            }
        }
    }

    long count = 0;
    long bad = 0;
    long rec = 0;   /* record index (serial runs only) */
    double *tbins = NULL;
    if (nthreads > 1) {
        tbins = (double *)calloc((size_t)nthreads * 3 * nbin * NPOLES, sizeof(double));
        if (!tbins) { free(C.cache); return -2; }
    }

    long *roff = NULL;   /* threaded runs: first record index of each ring triple */
    char *need = NULL;   /* threaded runs: fields a p1 reads */
    if (nthreads > 1) {
        need = (char *)calloc((size_t)Nmax * Nr, 1);
        if (contr_out) roff = (long *)calloc((size_t)rcount + 1, sizeof(long));
        if (!need || (contr_out && !roff)) {
            free(need); free(roff); free(tbins); free(C.cache);
            return -2;
        }
    }

    for (int p1 = p1_begin; p1 < p1_end; ++p1) {
        /* Fields used by this p1 are built up front so the parallel loop
         * below only reads the cache. */
        for (int rp = 0; rp < rcount && !dir; ++rp) {
            int r1 = (int)rdata[rp * 4 + 0];
            (void)cache_get(&C, p1, r1);
        }
        if (nthreads > 1) {
            /* build every field this p1 touches in a serial pre-pass that
             * mirrors the enumeration below */
            for (int rp = 0; rp < rcount; ++rp) {
                int r1 = (int)rdata[rp * 4 + 0], r2 = (int)rdata[rp * 4 + 1], r3 = (int)rdata[rp * 4 + 2];
                int p2_min, p2_max;
                if (r1 == r2 && r2 == r3) { p2_min = 0; p2_max = p1; }
                else if (r1 == r2 && r2 != r3) { p2_min = -p1; p2_max = p1; }
                else if (r1 != r2 && r2 == r3) { p2_min = (int)(-p1 / 2); p2_max = Nmax; }
                else { p2_min = (p1 == 0) ? 0 : -Nmax + 1; p2_max = Nmax; }
                long nrp = 0;
                for (int p2 = p2_min; p2 <= p2_max; ++p2) {
                    int p3 = -p1 - p2;
                    int s2 = p2 < 0 ? -1 : 1, s3 = p3 < 0 ? -1 : 1;
                    if (s3 * p3 < Nmax) {
                        double K[3][2] = {{sqrt(avg[r1] * avg[r1] + p1 * p1), p1},
                                          {sqrt(avg[r2] * avg[r2] + p2 * p2), p2},
                                          {sqrt(avg[r3] * avg[r3] + p3 * p3), p3}};
                        orc_sort(K);
                        if (K[0][0] > (LL_k1) && K[0][0] < UL_k1) {
                            need[(long)(s2 * p2) * Nr + r2] = 1;
                            need[(long)(s3 * p3) * Nr + r3] = 1;
                            ++nrp;
                        }
                    }
                }
                if (roff) roff[rp + 1] = roff[rp] + nrp;
            }
            /* the marked fields are independent: build them in parallel */
            long nslot = (long)Nmax * Nr, nbuilt = 0;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1) reduction(+:nbuilt)
#endif
            for (long k = 0; k < nslot; ++k)
                if (need[k] && !C.cache[k]) {
                    nbuilt += cache_get(&C, (int)(k / Nr), (int)(k % Nr)) != NULL;
                    need[k] = 0;
                }
            (void)nbuilt;
        }

#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 4) reduction(+:count,bad) if(nthreads > 1)
#endif
        for (int rp = 0; rp < rcount; ++rp) {
            int tid = 0;
#ifdef _OPENMP
            tid = omp_get_thread_num();
#endif
            double *bbs = bin_bs, *bnorm = bin_norm, *bbia = bin_biana;
            if (nthreads > 1) {
                bbs = tbins + ((size_t)tid * 3 + 0) * nbin * NPOLES;
                bnorm = tbins + ((size_t)tid * 3 + 1) * nbin * NPOLES;
                bbia = tbins + ((size_t)tid * 3 + 2) * nbin * NPOLES;
            }
            int r1 = (int)rdata[rp * 4 + 0];
            int r2 = (int)rdata[rp * 4 + 1];
            int r3 = (int)rdata[rp * 4 + 2];
            int p2_min, p2_max;
            /* barbie.c:257-268 */
            if (r1 == r2 && r2 == r3) { p2_min = 0; p2_max = p1; }
            else if (r1 == r2 && r2 != r3) { p2_min = -p1; p2_max = p1; }
            else if (r1 != r2 && r2 == r3) { p2_min = (int)(-p1 / 2); p2_max = Nmax; }
            else { p2_min = (p1 == 0) ? 0 : -Nmax + 1; p2_max = Nmax; }

            const double *buffer1 = field_at(&C, p1, r1, 0);
            if (!buffer1) { bad++; continue; }
            long trec = roff ? rec + roff[rp] : 0;   /* this triple's records (threaded) */
            for (int p2 = p2_min; p2 <= p2_max; ++p2) {
                int p3 = -p1 - p2;
                int sign_2 = p2 < 0 ? -1 : 1;
                int sign_3 = p3 < 0 ? -1 : 1;
                if (sign_3 * p3 < Nmax) {
                    /* barbie.c:287-298 */
                    double K[3][2] = {{sqrt(avg[r1] * avg[r1] + p1 * p1), p1},
                                      {sqrt(avg[r2] * avg[r2] + p2 * p2), p2},
                                      {sqrt(avg[r3] * avg[r3] + p3 * p3), p3}};
                    orc_sort(K);
                    if (K[0][0] > (LL_k1) && K[0][0] < UL_k1) {
                        double tp[5];
                        orc_triangle_params((const double(*)[2])K, tp);
                        double K1 = tp[0], T = tp[1], MU = tp[2], pz = tp[3], px = tp[4];
                        double Num = rdata[rp * 4 + 3];
                        /* barbie.c:301-305 */
                        double mu_1 = K[0][1] / K[0][0];
                        double mu_2 = K[1][1] / K[1][0];
                        double mu_3 = K[2][1] / K[2][0];
                        double Biana = 2. * P->ff * (input_ps(Cunits * K[0][0]) * input_ps(Cunits * K[1][0]) + input_ps(Cunits * K[1][0]) * input_ps(Cunits * K[2][0]) + input_ps(Cunits * K[2][0]) * input_ps(Cunits * K[0][0]));
                        Biana = Biana * (1 + P->beta * mu_1 * mu_1) * (1 + P->beta * mu_2 * mu_2) * (1 + P->beta * mu_3 * mu_3);

                        /* barbie.c:326-337 */
                        const double *buffer2 = C.cache[(long)(sign_2 * p2) * Nr + r2];
                        const double *buffer3 = C.cache[(long)(sign_3 * p3) * Nr + r3];
                        if (nthreads == 1) {
                            buffer2 = field_at(&C, sign_2 * p2, r2, 1);
                            buffer3 = field_at(&C, sign_3 * p3, r3, 2);
                        }
                        if (!buffer2 || !buffer3) { bad++; continue; }
                        double Bispec_real = 0.;
                        for (long index = 0; index < NN2; ++index) {
                            Bispec_real += buffer1[2 * index] * buffer2[2 * index] * buffer3[2 * index]
                                - buffer1[2 * index] * sign_2 * buffer2[2 * index + 1] * sign_3 * buffer3[2 * index + 1]
                                - buffer2[2 * index] * buffer1[2 * index + 1] * sign_3 * buffer3[2 * index + 1]
                                - buffer3[2 * index] * buffer1[2 * index + 1] * sign_2 * buffer2[2 * index + 1];
                        }
                        /* barbie.c:343-345 */
                        int K1_bin = floor((K1 - (LL_k1)) / (1. * BW_k1));
                        int T_bin = (T == 1) ? (n_mtbin - 1) : floor((T - .5) / (.5 / (1. * n_mtbin)));
                        int MU_bin = (MU == 1) ? (n_mtbin - 1) : floor((MU - .5) / (.5 / (1. * n_mtbin)));
                        double bs[NPOLES], norm[NPOLES];
                        orc_cal_coeff(px, pz, bs, norm);
                        /* barbie.c:355-357 */
                        if (Bispec_real == 0.) { Biana = 0.; Num = 0.; }
                        long flat = -1;
                        if (K1_bin >= 0 && K1_bin < n_k1bin && T_bin >= 0 && T_bin < n_mtbin &&
                            MU_bin >= 0 && MU_bin < n_mtbin) {
                            flat = ((long)K1_bin * n_mtbin + T_bin) * n_mtbin + MU_bin;
                            /* barbie.c:358-363 */
                            for (int ii = 0; ii < NPOLES; ++ii) {
                                bbs[flat * NPOLES + ii] += Bispec_real * bs[ii];
                                bbia[flat * NPOLES + ii] += Biana * Num * bs[ii];
                                bnorm[flat * NPOLES + ii] += Num * norm[ii];
                            }
                        } else {
                            bad++;   /* the reference would index out of bounds */
                        }
                        if (contr_out) {
                            long *ri = roff ? &trec : &rec;
                            if (*ri < contr_cap) {
                                double *o = contr_out + 8 * *ri;
                                o[0] = p1; o[1] = r1; o[2] = p2; o[3] = r2; o[4] = p3; o[5] = r3;
                                o[6] = Bispec_real; o[7] = (double)flat;
                            }
                            ++*ri;
                        }
                        count++;
                    }
                }
            }
        }
        if (roff) rec += roff[rcount];
    }
    free(roff);
    free(need);
    if (nthreads > 1) {
        for (int t = 0; t < nthreads; ++t)
            for (long k = 0; k < nbin * NPOLES; ++k) {
                bin_bs[k] += tbins[((size_t)t * 3 + 0) * nbin * NPOLES + k];
                bin_norm[k] += tbins[((size_t)t * 3 + 1) * nbin * NPOLES + k];
                bin_biana[k] += tbins[((size_t)t * 3 + 2) * nbin * NPOLES + k];
            }
        free(tbins);
    }
    for (long k = 0; k < (long)Nmax * Nr; ++k) free(C.cache[k]);
    free(C.cache);
    for (int k = 0; k < 3; ++k) free(C.buf[k]);
    if (bad) return -1000 - bad;
    return count;
}

/* ---------------------------------------------------------------------- */
/* Normalisation and the 16 output files, barbie.c:383-397.
 * Writes <dir>/BS_test_ii<ii>_N<NN>_R<NRi>.txt.  The bins are not modified
 * (the reference divides in place; the values printed are the same). */
int orc_write_outputs(const orc_params *P, const double *bin_bs, const double *bin_norm,
                      const double *bin_biana, const char *dir, int NRi)
{
    const int NN = P->NN, n_mtbin = P->n_mtbin;
    const int n_k1bin = (int)((P->UL_k1 - P->LL_k1) / (1. * P->BW_k1));
    const double vol = pow(NN * P->LL, 3);
    char fname[512];
    for (int ii = 0; ii < NPOLES; ++ii) {
        snprintf(fname, sizeof fname, "%s/BS_test_ii%d_N%d_R%d.txt", dir, ii, NN, NRi);
        FILE *fp = fopen(fname, "w");
        if (!fp) return -1;
        for (int i = 0; i < n_k1bin; i++)
            for (int j = 0; j < n_mtbin; j++)
                for (int l = 0; l < n_mtbin; l++) {
                    long b = (((long)i * n_mtbin + j) * n_mtbin + l) * NPOLES;
                    double v = bin_bs[b + ii] / (vol * bin_norm[b + ii]);
                    double a = bin_biana[b + ii] / (bin_norm[b + ii]);
                    if (bin_norm[b + 0] > .1)
                        fprintf(fp, "%d\t%d\t%d\t%e\t%e\t%f\n", i, j, l, v, a, (bin_norm[b + ii]) / (NN * NN * 1.));
                }
        fclose(fp);
    }
    return 0;
}
