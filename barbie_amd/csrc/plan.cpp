// Host-side geometry planner: ring edges, ring masks, ring triples and the
// contraction list of the estimator, in the reference's enumeration order.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <set>
#include <string>
#include <unordered_set>

#include "plan.hpp"

namespace bb {

static thread_local std::string g_err;

int set_error(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

const char *last_error() { return g_err.c_str(); }

static bool good_fft_size(int n)
{
    if (n < 2) return false;
    for (int f : {2, 3, 5})
        while (n % f == 0) n /= f;
    return n == 1;
}

// Ring edges.  BB_RINGS_REFERENCE follows barbie.c:110-130 (blocks of ten
// with width b+1 in block b; every edge but the first shifted by -0.1; the
// upper edge of ring i is the lower edge of ring i+1; the last ring closes at
// Nmax; avg is the midpoint).  For Nr not a multiple of ten the rule is cut
// at Nr.  BB_RINGS_UNIFORM applies the same construction with a fixed width.
static int make_edges(const bb_geometry &g, std::vector<double> &lo,
                      std::vector<double> &hi, std::vector<double> &av)
{
    const int Nr = g.Nr;
    lo.assign(Nr, 0.0);
    hi.assign(Nr, 0.0);
    av.assign(Nr, 0.0);
    if (g.ring_rule == BB_RINGS_EXPLICIT) {
        if (!g.ring_low || !g.ring_high)
            return set_error(BB_E_INVALID, "explicit ring rule needs ring_low and ring_high");
        for (int i = 0; i < Nr; ++i) {
            lo[i] = g.ring_low[i];
            hi[i] = g.ring_high[i];
        }
    } else if (g.ring_rule == BB_RINGS_REFERENCE || g.ring_rule == BB_RINGS_UNIFORM) {
        double current = 0.0;
        for (int i = 0; i < Nr; ++i) {
            lo[i] = current;
            current += (g.ring_rule == BB_RINGS_REFERENCE) ? (double)(i / 10 + 1) : g.ring_width;
        }
        for (int i = 1; i < Nr; ++i) {
            lo[i] -= .1;
            hi[i - 1] = lo[i];
        }
        if (g.ring_rule == BB_RINGS_UNIFORM && !g.last_ring_to_nmax)
            hi[Nr - 1] = current - .1;
        else
            hi[Nr - 1] = 1.0 * g.Nmax;
    } else {
        return set_error(BB_E_INVALID, "unknown ring rule");
    }
    for (int i = 0; i < Nr; ++i) {
        if (!(std::isfinite(lo[i]) && std::isfinite(hi[i]) && lo[i] < hi[i] && lo[i] >= 0.0))
            return set_error(BB_E_INVALID, "ring " + std::to_string(i) + " has invalid edges");
        if (i > 0 && lo[i] < hi[i - 1] - 1e-12)
            return set_error(BB_E_INVALID, "rings overlap");
        av[i] = (hi[i] + lo[i]) / 2.;
        if (!(av[i] > 0.0))
            return set_error(BB_E_INVALID, "ring " + std::to_string(i) + " has zero mean radius");
    }
    return BB_OK;
}

int plan_create(const bb_geometry *gp, PlanData &d)
{
    if (!gp) return set_error(BB_E_INVALID, "null geometry");
    const bb_geometry &g = *gp;
    d.g = g;
    if (g.NN < 4 || (g.NN & 1) || !good_fft_size(g.NN))
        return set_error(BB_E_INVALID, "NN must be even, >= 4, with prime factors 2, 3, 5 only");
    if (g.NN > 600)
        return set_error(BB_E_INVALID, "NN above 600 exceeds the FFT kernels' LDS tiles");
    if (g.Nmax < 1 || g.Nmax > g.NN / 2)
        return set_error(BB_E_INVALID, "Nmax must be in [1, NN/2]");
    if (g.Nr < 1 || g.Nr > 4096) return set_error(BB_E_INVALID, "Nr must be in [1, 4096]");
    if (g.n_mtbin < 1) return set_error(BB_E_INVALID, "n_mtbin must be >= 1");
    if (g.BW_k1 <= 0) return set_error(BB_E_INVALID, "BW_k1 must be positive");
    if (g.UL_k1 <= g.LL_k1) return set_error(BB_E_INVALID, "UL_k1 must exceed LL_k1");
    if ((g.UL_k1 - g.LL_k1) % g.BW_k1 != 0)
        return set_error(BB_E_INVALID, "UL_k1 - LL_k1 must be divisible by BW_k1");
    if (!(g.LL > 0)) return set_error(BB_E_INVALID, "cell size LL must be positive");

    int rc = make_edges(g, d.low, d.high, d.avg);
    if (rc) return rc;
    const int NN = g.NN, Nr = g.Nr, Nmax = g.Nmax;
    if (d.high[Nr - 1] > NN / 2)
        return set_error(BB_E_INVALID, "outermost ring edge exceeds NN/2");

    // Ring masks (barbie.c:135-148) as integer bounds on n = ia^2 + ja^2: the
    // reference predicate sqrt(n) >= Low && sqrt(n) < High is monotone in n,
    // so it holds exactly on an interval [nlo, nhi).
    const int nmax_sq = 2 * (NN / 2) * (NN / 2);
    d.nlo.assign(Nr, 0);
    d.nhi.assign(Nr, 0);
    for (int r = 0; r < Nr; ++r) {
        int a = -1, b = -1;
        for (int n = 0; n <= nmax_sq + 1; ++n) {
            double kk = std::sqrt(1. * n);
            bool in = (kk >= d.low[r] && kk < d.high[r]);
            if (in && a < 0) a = n;
            if (in) b = n + 1;
        }
        if (a < 0) { a = 0; b = 0; }
        d.nlo[r] = a;
        d.nhi[r] = b;
    }

    // Ring triples, funcs.c:389-404.
    d.tri.clear();
    for (int r1 = 0; r1 < Nr; ++r1)
        for (int r2 = 0; r2 <= r1; ++r2) {
            double tt = d.avg[r2] / d.avg[r1];
            if (tt >= 0.5 && tt <= 1.) {
                for (int r3 = 0; r3 <= r2; ++r3) {
                    double mu = 0.5 * (tt + 1. / tt - (d.avg[r3] / d.avg[r1]) * (d.avg[r3] / d.avg[r2]));
                    if (mu >= 0.5 / tt && mu <= 1.) {
                        d.tri.push_back(r1);
                        d.tri.push_back(r2);
                        d.tri.push_back(r3);
                    }
                }
            }
        }
    const int rcount = (int)(d.tri.size() / 3);

    // Contractions, barbie.c:247-292, and their bins, barbie.c:343-345.
    d.n_k1bin = (int)((g.UL_k1 - g.LL_k1) / (1. * g.BW_k1));
    d.nbins = d.n_k1bin * g.n_mtbin * g.n_mtbin;
    const int n_mtbin = g.n_mtbin;
    d.contr.clear();
    d.field_used.assign((size_t)Nmax * Nr, 0);
    long long bad_bin = 0, bad_mu = 0;
    for (int p1 = 0; p1 < Nmax; ++p1) {
        for (int rp = 0; rp < rcount; ++rp) {
            const int r1 = d.tri[3 * rp], r2 = d.tri[3 * rp + 1], r3 = d.tri[3 * rp + 2];
            int p2_min, p2_max;
            if (r1 == r2 && r2 == r3) { p2_min = 0; p2_max = p1; }
            else if (r1 == r2 && r2 != r3) { p2_min = -p1; p2_max = p1; }
            else if (r1 != r2 && r2 == r3) { p2_min = (int)(-p1 / 2); p2_max = Nmax; }
            else { p2_min = (p1 == 0) ? 0 : -Nmax + 1; p2_max = Nmax; }
            for (int p2 = p2_min; p2 <= p2_max; ++p2) {
                const int p3 = -p1 - p2;
                const int s3 = p3 < 0 ? -1 : 1;
                if (s3 * p3 >= Nmax) continue;
                double K[3][2] = {{std::sqrt(d.avg[r1] * d.avg[r1] + p1 * p1), (double)p1},
                                  {std::sqrt(d.avg[r2] * d.avg[r2] + p2 * p2), (double)p2},
                                  {std::sqrt(d.avg[r3] * d.avg[r3] + p3 * p3), (double)p3}};
                for (int i = 0; i < 2; i++)
                    for (int j = 0; j < 2 - i; j++)
                        if (K[j][0] < K[j + 1][0]) {
                            double s1 = K[j][0], s2 = K[j][1];
                            K[j][0] = K[j + 1][0]; K[j][1] = K[j + 1][1];
                            K[j + 1][0] = s1; K[j + 1][1] = s2;
                        }
                if (!(K[0][0] > g.LL_k1 && K[0][0] < g.UL_k1)) continue;
                const double K1 = K[0][0];
                const double T = K[1][0] / K[0][0];
                const double MU = 0.5 * (T + 1. / T - (K[2][0] / K[0][0]) * (K[2][0] / K[1][0]));
                const int kb = (int)std::floor((K1 - g.LL_k1) / (1. * g.BW_k1));
                const int tb = (T == 1) ? (n_mtbin - 1) : (int)std::floor((T - .5) / (.5 / (1. * n_mtbin)));
                const int mb = (MU == 1) ? (n_mtbin - 1) : (int)std::floor((MU - .5) / (.5 / (1. * n_mtbin)));
                if (kb < 0 || kb >= d.n_k1bin || tb < 0 || tb >= n_mtbin || mb < 0 || mb >= n_mtbin) {
                    bad_bin++;
                    continue;
                }
                if (!(1. - MU * MU > 0.)) bad_mu++;
                Contraction c;
                c.p1 = (int16_t)p1; c.r1 = (int16_t)r1;
                c.p2 = (int16_t)p2; c.r2 = (int16_t)r2;
                c.p3 = (int16_t)p3; c.r3 = (int16_t)r3;
                c.rp = rp;
                c.bin = (kb * n_mtbin + tb) * n_mtbin + mb;
                d.contr.push_back(c);
                d.field_used[(size_t)p1 * Nr + r1] = 1;
                d.field_used[(size_t)std::abs(p2) * Nr + r2] = 1;
                d.field_used[(size_t)std::abs(p3) * Nr + r3] = 1;
            }
        }
    }
    if (bad_bin)
        return set_error(BB_E_RANGE, std::to_string(bad_bin) +
                                         " triangles fall outside the (k1, t, mu) bin grid");
    if (bad_mu)
        return set_error(BB_E_RANGE, std::to_string(bad_mu) +
                                         " triangles have mu = 1 (orientation undefined)");

    std::unordered_set<long long> pairs;
    std::set<std::pair<int, int>> groups;
    for (const Contraction &c : d.contr) {
        long long key = ((((long long)c.p1 * 4096 + c.r1) * 8192 + (c.p2 + 4096)) * 4096) + c.r2;
        pairs.insert(key);
        groups.insert({c.p1, c.p2});
    }
    d.npairs = (long long)pairs.size();
    d.ngroups = (int)groups.size();
    d.nfields_used = 0;
    for (uint8_t u : d.field_used) d.nfields_used += u;
    return BB_OK;
}

}  // namespace bb

extern "C" {

const char *bb_last_error(void) { return bb::last_error(); }

const char *bb_version(void) { return "barbie_amd 0.1 (gfx950)"; }

int bb_geometry_reference(int NN, int UL_k1, int LL_k1, int BW_k1, bb_geometry *g)
{
    if (!g) return bb::set_error(BB_E_INVALID, "null geometry");
    std::memset(g, 0, sizeof *g);
    g->NN = NN;
    g->Nmax = 100;
    g->Nr = 40;
    g->ring_rule = BB_RINGS_REFERENCE;
    g->ring_width = 1.0;
    g->last_ring_to_nmax = 1;
    g->n_mtbin = 5;
    g->UL_k1 = UL_k1;
    g->LL_k1 = LL_k1;
    g->BW_k1 = BW_k1;
    g->LL = 1.0;
    g->ff = 0.1;
    g->beta = 1.0;
    return BB_OK;
}

int bb_plan_create(const bb_geometry *g, bb_plan **out)
{
    if (!out) return bb::set_error(BB_E_INVALID, "null output pointer");
    *out = nullptr;
    bb_plan *p = new (std::nothrow) bb_plan;
    if (!p) return bb::set_error(BB_E_NOMEM, "out of host memory");
    int rc;
    try {
        rc = bb::plan_create(g, p->d);
    } catch (const std::bad_alloc &) {
        rc = bb::set_error(BB_E_NOMEM, "out of host memory while planning");
    }
    if (rc) {
        delete p;
        return rc;
    }
    *out = p;
    return BB_OK;
}

void bb_plan_destroy(bb_plan *p) { delete p; }

int bb_plan_stats_get(const bb_plan *p, bb_plan_stats *s)
{
    if (!p || !s) return bb::set_error(BB_E_INVALID, "null argument");
    const bb::PlanData &d = p->d;
    s->NN = d.g.NN;
    s->Nmax = d.g.Nmax;
    s->Nr = d.g.Nr;
    s->n_mtbin = d.g.n_mtbin;
    s->n_k1bin = d.n_k1bin;
    s->rcount = (int)(d.tri.size() / 3);
    s->ncontr = (long long)d.contr.size();
    s->npairs = d.npairs;
    s->ngroups = d.ngroups;
    s->nfields_used = d.nfields_used;
    s->nbins = d.nbins;
    return BB_OK;
}

int bb_plan_ring_edges(const bb_plan *p, double *low, double *high, double *avg)
{
    if (!p) return bb::set_error(BB_E_INVALID, "null plan");
    const bb::PlanData &d = p->d;
    for (int i = 0; i < d.g.Nr; ++i) {
        if (low) low[i] = d.low[i];
        if (high) high[i] = d.high[i];
        if (avg) avg[i] = d.avg[i];
    }
    return BB_OK;
}

int bb_plan_triples(const bb_plan *p, int *r123)
{
    if (!p || !r123) return bb::set_error(BB_E_INVALID, "null argument");
    std::memcpy(r123, p->d.tri.data(), sizeof(int) * p->d.tri.size());
    return BB_OK;
}

int bb_plan_contractions(const bb_plan *p, int *rec)
{
    if (!p || !rec) return bb::set_error(BB_E_INVALID, "null argument");
    size_t k = 0;
    for (const bb::Contraction &c : p->d.contr) {
        rec[k++] = c.p1; rec[k++] = c.r1; rec[k++] = c.p2; rec[k++] = c.r2;
        rec[k++] = c.p3; rec[k++] = c.r3; rec[k++] = c.rp; rec[k++] = c.bin;
    }
    return BB_OK;
}

}  // extern "C"
