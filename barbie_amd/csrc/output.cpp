// Normalisation and the 16 multipole files of barbie.c:383-397.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <string>

#include "plan.hpp"

namespace {

// One multipole file's value columns.  checked: a (bin, multipole) whose own
// denominator is not positive and finite, or whose quotient is not finite,
// is written as "nan" in both value columns and counted; unchecked: the
// reference's arithmetic, whatever the division gives.
int write_files(const bb::PlanData &d, const double *bin_bs, const double *bin_norm,
                const double *bin_biana, const char *dir, int NRi, bool checked,
                long long *n_unsupported)
{
    const int NN = d.g.NN, n_mtbin = d.g.n_mtbin, n_k1bin = d.n_k1bin;
    const double vol = std::pow(NN * d.g.LL, 3);
    long long bad = 0;
    for (int ii = 0; ii < BB_NPOLES; ++ii) {
        char name[4096];
        std::snprintf(name, sizeof name, "%s/BS_test_ii%d_N%d_R%d.txt", dir, ii, NN, NRi);
        FILE *fp = std::fopen(name, "w");
        if (!fp) return bb::set_error(BB_E_IO, std::string("cannot open ") + name);
        for (int i = 0; i < n_k1bin; i++)
            for (int j = 0; j < n_mtbin; j++)
                for (int l = 0; l < n_mtbin; l++) {
                    const long b = (((long)i * n_mtbin + j) * n_mtbin + l) * BB_NPOLES;
                    const double v = bin_bs[b + ii] / (vol * bin_norm[b + ii]);
                    const double a = bin_biana ? bin_biana[b + ii] / (bin_norm[b + ii]) : 0.0;
                    if (!(bin_norm[b + 0] > .1)) continue;
                    const double nrm = bin_norm[b + ii];
                    if (checked && !(nrm > 0 && std::isfinite(nrm) && std::isfinite(v) &&
                                     std::isfinite(a))) {
                        ++bad;
                        std::fprintf(fp, "%d\t%d\t%d\tnan\tnan\t%f\n", i, j, l, nrm / (NN * NN * 1.));
                    } else {
                        std::fprintf(fp, "%d\t%d\t%d\t%e\t%e\t%f\n", i, j, l, v, a,
                                     nrm / (NN * NN * 1.));
                    }
                }
        if (std::fclose(fp) != 0) return bb::set_error(BB_E_IO, std::string("write failed: ") + name);
    }
    if (n_unsupported) *n_unsupported = bad;
    return BB_OK;
}

}  // namespace

extern "C" int bb_write_outputs(const bb_plan *p, const double *bin_bs, const double *bin_norm,
                                const double *bin_biana, const char *dir, int NRi)
{
    if (!p || !bin_bs || !bin_norm || !dir)
        return bb::set_error(BB_E_INVALID, "null argument");
    return write_files(p->d, bin_bs, bin_norm, bin_biana, dir, NRi, false, nullptr);
}

extern "C" int bb_write_outputs_checked(const bb_plan *p, const double *bin_bs,
                                        const double *bin_norm, const double *bin_biana,
                                        const char *dir, int NRi, long long *n_unsupported)
{
    if (!p || !bin_bs || !bin_norm || !dir)
        return bb::set_error(BB_E_INVALID, "null argument");
    return write_files(p->d, bin_bs, bin_norm, bin_biana, dir, NRi, true, n_unsupported);
}
