// Normalisation and the 16 multipole files of barbie.c:383-397.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <string>

#include "plan.hpp"

extern "C" int bb_write_outputs(const bb_plan *p, const double *bin_bs, const double *bin_norm,
                                const double *bin_biana, const char *dir, int NRi)
{
    if (!p || !bin_bs || !bin_norm || !dir)
        return bb::set_error(BB_E_INVALID, "null argument");
    const bb::PlanData &d = p->d;
    const int NN = d.g.NN, n_mtbin = d.g.n_mtbin, n_k1bin = d.n_k1bin;
    const double vol = std::pow(NN * d.g.LL, 3);
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
                    if (bin_norm[b + 0] > .1)
                        std::fprintf(fp, "%d\t%d\t%d\t%e\t%e\t%f\n", i, j, l, v, a,
                                     (bin_norm[b + ii]) / (NN * NN * 1.));
                }
        if (std::fclose(fp) != 0) return bb::set_error(BB_E_IO, std::string("write failed: ") + name);
    }
    return BB_OK;
}
