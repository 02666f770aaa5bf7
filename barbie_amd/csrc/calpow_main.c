/*
 * calpow_mom -- drop-in replacement of the reference's power-spectrum
 * multipoles program (calpow_mom.c:22-117) running on one MI355X through
 * libbarbie_amd (bb_power_multipoles).
 *
 *   calpow_mom NN NRi
 *
 * Same arguments, usage message and exit status for a wrong argument count
 * (calpow_mom.c:43-45), same constants (LL = 0.56, 10 bins,
 * calpow_mom.c:50,85), same input file
 * ResFiles_N<NN>/data_files/field_N<NN>_R<NRi>.txt (raw FP64 r2c buffer,
 * calpow_mom.c:73-81) and same output ResFiles_N<NN>/psoutN<NN>_R<NRi>.dat:
 * rows "%e\t%e\t%e\t%e\t%e\t%e\n" of kmean, P0, P2, P4, P6, num for the bins
 * holding modes (calpow_mom.c:104-113).  A missing or short input exits with
 * status 3, device errors with 4 (the reference does not check).
 */
#include <stdio.h>
#include <stdlib.h>

#include "barbie.h"

int main(int argc, char *argv[])
{
    long NN;
    int NRi;
    if (argc != 3) {
        printf("Usage: %s NN(384) NRi(1) UL_k1(Nmax) LL_k1(Nmax-10) BW_k1(1)\n", argv[0]);
        return 1;
    }
    sscanf(argv[1], "%ld", &NN);
    sscanf(argv[2], "%d", &NRi);
    printf("\n Calculating PS Multipoles for Realization(NRi) = %d and Grids(NN) = %ld\n", NRi, NN);
    const double LL = .56;
    const int NB = 10;
    if (NN < 4 || NN > 4096 || (NN & 1)) {
        fprintf(stderr, "calpow_mom: NN must be even and in [4, 4096]\n");
        return 2;
    }
    const size_t n = (size_t)NN * NN * (NN + 2);
    double *field = malloc(n * sizeof(double));
    if (!field) {
        fprintf(stderr, "calpow_mom: out of memory\n");
        return 3;
    }
    char file1[256], file3[256];
    snprintf(file1, sizeof file1, "ResFiles_N%ld/data_files/field_N%ld_R%d.txt", NN, NN, NRi);
    FILE *fp1 = fopen(file1, "rb");
    if (!fp1 || fread(field, sizeof(double), n, fp1) != n) {
        fprintf(stderr, "calpow_mom: cannot read %s\n", file1);
        if (fp1) fclose(fp1);
        free(field);
        return 3;
    }
    fclose(fp1);
    double out[6 * 10];
    if (bb_power_multipoles(field, 0, (int)NN, NB, LL, out)) {
        fprintf(stderr, "calpow_mom: %s\n", bb_last_error());
        free(field);
        return 4;
    }
    free(field);
    snprintf(file3, sizeof file3, "ResFiles_N%ld/psoutN%ld_R%d.dat", NN, NN, NRi);
    FILE *fp3 = fopen(file3, "w");
    if (!fp3) {
        fprintf(stderr, "calpow_mom: cannot create %s\n", file3);
        return 3;
    }
    for (int i = 0; i < NB; ++i)
        if (out[6 * i + 5] > 0)
            fprintf(fp3, "%e\t%e\t%e\t%e\t%e\t%e\n", out[6 * i], out[6 * i + 1], out[6 * i + 2],
                    out[6 * i + 3], out[6 * i + 4], out[6 * i + 5]);
    fclose(fp3);
    return 0;
}
