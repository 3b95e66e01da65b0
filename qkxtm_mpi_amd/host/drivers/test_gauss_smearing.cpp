// test_gauss_smearing.cpp -- counterpart of the reference's
// bin/test_gauss_smearing.cpp on synthetic data: plaquette, APE smearing of a
// copy of the links (NsmearAPE 20, alpha 0.5), the smeared plaquette, a unit
// point source at site 0, Gaussian smearing with the APE links (NsmearGauss
// 50, alpha 4), its norm, and the time slice t = 0 of the result written to a
// file as "x y z mu c re im" (the reference writes the same slice to kale.dat).
// Synthetic gauge (generator, eps 0.4, 8^3x16), passed through an ILDG/LIME
// file when a second argument names one; a third argument overrides the
// Gaussian step count.  The norm of vec_tmp after the smearing is printed too
// (the reference leaves the last or the next-to-last iterate there).
#include "qkxtm/qkxTM.h"
#include "qkxtm/lattice_util.h"
#include "qkx_hip.h"
#include <cstdio>
#include <cstdlib>

using namespace qkxTM;

int main(int argc, char *argv[])
{
    LatticeInfo latInfo;
    latInfo.L[0] = 8;
    latInfo.L[1] = 8;
    latInfo.L[2] = 8;
    latInfo.L[3] = 16;
    latInfo.boundaryT = +1;
    latInfo.NsmearAPE = 20;
    latInfo.alphaAPE = 0.5;
    latInfo.NsmearGauss = 50;
    latInfo.alphaGauss = 4.0;
    if (argc > 3) latInfo.NsmearGauss = atoi(argv[3]);
    int rank = 0;

    Gauge *gauge = new Gauge(&latInfo);
    Gauge *gauge_tmp = new Gauge(&latInfo);
    Gauge *gauge_ape = new Gauge(&latInfo);
    Vector *vec = new Vector(&latInfo);
    Vector *vec_tmp = new Vector(&latInfo);

    gauge->zero();
    if (qkx_gen_gauge(&(gauge->M[0][0][0][0].real()), latInfo.L, 0.4, 1)) errorQKXTM("gauge generation failed\n");
    if (argc > 2) {
        // through an ILDG/LIME file, as the reference driver reads its
        // configuration (bin/test_gauss_smearing.cpp:59)
        char msg[] = "synthetic configuration, eps 0.4, seed 1";
        qkxTM_writeGaugeLime(argv[2], *gauge, msg);
        gauge->zero();
        qkxTM_MPI_getGaugeLime(argv[2], *gauge);
    }
    gauge_tmp->copy(*gauge);
    double plaq = gauge->calculatePlaquette();
    printfQKXTM("plaquette is %.15f\n", plaq);

    gauge_ape->APE_smearing(*gauge_tmp, &latInfo);
    delete gauge_tmp;
    plaq = gauge_ape->calculatePlaquette();
    printfQKXTM("plaquette smeared is %.15f\n", plaq);

    vec_tmp->zero();
    vec_tmp->M[0][0][0] = Complex(1.0, 0.0);
    vec->Gaussian_smearing(*vec_tmp, *gauge_ape, &latInfo);
    printfQKXTM("the norm of vec_tmp is %.17e\n", vec_tmp->norm2());
    delete vec_tmp;
    double vec_norm = vec->norm2();
    printfQKXTM("the norm of vector is %.17e\n", vec_norm);

    if (argc > 1) {
        FILE *f = fopen(argv[1], "w");
        if (!f) errorQKXTM("cannot open %s\n", argv[1]);
        for (int z = 0; z < latInfo.L[2]; z++)
            for (int y = 0; y < latInfo.L[1]; y++)
                for (int x = 0; x < latInfo.L[0]; x++) {
                    long iv = (z * latInfo.L[1] + y) * latInfo.L[0] + x;
                    for (int mu = 0; mu < 4; mu++)
                        for (int c = 0; c < 3; c++)
                            fprintf(f, "%d %d %d %d %d %+.17e %+.17e\n", x, y, z, mu, c, vec->M[iv][mu][c].real(),
                                    vec->M[iv][mu][c].imag());
                }
        fclose(f);
    }
    delete vec;
    delete gauge_ape;
    delete gauge;
    return 0;
}
