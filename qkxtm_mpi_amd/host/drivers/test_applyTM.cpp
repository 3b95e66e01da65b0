// test_applyTM.cpp -- counterpart of the reference's bin/test_applyTM.cpp on
// synthetic data: apply the even-odd normal operator tm_xxDagTm_xx(OO) to a
// vector and print every site (t z y x, input re im, output re im).
// The gauge comes from the library's generator instead of a LIME file, the
// input vector is Z2 noise instead of an eigenvector file.
//   usage: test_applyTM [dump-file]
#include "qkxtm/qkxTM.h"
#include "qkxtm/lattice_util.h"
#include "qkx_hip.h"

using namespace qkxTM;

int main(int argc, char *argv[])
{
    LatticeInfo latInfo;
    latInfo.L[0] = 8;
    latInfo.L[1] = 8;
    latInfo.L[2] = 8;
    latInfo.L[3] = 8;
    latInfo.kappa = 0.156;
    latInfo.mu = 0.1;
    latInfo.twistSign = +1;
    latInfo.boundaryT = +1;
    int rank = 0;

    Gauge *gauge = new Gauge(&latInfo);
    Vector *vec = new Vector(&latInfo);
    Vector *vecIn = new Vector(&latInfo);
    printfQKXTM("global lattice (x,y,z,t) =  %d x %d x %d x %d\n", gauge->L[0], gauge->L[1], gauge->L[2], gauge->L[3]);

    gauge->zero();
    if (qkx_gen_gauge(&(gauge->M[0][0][0][0].real()), latInfo.L, 0.4, 11)) errorQKXTM("gauge generation failed\n");
    gauge->applyBoundaryCondition(&latInfo);

    vecIn->zero();
    if (qkx_gen_z2(&(vecIn->M[0][0][0].real()), latInfo.L, 12)) errorQKXTM("source generation failed\n");
    gauge->communicateToPlus();
    gauge->communicateToMinus();
    vecIn->communicateToPlus();
    vecIn->communicateToMinus();
    vec->tm_xxDagTm_xx(*vecIn, *gauge, &latInfo, OO);

    FILE *out = argc > 1 ? fopen(argv[1], "w") : stdout;
    if (!out) errorQKXTM("cannot open %s\n", argv[1]);
    for (int t = 0; t < vecIn->lL[3]; t++)
        for (int z = 0; z < vecIn->lL[2]; z++)
            for (int y = 0; y < vecIn->lL[1]; y++)
                for (int x = 0; x < vecIn->lL[0]; x++)
                    for (int mu = 0; mu < 4; mu++)
                        for (int c1 = 0; c1 < 3; c1++)
                            fprintf(out, "%d %d %d %d %+.17e %+.17e \t %+.17e %+.17e\n", t, z, y, x,
                                    vecIn->M[LEXIC(t, z, y, x, vecIn->lL)][mu][c1].real(),
                                    vecIn->M[LEXIC(t, z, y, x, vecIn->lL)][mu][c1].imag(),
                                    vec->M[LEXIC(t, z, y, x, vecIn->lL)][mu][c1].real(),
                                    vec->M[LEXIC(t, z, y, x, vecIn->lL)][mu][c1].imag());
    if (out != stdout) fclose(out);
    vec->normalize();
    printfQKXTM("norm of the normalized vector is %.17e\n", vec->norm2());
    delete gauge;
    delete vecIn;
    delete vec;
    return 0;
}
