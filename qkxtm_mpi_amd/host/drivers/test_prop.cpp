// test_prop.cpp -- counterpart of the propagator part of the reference's
// bin/test_seq_source.cpp (lines 81-117 and 143-144) on synthetic data: for
// each spin-colour (mu, ic) a unit point source at site 0, Gaussian smearing
// with APE-smeared links, tmDag, CG on D^dag D, rescale(2 kappa) and
// absorbVector, for twistSign +1 (up) and -1 (down); then the rotations to the
// physical basis.  Prints both propagators' norms and writes their site-0
// matrices to a file as "quark mu nu c1 c2 re im".  Synthetic gauge
// (generator, eps 0.4, 8^3x16) instead of a LIME file; APE 5 x 0.5 and
// Gaussian 10 x 4.0 to keep the run short.
#include "qkxtm/qkxTM.h"
#include "qkxtm/lattice_util.h"
#include "qkxtm/solvers.h"
#include "qkx_hip.h"
#include <cstdio>

using namespace qkxTM;

int main(int argc, char *argv[])
{
    LatticeInfo latInfo;
    latInfo.L[0] = 8;
    latInfo.L[1] = 8;
    latInfo.L[2] = 8;
    latInfo.L[3] = 16;
    latInfo.kappa = 0.156;
    latInfo.mu = 0.01;
    latInfo.twistSign = +1;
    latInfo.tol = 1e-12;
    latInfo.maxIter = 20000;
    latInfo.boundaryT = -1;
    latInfo.NsmearAPE = 5;
    latInfo.alphaAPE = 0.5;
    latInfo.NsmearGauss = 10;
    latInfo.alphaGauss = 4.0;
    int rank = 0;

    Gauge *gauge = new Gauge(&latInfo);
    Gauge *gauge_tmp = new Gauge(&latInfo);
    Gauge *gauge_APE = new Gauge(&latInfo);
    Vector *vec_in = new Vector(&latInfo);
    Vector *vec_out = new Vector(&latInfo);
    Vector *vec_tmp = new Vector(&latInfo);
    Propagator *uprop = new Propagator(&latInfo);
    Propagator *dprop = new Propagator(&latInfo);

    gauge->zero();
    if (qkx_gen_gauge(&(gauge->M[0][0][0][0].real()), latInfo.L, 0.4, 1)) errorQKXTM("gauge generation failed\n");
    gauge_tmp->copy(*gauge);
    gauge_APE->APE_smearing(*gauge_tmp, &latInfo);
    gauge->applyBoundaryCondition(&latInfo);
    gauge->constGauge = true;
    gauge_APE->constGauge = true;

    CG *cg_P = new CG(&latInfo, *gauge);
    CG &cg = *cg_P;
    for (int quark = 0; quark < 2; quark++) {
        latInfo.twistSign = quark == 0 ? +1 : -1;
        Propagator *prop = quark == 0 ? uprop : dprop;
        for (int mu = 0; mu < 4; mu++)
            for (int ic = 0; ic < 3; ic++) {
                vec_tmp->zero();
                vec_out->zero();
                vec_tmp->M[0][mu][ic] = Complex(1.0, 0.0);
                vec_out->Gaussian_smearing(*vec_tmp, *gauge_APE, &latInfo);
                vec_in->tmDag(*vec_out, *gauge, &latInfo);
                vec_out->zero();
                cg(*vec_out, *vec_in, &latInfo);
                vec_out->rescale(2 * latInfo.kappa);
                prop->absorbVector(*vec_out, mu, ic);
            }
    }
    uprop->rotateToPhysicalBasePlus();
    dprop->rotateToPhysicalBaseMinus();
    printfQKXTM("up norm %.17e\n", uprop->norm2());
    printfQKXTM("down norm %.17e\n", dprop->norm2());
    if (argc > 1) {
        FILE *f = fopen(argv[1], "w");
        if (!f) errorQKXTM("cannot open %s\n", argv[1]);
        for (int quark = 0; quark < 2; quark++) {
            Propagator *prop = quark == 0 ? uprop : dprop;
            for (int mu = 0; mu < 4; mu++)
                for (int nu = 0; nu < 4; nu++)
                    for (int c1 = 0; c1 < 3; c1++)
                        for (int c2 = 0; c2 < 3; c2++)
                            fprintf(f, "%d %d %d %d %d %+.17e %+.17e\n", quark, mu, nu, c1, c2,
                                    prop->M[0][mu][nu][c1][c2].real(), prop->M[0][mu][nu][c1][c2].imag());
        }
        fclose(f);
    }
    delete cg_P;
    delete uprop;
    delete dprop;
    delete vec_tmp;
    delete vec_out;
    delete vec_in;
    delete gauge_APE;
    delete gauge_tmp;
    delete gauge;
    return 0;
}
