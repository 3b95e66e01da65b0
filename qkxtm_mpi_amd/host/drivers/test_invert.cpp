// test_invert.cpp -- counterpart of the reference's bin/test_invert.cpp /
// bin/test_invert_prop.cpp on synthetic data, in the consistent order of the
// latter (boundary condition first): source -> tmDag -> CG on D^dag D -> x,
// then the true residual |D x - source| / |source|.
// Synthetic gauge (generator, eps 0.4) instead of a LIME file, Z2 source
// instead of a smeared point source.
#include "qkxtm/qkxTM.h"
#include "qkxtm/lattice_util.h"
#include "qkxtm/solvers.h"
#include "qkxtm/blas_qkxTM.h"
#include "qkx_hip.h"
#include <cmath>

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
    latInfo.tol = 1e-10;
    latInfo.maxIter = 5000;
    latInfo.boundaryT = -1;
    int rank = 0;

    Gauge *gauge = new Gauge(&latInfo);
    Vector *vec_out = new Vector(&latInfo);
    Vector *vec_tmp = new Vector(&latInfo);
    Vector *vec_in = new Vector(&latInfo);

    gauge->zero();
    if (qkx_gen_gauge(&(gauge->M[0][0][0][0].real()), latInfo.L, 0.4, 1)) errorQKXTM("gauge generation failed\n");
    gauge->applyBoundaryCondition(&latInfo);
    gauge->constGauge = true;

    vec_tmp->zero();
    if (qkx_gen_z2(&(vec_tmp->M[0][0][0].real()), latInfo.L, 2)) errorQKXTM("source generation failed\n");

    vec_in->tmDag(*vec_tmp, *gauge, &latInfo);
    vec_out->zero();

    CG *cg_P = new CG(&latInfo, *gauge);
    CG &cg = *cg_P;
    cg(*vec_out, *vec_in, &latInfo);

    double vec_norm = vec_out->norm2();
    printfQKXTM("the norm of vector is %.17e\n", vec_norm);
    printfQKXTM("iterations %d\n", cg.lastIterations);

    Vector *chk = new Vector(&latInfo);
    chk->tm(*vec_out, *gauge, &latInfo);
    X_eq_X_p_aY(*chk, -1.0, *vec_tmp);
    double res = std::sqrt(chk->norm2() / vec_tmp->norm2());
    printfQKXTM("true residual %.6e\n", res);

    delete chk;
    delete cg_P;
    delete vec_in;
    delete vec_tmp;
    delete vec_out;
    delete gauge;
    return 0;
}
