/*
 * solvers.h -- drop-in counterpart of the reference's include/solvers.h:
 * CG on D^dag D (lib/solvers.cpp:8-81), run on the GPU.
 */
#ifndef QKXTM_DROPIN_SOLVERS_H
#define QKXTM_DROPIN_SOLVERS_H

#include "lattice_util.h"

class CG {
  private:
    int maxIter;
    double tol;
    double reliableDelta;
    qkxTM::Gauge &u;

  public:
    CG(LatticeInfo *param, qkxTM::Gauge &gauge);
    ~CG() {}

    int MaxIter() { return maxIter; }
    double Tol() { return tol; }
    double ReliableDelta() { return reliableDelta; }

    /* solves D^dag D out = in; out accumulates onto its value, r0 = in
     * (the reference assumes out = 0) */
    void operator()(qkxTM::Vector &out, qkxTM::Vector &in, LatticeInfo *param);

    /* iterations of the last solve (not in the reference) */
    int lastIterations;
};

#endif
