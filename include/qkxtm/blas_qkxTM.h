/*
 * blas_qkxTM.h -- drop-in counterpart of the reference's
 * include/blas_qkxTM.h (lib/blas_qkxTM.cpp), run on the GPU.
 */
#ifndef QKXTM_DROPIN_BLAS_QKXTM_H
#define QKXTM_DROPIN_BLAS_QKXTM_H

#include "lattice_util.h"

double reCdotXdagY(qkxTM::Vector &x, qkxTM::Vector &y);
void X_eq_X_p_aY(qkxTM::Vector &x, double a, qkxTM::Vector &y);
void X_eq_aX_p_Y(qkxTM::Vector &x, double a, qkxTM::Vector &y);

#endif
