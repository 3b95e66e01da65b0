/*
 * qkxTM.h -- drop-in counterpart of the reference's include/qkxTM.h for the
 * twisted-mass operator / CG path, backed by libqkx_hip.so (one MI355X).
 *
 * C++98 (gnu++98) like the reference, so existing drivers keep compiling;
 * no HIP or MPI header is included.  Only the hot-path part of the reference
 * header is provided: Complex, the size/index macros
 * (include/qkxTM.h:117-134), the error/print macros (:137-146) and
 * LatticeInfo (:156-204), whose fields, meanings and defaults are kept so a
 * driver that fills one keeps working.
 */
#ifndef QKXTM_DROPIN_QKXTM_H
#define QKXTM_DROPIN_QKXTM_H

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <complex>

typedef std::complex<double> Complex;

#define NDIM 4
#define NSPINS 4
#define NCOLORS 3
#define NDF 2
#define MOM_MAX 1000
#define NSPINOR (NSPINS * NCOLORS * NDF)
#define NLINKS (NDIM * NCOLORS * NCOLORS * NDF)
#define NPROPS (NSPINS * NSPINS * NCOLORS * NCOLORS * NDF)

#define MS(spin, color) ((spin) * NCOLORS + color)
#define MG(v, color1, color2) ((v) * NCOLORS * NCOLORS + (color1) * NCOLORS + color2)
#define MU(color1, color2) ((color1) * NCOLORS + color2)
#define LEXIC(it, iz, iy, ix, L) ((it) * L[0] * L[1] * L[2] + (iz) * L[0] * L[1] + (iy) * L[0] + (ix))
/* lexicographic indices within the 3-d faces (ghost-zone order of the reference) */
#define LEXIC_TZY(it, iz, iy, L) ((it) * L[1] * L[2] + (iz) * L[1] + (iy))
#define LEXIC_TZX(it, iz, ix, L) ((it) * L[0] * L[2] + (iz) * L[0] + (ix))
#define LEXIC_TYX(it, iy, ix, L) ((it) * L[0] * L[1] + (iy) * L[0] + (ix))
#define LEXIC_ZYX(iz, iy, ix, L) ((iz) * L[0] * L[1] + (iy) * L[0] + (ix))

/* The reference aborts the MPI job; one GPU process exits. */
#define errorQKXTM(...)               \
    do {                              \
        fprintf(stderr, __VA_ARGS__); \
        exit(-1);                     \
    } while (0)

#define printfQKXTM(...)              \
    do {                              \
        if (rank == 0)                \
            fprintf(stdout, __VA_ARGS__); \
    } while (0)

struct LatticeInfo {
    /* index 0 -> x, 1 -> y, 2 -> z, 3 -> t */
    int L[4];
    int P[4];          /* must be 1,1,1,1: the whole lattice sits on one GPU */
    int x_src[4];
    int Nprocs;
    int rank;
    double kappa;
    double csw;
    double mu;
    int twistSign;     /* +1 or -1 */
    double tol;
    double maxIter;
    double reliableDelta;
    int NsmearAPE;
    int NsmearGauss;
    double alphaAPE;
    double alphaGauss;
    int boundaryT;     /* +1 periodic, -1 antiperiodic */
    int Q_sq;

    LatticeInfo()
    {
        L[0] = 0; L[1] = 0; L[2] = 0; L[3] = 0;
        P[0] = 1; P[1] = 1; P[2] = 1; P[3] = 1;
        x_src[0] = 0; x_src[1] = 0; x_src[2] = 0; x_src[3] = 0;
        Nprocs = 1;
        rank = 0;
        kappa = 0.16;
        mu = 0.008;
        twistSign = +1;
        tol = 1e-1;
        maxIter = 1000;
        reliableDelta = 0.1;
        NsmearAPE = 20;
        NsmearGauss = 50;
        alphaAPE = 0.5;
        alphaGauss = 4.0;
        boundaryT = -1;
        Q_sq = 0;
        csw = 1.57551;
    }
};

#endif
