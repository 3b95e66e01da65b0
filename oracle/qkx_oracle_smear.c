/*
 * qkx_oracle_smear.c -- CPU restatement of the reference's source-preparation
 * steps: the SU(3) projection (lib/blas.cpp:6-422), APE smearing of the
 * spatial links (lib/lattice_util.cpp:950-1078), Gaussian smearing of a
 * spinor (lib/lattice_util.cpp:1080-1146), the plaquette
 * (lib/lattice_util.cpp:923-946) and the rotation of a propagator to the
 * physical basis (lib/lattice_util.cpp:2323-2453).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the library's smearing kernels.
 *
 * The operations are written in the reference's order of evaluation (sums
 * left to right, the same association of complex products) so that, built
 * with -ffp-contract=off, the results agree with the reference to the last
 * bits; the projection's formulas have to follow the reference's closed-form
 * eigensystem for the same reason.  Layout as in qkx_oracle.h (LEXIC sites,
 * U[iv][mu][row][col], psi[iv][spin][colour]).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "qkx_oracle.h"

typedef struct { double re, im; } cplx;

static cplx C(double r, double i) { cplx z = {r, i}; return z; }
static cplx cadd(cplx a, cplx b) { return C(a.re + b.re, a.im + b.im); }
static cplx csub(cplx a, cplx b) { return C(a.re - b.re, a.im - b.im); }
static cplx cmulz(cplx a, cplx b) { return C(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static cplx conjz(cplx a) { return C(a.re, -a.im); }
static double absz(cplx a) { return sqrt(a.re * a.re + a.im * a.im); }
static double argz(cplx a) { return atan2(a.im, a.re); }
/* principal power |z|^p e^{i p arg z}  (reference include/blas.h:320) */
static cplx cpowz(cplx z, double p)
{
    double r = pow(absz(z), p), t = argz(z) * p;
    return C(r * cos(t), r * sin(t));
}
/* Re(x / y) as the reference forms it (include/blas.h:314) */
static double cdivre(cplx x, cplx y)
{
    return (x.re * y.re + x.im * y.im) / (y.re * y.re + y.im * y.im);
}

/* ---- site neighbours ---------------------------------------------------- */
static long nbr(long iv, int mu, int dir, const int L[4])
{
    int c[4];
    long r = iv;
    for (int m = 0; m < 4; m++) { c[m] = (int)(r % L[m]); r /= L[m]; }
    c[mu] = (c[mu] + dir + L[mu]) % L[mu];
    return ((long)(c[3] * L[2] + c[2]) * L[1] + c[1]) * L[0] + c[0];
}

/* ---- 3x3 products in the reference's element order ---------------------- */
typedef cplx mat3[3][3];

static void ld3(mat3 a, const double *p)
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) a[i][j] = C(p[2 * (3 * i + j)], p[2 * (3 * i + j) + 1]);
}
static void st3(double *p, mat3 a)
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) { p[2 * (3 * i + j)] = a[i][j].re; p[2 * (3 * i + j) + 1] = a[i][j].im; }
}
/* a = b c   (include/blas.h:4: each element a sum over k = 0, 1, 2) */
static void mul(mat3 a, mat3 b, mat3 c)
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double re = 0, im = 0;
            re = b[i][0].re * c[0][j].re - b[i][0].im * c[0][j].im + b[i][1].re * c[1][j].re -
                 b[i][1].im * c[1][j].im + b[i][2].re * c[2][j].re - b[i][2].im * c[2][j].im;
            im = b[i][0].re * c[0][j].im + b[i][0].im * c[0][j].re + b[i][1].re * c[1][j].im +
                 b[i][1].im * c[1][j].re + b[i][2].re * c[2][j].im + b[i][2].im * c[2][j].re;
            a[i][j] = C(re, im);
        }
}
/* a = b c^dag   (include/blas.h:64) */
static void mul_dag(mat3 a, mat3 b, mat3 c)
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double re = b[i][0].re * c[j][0].re + b[i][0].im * c[j][0].im + b[i][1].re * c[j][1].re +
                        b[i][1].im * c[j][1].im + b[i][2].re * c[j][2].re + b[i][2].im * c[j][2].im;
            double im = -b[i][0].re * c[j][0].im + b[i][0].im * c[j][0].re - b[i][1].re * c[j][1].im +
                        b[i][1].im * c[j][1].re - b[i][2].re * c[j][2].im + b[i][2].im * c[j][2].re;
            a[i][j] = C(re, im);
        }
}
/* a = b^dag c   (include/blas.h:125) */
static void dag_mul(mat3 a, mat3 b, mat3 c)
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double re = b[0][i].re * c[0][j].re + b[0][i].im * c[0][j].im + b[1][i].re * c[1][j].re +
                        b[1][i].im * c[1][j].im + b[2][i].re * c[2][j].re + b[2][i].im * c[2][j].im;
            double im = b[0][i].re * c[0][j].im - b[0][i].im * c[0][j].re + b[1][i].re * c[1][j].im -
                        b[1][i].im * c[1][j].re + b[2][i].re * c[2][j].im - b[2][i].im * c[2][j].re;
            a[i][j] = C(re, im);
        }
}

/* ---- SU(3) projection of one matrix (lib/blas.cpp:32-420) ----------------
 * U = M (M^dag M)^{-1/2} exp(-i arg(det M)/3), with (M^dag M)^{-1/2} from the
 * closed-form (Cardano) eigensystem of the Hermitian H = M^dag M, followed by
 * a Gram-Schmidt pass over the first two columns and column 3 = conj(col1 x
 * col2).  When H is diagonal to 1e-8 (sum of |off-diagonal| moduli) row i of
 * M is only scaled by 1/sqrt(H_ii) (lib/blas.cpp:99-123). */
void orc_su3_project(double *m18)
{
    mat3 M, H, U, v, vr, HM;
    ld3(M, m18);
    /* det M by the rule of Sarrus (lib/blas.cpp:41-49) */
    cplx det = cadd(cadd(cmulz(M[0][0], cmulz(M[1][1], M[2][2])), cmulz(M[0][1], cmulz(M[1][2], M[2][0]))),
                    cmulz(M[0][2], cmulz(M[1][0], M[2][1])));
    det = csub(det, cadd(cadd(cmulz(M[0][2], cmulz(M[1][1], M[2][0])), cmulz(M[0][0], cmulz(M[1][2], M[2][1]))),
                         cmulz(M[0][1], cmulz(M[1][0], M[2][2]))));
    const double phase = argz(det) / 3.;

    /* H = M^dag M, column-by-column sums over k = 0, 1, 2 (lib/blas.cpp:54-72) */
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            cplx a0 = cmulz(conjz(M[0][i]), M[0][j]), a1 = cmulz(conjz(M[1][i]), M[1][j]),
                 a2 = cmulz(conjz(M[2][i]), M[2][j]);
            H[i][j] = C(a0.re + a1.re + a2.re, a0.im + a1.im + a2.im);
        }
    /* exact Hermiticity of the off-diagonal pairs (lib/blas.cpp:79-92) */
    for (int i = 0; i < 3; i++)
        for (int j = i + 1; j < 3; j++) {
            H[i][j] = C((H[i][j].re + H[j][i].re) / 2., (H[i][j].im - H[j][i].im) / 2.);
            H[j][i] = conjz(H[i][j]);
        }

    const double offd = absz(H[0][1]) + absz(H[0][2]) + absz(H[1][2]);
    if (offd <= 1e-08) {
        for (int i = 0; i < 3; i++) {
            double e = 1. / sqrt(H[i][i].re);
            for (int j = 0; j < 3; j++) U[i][j] = C(M[i][j].re * e, M[i][j].im * e);
        }
        st3(m18, U);
        return;
    }

    /* eigenvalues of the traceless part: e^3 + a e + b = 0 (lib/blas.cpp:129-178) */
    const double tr = (H[0][0].re + H[1][1].re + H[2][2].re) / 3.;
    H[0][0].re -= tr; H[1][1].re -= tr; H[2][2].re -= tr;
    const double h00 = H[0][0].re, h11 = H[1][1].re, h22 = H[2][2].re;
    const double n01 = cmulz(H[0][1], conjz(H[0][1])).re, n02 = cmulz(H[0][2], conjz(H[0][2])).re,
                 n12 = cmulz(H[1][2], conjz(H[1][2])).re;
    const double a = -(h22 * h22 - h00 * h11 + n01 + n02 + n12);
    cplx b, t1, t2;
    t1 = cmulz(H[0][1], cmulz(H[1][2], conjz(H[0][2])));
    b.re = -h00 * h11 * h22 + h22 * n01 - t1.re + h11 * n02;
    b.im = h22 * cmulz(H[0][1], conjz(H[0][1])).im - t1.im + h11 * cmulz(H[0][2], conjz(H[0][2])).im;
    t2 = cmulz(H[0][2], cmulz(conjz(H[0][1]), conjz(H[1][2])));
    b.re += h00 * n12 - t2.re;
    b.im += h00 * cmulz(H[1][2], conjz(H[1][2])).im - t2.im;

    const cplx bb = cmulz(b, b);
    const cplx disc = C(12. * a * a * a + 81. * bb.re, 81. * bb.im);
    const cplx w = cpowz(disc, 0.5);
    const cplx D = cpowz(C(-9. * b.re + w.re, -9. * b.im + w.im), 1. / 3.);
    const double r18 = pow(18., 1. / 3.), r12 = pow(12., 1. / 3.), r23 = pow(2. / 3., 1. / 3.);
    double e[3];
    e[0] = D.re / r18 - cdivre(C(a * r23, 0), D);
    e[1] = a * cdivre(C(1, sqrt(3.)), C(D.re * r12, D.im * r12)) - cmulz(C(1, -sqrt(3.)), D).re / (r18 * 2.);
    e[2] = -e[0] - e[1];
    for (int k = 0; k < 3; k++) e[k] += tr;
    H[0][0].re += tr; H[1][1].re += tr; H[2][2].re += tr;

    /* eigenvectors 1 and 2 from rows 1-2 of H - e (lib/blas.cpp:188-208): the
     * reference stores conj(x) for H x = e x, so sum_k conj(v_k) v_k^T / sqrt(e_k)
     * is H^{-1/2} */
    for (int k = 0; k < 2; k++) {
        const double ek = e[k];
        cplx p = cmulz(H[1][0], H[2][1]);
        v[k][0] = C(-(ek * H[2][0].re - H[2][0].re * H[1][1].re + p.re),
                    -(ek * H[2][0].im - H[2][0].im * H[1][1].re + p.im));
        cplx q = cmulz(H[2][0], H[0][1]);
        v[k][1] = C(-(q.re + ek * H[2][1].re - H[0][0].re * H[2][1].re),
                    -(q.im + ek * H[2][1].im - H[0][0].re * H[2][1].im));
        v[k][2] = C(-ek * ek + ek * H[0][0].re + cmulz(H[0][1], conjz(H[0][1])).re + ek * H[1][1].re -
                        H[0][0].re * H[1][1].re,
                    0.);
    }
    /* Gram-Schmidt of v1 against v0, normalise both (lib/blas.cpp:217-259) */
    double nrm = cmulz(v[0][0], conjz(v[0][0])).re + cmulz(v[0][1], conjz(v[0][1])).re +
                 cmulz(v[0][2], conjz(v[0][2])).re;
    cplx ov;
    {
        cplx o0 = cmulz(v[0][0], conjz(v[1][0])), o1 = cmulz(v[0][1], conjz(v[1][1])),
             o2 = cmulz(v[0][2], conjz(v[1][2]));
        ov = C((o0.re + o1.re + o2.re) / nrm, (o0.im + o1.im + o2.im) / nrm);
    }
    for (int j = 0; j < 3; j++) v[1][j] = csub(v[1][j], cmulz(ov, v[0][j]));
    nrm = 1. / sqrt(nrm);
    for (int j = 0; j < 3; j++) v[0][j] = C(v[0][j].re * nrm, v[0][j].im * nrm);
    nrm = cmulz(v[1][0], conjz(v[1][0])).re + cmulz(v[1][1], conjz(v[1][1])).re +
          cmulz(v[1][2], conjz(v[1][2])).re;
    nrm = 1. / sqrt(nrm);
    for (int j = 0; j < 3; j++) v[1][j] = C(v[1][j].re * nrm, v[1][j].im * nrm);
    /* v2 = conj(v0 x v1) (lib/blas.cpp:265-275) */
    {
        cplx x;
        x = csub(cmulz(v[0][1], v[1][2]), cmulz(v[0][2], v[1][1])); v[2][0] = conjz(x);
        x = csub(cmulz(v[0][2], v[1][0]), cmulz(v[0][0], v[1][2])); v[2][1] = conjz(x);
        x = csub(cmulz(v[0][0], v[1][1]), cmulz(v[0][1], v[1][0])); v[2][2] = conjz(x);
    }
    /* vr_k = v_k e^{-i phase} / sqrt(e_k)  (lib/blas.cpp:285-309) */
    for (int k = 0; k < 3; k++) {
        double de = 1. / sqrt(e[k]);
        cplx s = C(de * cos(phase), -de * sin(phase));
        for (int j = 0; j < 3; j++) vr[k][j] = cmulz(s, v[k][j]);
    }
    /* U = (M conj(v)^T) vr  (lib/blas.cpp:312-356) */
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++) {
            cplx a0 = cmulz(M[i][0], conjz(v[k][0])), a1 = cmulz(M[i][1], conjz(v[k][1])),
                 a2 = cmulz(M[i][2], conjz(v[k][2]));
            HM[i][k] = C(a0.re + a1.re + a2.re, a0.im + a1.im + a2.im);
        }
    for (int i = 0; i < 3; i++)
        for (int l = 0; l < 3; l++) {
            cplx a0 = cmulz(HM[i][0], vr[0][l]), a1 = cmulz(HM[i][1], vr[1][l]), a2 = cmulz(HM[i][2], vr[2][l]);
            U[i][l] = C(a0.re + a1.re + a2.re, a0.im + a1.im + a2.im);
        }
    /* column Gram-Schmidt: the reference subtracts (sum_i U_i0 conj U_i1)/|col0|^2
     * times column 0 from column 1 (lib/blas.cpp:364-381) */
    nrm = cmulz(U[0][0], conjz(U[0][0])).re + cmulz(U[1][0], conjz(U[1][0])).re +
          cmulz(U[2][0], conjz(U[2][0])).re;
    {
        cplx o0 = cmulz(U[0][0], conjz(U[0][1])), o1 = cmulz(U[1][0], conjz(U[1][1])),
             o2 = cmulz(U[2][0], conjz(U[2][1]));
        ov = C((o0.re + o1.re + o2.re) / nrm, (o0.im + o1.im + o2.im) / nrm);
    }
    for (int i = 0; i < 3; i++) U[i][1] = csub(U[i][1], cmulz(ov, U[i][0]));
    nrm = 1. / sqrt(nrm);
    for (int i = 0; i < 3; i++) U[i][0] = C(U[i][0].re * nrm, U[i][0].im * nrm);
    nrm = cmulz(U[0][1], conjz(U[0][1])).re + cmulz(U[1][1], conjz(U[1][1])).re +
          cmulz(U[2][1], conjz(U[2][1])).re;
    nrm = 1. / sqrt(nrm);
    for (int i = 0; i < 3; i++) U[i][1] = C(U[i][1].re * nrm, U[i][1].im * nrm);
    /* column 3 = conj(col1 x col2) (lib/blas.cpp:407-414) */
    {
        cplx x;
        x = csub(cmulz(U[1][0], U[2][1]), cmulz(U[2][0], U[1][1])); U[0][2] = conjz(x);
        x = csub(cmulz(U[2][0], U[0][1]), cmulz(U[0][0], U[2][1])); U[1][2] = conjz(x);
        x = csub(cmulz(U[0][0], U[1][1]), cmulz(U[1][0], U[0][1])); U[2][2] = conjz(x);
    }
    st3(m18, U);
}

/* Projection of the spatial links of a whole field (lib/blas.cpp:35: mu < 3). */
void orc_project_su3_field(double *U, const int L[4])
{
    long V = (long)L[0] * L[1] * L[2] * L[3];
    for (long iv = 0; iv < V; iv++)
        for (int mu = 0; mu < 3; mu++) orc_su3_project(U + (iv * 4 + mu) * 18);
}

#define LNK(U, iv, mu) ((U) + ((long)(iv) * 4 + (mu)) * 18)

/* One APE step on the spatial links: out_mu(x) = P[ U_mu(x) + alpha (sum of the
 * forward staples over nu != mu, then the backward ones) ]
 *   forward  U_nu(x) U_mu(x+nu) U_nu(x+mu)^dag          (lattice_util.cpp:988-989)
 *   backward U_nu(x-nu)^dag U_mu(x-nu) U_nu(x-nu+mu)    (:983-984, added at :1029)
 * The temporal links of out are left as they are. */
static void ape_step(double *out, const double *in, const int L[4], double alpha)
{
    long V = (long)L[0] * L[1] * L[2] * L[3];
    for (long iv = 0; iv < V; iv++)
        for (int mu = 0; mu < 3; mu++) {
            mat3 S, A, B, T, Uf;
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) S[i][j] = C(0, 0);
            long xpm = nbr(iv, mu, +1, L);
            for (int nu = 0; nu < 3; nu++) {
                if (nu == mu) continue;
                ld3(A, LNK(in, iv, nu));
                ld3(B, LNK(in, nbr(iv, nu, +1, L), mu));
                mul(T, A, B);
                ld3(A, LNK(in, xpm, nu));
                mul_dag(Uf, T, A);
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) S[i][j] = cadd(S[i][j], Uf[i][j]);
            }
            for (int nu = 0; nu < 3; nu++) {
                if (nu == mu) continue;
                long xm = nbr(iv, nu, -1, L);
                ld3(A, LNK(in, xm, mu));
                ld3(B, LNK(in, nbr(xm, mu, +1, L), nu));
                mul(T, A, B);
                ld3(A, LNK(in, xm, nu));
                dag_mul(Uf, A, T);
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) S[i][j] = cadd(S[i][j], Uf[i][j]);
            }
            ld3(A, LNK(in, iv, mu));
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++)
                    S[i][j] = C(alpha * S[i][j].re + A[i][j].re, alpha * S[i][j].im + A[i][j].im);
            st3(LNK(out, iv, mu), S);
            orc_su3_project(LNK(out, iv, mu));
        }
}

/* Gauge::APE_smearing (lattice_util.cpp:950-1078): out = nsmear APE steps of
 * in on the spatial links, temporal links of in.  nsmear = 0 leaves out as it
 * is (lattice_util.cpp:956).  in is not modified. */
void orc_ape_smear(double *out, const double *in, const int L[4], int nsmear, double alpha)
{
    if (nsmear == 0) return;
    long n = (long)L[0] * L[1] * L[2] * L[3] * 72;
    double *a = malloc(n * sizeof(double)), *b = malloc(n * sizeof(double));
    memcpy(a, in, n * sizeof(double));
    memcpy(b, in, n * sizeof(double));
    for (int it = 0; it < nsmear; it++) {
        ape_step(b, a, L, alpha);
        double *t = a; a = b; b = t;
    }
    memcpy(out, a, n * sizeof(double));
    free(a);
    free(b);
}

/* Vector::Gaussian_smearing (lattice_util.cpp:1080-1146): nsmear steps of
 *   out(x) = (in(x) + alpha sum_{mu<3} [U_mu(x) in(x+mu) + U_mu(x-mu)^dag in(x-mu)]) / (1 + 6 alpha)
 * with the hop sum accumulated as in lattice_util.cpp:1118-1131.  nsmear = 0
 * leaves out as it is. */
void orc_gauss_smear(double *out, const double *in, const double *U, const int L[4], int nsmear, double alpha)
{
    if (nsmear == 0) return;
    long V = (long)L[0] * L[1] * L[2] * L[3];
    const double normalize = 1. / (1 + 6 * alpha);
    double *a = malloc(V * 24 * sizeof(double)), *b = malloc(V * 24 * sizeof(double));
    memcpy(a, in, V * 24 * sizeof(double));
    for (int it = 0; it < nsmear; it++) {
        for (long iv = 0; iv < V; iv++) {
            double tot[24];
            for (int k = 0; k < 24; k++) tot[k] = 0;
            for (int mu = 0; mu < 3; mu++) {
                const double *uf = LNK(U, iv, mu), *pf = a + nbr(iv, mu, +1, L) * 24;
                long xm = nbr(iv, mu, -1, L);
                const double *ub = LNK(U, xm, mu), *pb = a + xm * 24;
                double up[24], ud[24];
                for (int s = 0; s < 4; s++)
                    for (int i = 0; i < 3; i++) {
                        /* (U psi)_i and (U^dag psi)_i as one left-to-right chain of the
                         * six real products each (include/blas.h:203-261) */
                        const double *u = uf + 6 * i, *y = pf + 6 * s, *z = pb + 6 * s;
                        const double *w0 = ub + 2 * i, *w1 = ub + 6 + 2 * i, *w2 = ub + 12 + 2 * i;
                        up[2 * (3 * s + i)] =
                            u[0] * y[0] - u[1] * y[1] + u[2] * y[2] - u[3] * y[3] + u[4] * y[4] - u[5] * y[5];
                        up[2 * (3 * s + i) + 1] =
                            u[0] * y[1] + u[1] * y[0] + u[2] * y[3] + u[3] * y[2] + u[4] * y[5] + u[5] * y[4];
                        ud[2 * (3 * s + i)] =
                            w0[0] * z[0] + w0[1] * z[1] + w1[0] * z[2] + w1[1] * z[3] + w2[0] * z[4] + w2[1] * z[5];
                        ud[2 * (3 * s + i) + 1] =
                            w0[0] * z[1] - w0[1] * z[0] + w1[0] * z[3] - w1[1] * z[2] + w2[0] * z[5] - w2[1] * z[4];
                    }
                for (int k = 0; k < 24; k++) tot[k] += up[k] + ud[k];
            }
            double *o = b + iv * 24;
            const double *x = a + iv * 24;
            for (int k = 0; k < 24; k++) o[k] = normalize * (x[k] + alpha * tot[k]);
        }
        double *t = a; a = b; b = t;
    }
    memcpy(out, a, V * 24 * sizeof(double));
    free(a);
    free(b);
}

/* Gauge::calculatePlaquette (lattice_util.cpp:923-946): mean over sites and the
 * six planes mu < nu of Re tr[U_mu(x) U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag] / 3. */
double orc_plaquette(const double *U, const int L[4])
{
    long V = (long)L[0] * L[1] * L[2] * L[3];
    double acc = 0;
    for (long iv = 0; iv < V; iv++)
        for (int mu = 0; mu < 3; mu++)
            for (int nu = mu + 1; nu < 4; nu++) {
                mat3 A, B, P, T;
                ld3(A, LNK(U, iv, mu));
                ld3(B, LNK(U, nbr(iv, mu, +1, L), nu));
                mul(P, A, B);
                ld3(A, LNK(U, nbr(iv, nu, +1, L), mu));
                mul_dag(T, P, A);
                ld3(A, LNK(U, iv, nu));
                mul_dag(P, T, A);
                acc += P[0][0].re + P[1][1].re + P[2][2].re;
            }
    return acc / (V * 3.0 * 6.0);
}

/* Propagator::rotateToPhysicalBasePlus/Minus (lattice_util.cpp:2323-2453): per
 * site and colour pair, S -> 0.5 R S R with R = 1 + sign i g5 (UKQCD g5:
 * R[a][a] = 1, R[0][2] = R[1][3] = R[2][0] = R[3][1] = sign i), the products
 * summed over mu' then nu' as the reference's loops do.  M[V][4][4][3][3]. */
void orc_prop_rotate(double *M, long V, int sign)
{
    cplx R[4][4];
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) R[a][b] = C(a == b ? 1. : 0., 0.);
    R[0][2] = R[1][3] = R[2][0] = R[3][1] = C(0., (double)sign);
    for (long iv = 0; iv < V; iv++) {
        double *m = M + iv * 288;
        for (int c1 = 0; c1 < 3; c1++)
            for (int c2 = 0; c2 < 3; c2++) {
                cplx S[4][4], T[4][4];
                for (int mu = 0; mu < 4; mu++)
                    for (int nu = 0; nu < 4; nu++) {
                        const double *e = m + 2 * (((mu * 4 + nu) * 3 + c1) * 3 + c2);
                        S[mu][nu] = C(e[0], e[1]);
                        T[mu][nu] = C(0., 0.);
                    }
                for (int mu = 0; mu < 4; mu++)
                    for (int mp = 0; mp < 4; mp++)
                        for (int np = 0; np < 4; np++)
                            for (int nu = 0; nu < 4; nu++)
                                T[mu][nu] = cadd(T[mu][nu], cmulz(cmulz(R[mu][mp], S[mp][np]), R[np][nu]));
                for (int mu = 0; mu < 4; mu++)
                    for (int nu = 0; nu < 4; nu++) {
                        double *e = m + 2 * (((mu * 4 + nu) * 3 + c1) * 3 + c2);
                        e[0] = 0.5 * T[mu][nu].re;
                        e[1] = 0.5 * T[mu][nu].im;
                    }
            }
    }
}
