/*
 * qkx_oracle.c -- CPU restatement of the reference FP64 twisted-mass path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load this file's
 * library.  Plain C99, one thread, reference AoS layout, same operation order
 * as the reference so that results agree to rounding.
 *
 * Each function cites the reference lines it restates (paths relative to the
 * reference repository root).
 */
#include <complex.h>
#include <stdlib.h>
#include <string.h>
#include "qkx_oracle.h"

typedef double complex cplx;

#define SP(iv, s, c) ((((long)(iv)) * 4 + (s)) * 3 + (c))
#define GL(iv, mu, a, b) ((((((long)(iv)) * 4 + (mu)) * 3 + (a)) * 3 + (b)))

/* ---- exact unit-coefficient arithmetic ------------------------------------
 * The reference macros write every projection/reconstruction as explicit
 * re/im adds and subtracts (include/hopping_term.h:30-343).  Multiplying by
 * one of {0, +1, -1, +i, -i, 2} with these helpers produces exactly those
 * floating-point operations. */
enum { C0 = 0, CP1, CM1, CPI, CMI, C2 };

static inline cplx unit_mul(int k, cplx z)
{
    double re = creal(z), im = cimag(z);
    switch (k) {
    case CP1: return CMPLX(re, im);
    case CM1: return CMPLX(-re, -im);
    case CPI: return CMPLX(-im, re);
    case CMI: return CMPLX(im, -re);
    case C2:  return CMPLX(2 * re, 2 * im);
    default:  return CMPLX(0.0, 0.0);
    }
}

/* Half-spinor projector (h = P psi) and reconstruction (R += Q xi) for one
 * direction and one sign.  'minus' = the (1 - gamma_mu) projector used for the
 * forward hop of H; 'plus' = (1 + gamma_mu).
 *   h_k = a_k * psi[sa_k] + b_k * psi[sb_k]         k = 0,1
 *   R_s += q_s * xi[kq_s]                           s = 0..3 */
typedef struct {
    int sa[2], a[2], sb[2], b[2];
    int kq[4], q[4];
} proj_t;

/* [basis][mu][0 = minus (1-g), 1 = plus (1+g)] */
static const proj_t PROJ[2][4][2] = {
  { /* UKQCD basis: include/hopping_term.h */
    { /* X: PROJ_MINUS_X :30-43, COLLECT_MINUS_X :184-205 */
      { {0,1},{CP1,CP1},{3,2},{CMI,CMI}, {0,1,1,0},{CP1,CP1,CPI,CPI} },
      /* PROJ_PLUS_X :46-59, COLLECT_PLUS_X :207-228 */
      { {0,1},{CP1,CP1},{3,2},{CPI,CPI}, {0,1,1,0},{CP1,CP1,CMI,CMI} } },
    { /* Y: PROJ_MINUS_Y :63-76, COLLECT_MINUS_Y :231-252 */
      { {0,1},{CP1,CP1},{3,2},{CM1,CP1}, {0,1,1,0},{CP1,CP1,CP1,CM1} },
      /* PROJ_PLUS_Y :79-92, COLLECT_PLUS_Y :255-276 */
      { {0,1},{CP1,CP1},{3,2},{CP1,CM1}, {0,1,1,0},{CP1,CP1,CM1,CP1} } },
    { /* Z: PROJ_MINUS_Z :95-108, COLLECT_MINUS_Z :278-299 */
      { {0,1},{CP1,CP1},{2,3},{CMI,CPI}, {0,1,0,1},{CP1,CP1,CPI,CMI} },
      /* PROJ_PLUS_Z :111-124, COLLECT_PLUS_Z :302-323 */
      { {0,1},{CP1,CP1},{2,3},{CPI,CMI}, {0,1,0,1},{CP1,CP1,CMI,CPI} } },
    { /* T: PROJ_MINUS_T :127-140 (h = 2 psi_{2,3}), COLLECT_MINUS_T :326-333 */
      { {2,3},{C2,C2},{0,0},{C0,C0}, {0,0,0,1},{C0,C0,CP1,CP1} },
      /* PROJ_PLUS_T :143-156 (h = 2 psi_{0,1}), COLLECT_PLUS_T :335-342 */
      { {0,1},{C2,C2},{0,0},{C0,C0}, {0,1,0,0},{CP1,CP1,C0,C0} } } },
  { /* chiral basis: include/hopping_term_chiral.h */
    { /* X: PROJ_MINUS_X_CHIRAL :1-13, COLLECT_MINUS_X_CHIRAL :136-157 */
      { {0,1},{CP1,CP1},{3,2},{CPI,CPI}, {0,1,1,0},{CP1,CP1,CMI,CMI} },
      /* PROJ_PLUS_X_CHIRAL :15-27, COLLECT_PLUS_X_CHIRAL :113-134 */
      { {0,1},{CP1,CP1},{3,2},{CMI,CMI}, {0,1,1,0},{CP1,CP1,CPI,CPI} } },
    { /* Y: PROJ_MINUS_Y_CHIRAL :29-41, COLLECT_MINUS_Y_CHIRAL :184-205 */
      { {0,1},{CP1,CP1},{3,2},{CP1,CM1}, {0,1,1,0},{CP1,CP1,CM1,CP1} },
      /* PROJ_PLUS_Y_CHIRAL :43-55, COLLECT_PLUS_Y_CHIRAL :160-181 */
      { {0,1},{CP1,CP1},{3,2},{CM1,CP1}, {0,1,1,0},{CP1,CP1,CP1,CM1} } },
    { /* Z: PROJ_MINUS_Z_CHIRAL :57-69, COLLECT_MINUS_Z_CHIRAL :231-252 */
      { {0,1},{CP1,CP1},{2,3},{CPI,CMI}, {0,1,0,1},{CP1,CP1,CMI,CPI} },
      /* PROJ_PLUS_Z_CHIRAL :71-83, COLLECT_PLUS_Z_CHIRAL :207-228 */
      { {0,1},{CP1,CP1},{2,3},{CMI,CPI}, {0,1,0,1},{CP1,CP1,CPI,CMI} } },
    { /* T: PROJ_MINUS_T_CHIRAL :85-97, COLLECT_MINUS_T_CHIRAL :278-299 */
      { {0,1},{CP1,CP1},{2,3},{CP1,CP1}, {0,1,0,1},{CP1,CP1,CP1,CP1} },
      /* PROJ_PLUS_T_CHIRAL :99-111, COLLECT_PLUS_T_CHIRAL :255-276 */
      { {0,1},{CP1,CP1},{2,3},{CM1,CM1}, {0,1,0,1},{CP1,CP1,CM1,CM1} } } }
};

/* ---- geometry: lib/lattice_util.cpp:86-99 (periodic neighbour tables),
 * :1148-1155 (getOddBit).  Computed on the fly from coordinates. */
static inline void coords(long iv, const int L[4], int x[4])
{
    x[0] = (int)(iv % L[0]); iv /= L[0];
    x[1] = (int)(iv % L[1]); iv /= L[1];
    x[2] = (int)(iv % L[2]); iv /= L[2];
    x[3] = (int)iv;
}
static inline long lexic(const int x[4], const int L[4])
{
    return (((long)x[3] * L[2] + x[2]) * L[1] + x[1]) * L[0] + x[0];
}
static inline long neighbour(long iv, const int L[4], int mu, int dir)
{
    int x[4];
    coords(iv, L, x);
    x[mu] = (x[mu] + dir + L[mu]) % L[mu];
    return lexic(x, L);
}
static inline int odd_bit(long iv, const int L[4])
{
    int x[4];
    coords(iv, L, x);
    return (x[0] + x[1] + x[2] + x[3]) & 1;
}
static inline long volume(const int L[4]) { return (long)L[0] * L[1] * L[2] * L[3]; }

/* One hop: project the neighbour spinor, multiply by U (or U^dag), reconstruct
 * and accumulate into R.  APPLY_LINK include/hopping_term.h:162-169:
 * xi[c1] = sum_c2 U[c1][c2] h[c2]; APPLY_LINK_DAG :171-178 with conj(U[c2][c1]). */
static void hop_one(cplx R[12], const cplx *psi, const cplx *U, long nb,
                    long link_site, int mu, const proj_t *P, int link_dag)
{
    cplx h[2][3], xi[2][3];
    for (int k = 0; k < 2; k++)
        for (int c = 0; c < 3; c++)
            h[k][c] = unit_mul(P->a[k], psi[SP(nb, P->sa[k], c)])
                    + unit_mul(P->b[k], psi[SP(nb, P->sb[k], c)]);
    /* PROJ_*_T writes h = 2*psi directly (no second term); unit_mul(C0,..)
     * adds an exact +0 which leaves every finite value unchanged. */
    const cplx *u = U + GL(link_site, mu, 0, 0);
    for (int k = 0; k < 2; k++)
        for (int c1 = 0; c1 < 3; c1++) {
            if (!link_dag)
                xi[k][c1] = u[c1 * 3 + 0] * h[k][0] + u[c1 * 3 + 1] * h[k][1]
                          + u[c1 * 3 + 2] * h[k][2];
            else
                xi[k][c1] = conj(u[0 * 3 + c1]) * h[k][0] + conj(u[1 * 3 + c1]) * h[k][1]
                          + conj(u[2 * 3 + c1]) * h[k][2];
        }
    for (int s = 0; s < 4; s++) {
        if (P->q[s] == C0) continue;
        for (int c = 0; c < 3; c++)
            R[s * 3 + c] += unit_mul(P->q[s], xi[P->kq[s]][c]);
    }
}

/* applyDslash lib/lattice_util.cpp:1158-1213, applyDslash_chiral :1215-1270,
 * applyDslash_xx :1272-1344, applyDslash_xx_chiral :1346-1418,
 * applyDslashDag :1643-1698.  Direction order +x,-x,+y,-y,+z,-z,+t,-t.
 * H:     forward (1-g_mu) U_mu(x) psi(x+mu), backward (1+g_mu) U_mu^dag(x-mu) psi(x-mu)
 * H^dag: the two projectors swapped (DAG_PROJ_* :348-473). */
void orc_hop(double *out_, const double *psi_, const double *U_, const int L[4],
             int basis, int dagger, int parity)
{
    cplx *out = (cplx *)out_;
    const cplx *psi = (const cplx *)psi_;
    const cplx *U = (const cplx *)U_;
    long V = volume(L);
    for (long iv = 0; iv < V; iv++) {
        if (parity == ORC_OE || parity == ORC_EO) {
            int ob = odd_bit(iv, L);
            if ((parity == ORC_OE && !ob) || (parity == ORC_EO && ob)) {
                for (int k = 0; k < 12; k++) out[iv * 12 + k] = 0;
                continue;
            }
        }
        cplx R[12];
        for (int k = 0; k < 12; k++) R[k] = 0;
        for (int mu = 0; mu < 4; mu++) {
            long fw = neighbour(iv, L, mu, +1);
            long bw = neighbour(iv, L, mu, -1);
            int pf = dagger ? 1 : 0, pb = dagger ? 0 : 1;
            hop_one(R, psi, U, fw, iv, mu, &PROJ[basis][mu][pf], 0);
            hop_one(R, psi, U, bw, bw, mu, &PROJ[basis][mu][pb], 1);
        }
        for (int k = 0; k < 12; k++) out[iv * 12 + k] = R[k];
    }
}

/* gamma_5 in each basis: UKQCD = spin swap 0<->2, 1<->3 (lib/lattice_util.cpp:
 * 1778-1798); chiral = diag(1,1,-1,-1) (:1800-1820). */
static inline void g5_site(cplx phi[12], const cplx *p, int basis)
{
    for (int c = 0; c < 3; c++) {
        if (basis == ORC_UKQCD) {
            phi[0 * 3 + c] = p[2 * 3 + c];
            phi[1 * 3 + c] = p[3 * 3 + c];
            phi[2 * 3 + c] = p[0 * 3 + c];
            phi[3 * 3 + c] = p[1 * 3 + c];
        } else {
            phi[0 * 3 + c] = p[0 * 3 + c];
            phi[1 * 3 + c] = p[1 * 3 + c];
            phi[2 * 3 + c] = -p[2 * 3 + c];
            phi[3 * 3 + c] = -p[3 * 3 + c];
        }
    }
}

/* twistCoeff = i * 2*kappa*mu*twistSign (lib/lattice_util.cpp:1424-1425) */
static inline double twist_a(double kappa, double mu, int s) { return 2 * kappa * mu * s; }

/* applyTwistAddDslash :1420-1442, _chiral :1444-1466,
 * applyTwistDagAddDslashDag :1701-1723 (coefficient -i a). */
void orc_twist_add_dslash(double *out_, const double *d_, const double *psi_,
                          const int L[4], double kappa, double mu, int twistSign,
                          int basis, int dagger)
{
    cplx *out = (cplx *)out_;
    const cplx *d = (const cplx *)d_, *psi = (const cplx *)psi_;
    double a = dagger ? -2 * kappa * mu * twistSign : twist_a(kappa, mu, twistSign);
    cplx tc = CMPLX(0.0, a);
    long V = volume(L);
    for (long iv = 0; iv < V; iv++) {
        cplx phi[12];
        g5_site(phi, psi + iv * 12, basis);
        for (int k = 0; k < 12; k++)
            out[iv * 12 + k] = (-kappa) * d[iv * 12 + k] + psi[iv * 12 + k] + tc * phi[k];
    }
}

/* applyTwistAddDslash_PC :1468-1508, _PC_chiral :1510-1550 */
void orc_twist_add_dslash_pc(double *out_, const double *d_, const double *psi_,
                             const int L[4], double kappa, double mu, int twistSign,
                             int basis, int parity)
{
    cplx *out = (cplx *)out_;
    const cplx *d = (const cplx *)d_, *psi = (const cplx *)psi_;
    cplx tc = CMPLX(0.0, twist_a(kappa, mu, twistSign));
    long V = volume(L);
    for (long iv = 0; iv < V; iv++) {
        int ob = odd_bit(iv, L);
        if ((parity == ORC_OO && !ob) || (parity == ORC_EE && ob)) {
            for (int k = 0; k < 12; k++) out[iv * 12 + k] = 0;
            continue;
        }
        cplx phi[12];
        g5_site(phi, psi + iv * 12, basis);
        for (int k = 0; k < 12; k++)
            out[iv * 12 + k] = (-kappa * kappa) * d[iv * 12 + k]
                             + (psi[iv * 12 + k] + tc * phi[k]);
    }
}

/* applyDeltaAddDslash :1552-1559 */
void orc_delta_add_dslash(double *out_, const double *d_, const double *psi_,
                          const int L[4], double kappa)
{
    cplx *out = (cplx *)out_;
    const cplx *d = (const cplx *)d_, *psi = (const cplx *)psi_;
    long n = volume(L) * 12;
    for (long k = 0; k < n; k++) out[k] = psi[k] - kappa * d[k];
}

/* applyTwist :1563-1585 (+ia), applyTwistDag :1726-1748 (-ia),
 * applyTwistInv :1588-1613 and _chiral :1615-1640 ((psi - ia g5 psi)/(1+a^2)),
 * applyTwistDagInv :1750-1775 ((psi + ia g5 psi)/(1+a^2)). */
void orc_twist(double *out_, const double *psi_, const int L[4], double kappa,
               double mu, int twistSign, int basis, int kind)
{
    cplx *out = (cplx *)out_;
    const cplx *psi = (const cplx *)psi_;
    int dag = (kind == ORC_TWIST_DAG || kind == ORC_TWIST_DAG_INV);
    int inv = (kind == ORC_TWIST_INV || kind == ORC_TWIST_DAG_INV);
    double a = dag ? -2 * kappa * mu * twistSign : twist_a(kappa, mu, twistSign);
    cplx tc = CMPLX(0.0, a);
    double normalize = (1 + (2 * kappa * mu) * (2 * kappa * mu));
    long V = volume(L);
    for (long iv = 0; iv < V; iv++) {
        cplx phi[12];
        g5_site(phi, psi + iv * 12, basis);
        for (int k = 0; k < 12; k++) {
            if (!inv) {
                out[iv * 12 + k] = psi[iv * 12 + k] + tc * phi[k];
            } else {
                cplx v = psi[iv * 12 + k] - tc * phi[k];
                out[iv * 12 + k] = CMPLX(creal(v) / normalize, cimag(v) / normalize);
            }
        }
    }
}

/* applyGamma5 :1778-1798, applyGamma5_chiral :1800-1820 */
void orc_gamma5(double *out_, const double *psi_, const int L[4], int basis)
{
    cplx *out = (cplx *)out_;
    const cplx *psi = (const cplx *)psi_;
    long V = volume(L);
    for (long iv = 0; iv < V; iv++) {
        cplx phi[12];
        g5_site(phi, psi + iv * 12, basis);
        for (int k = 0; k < 12; k++) out[iv * 12 + k] = phi[k];
    }
}

/* tm :1925-1928, tm_chiral :1930-1933, tmDag :1935-1938, tmDagtm :1940-1947,
 * Wilson :1949-1952.  The reference tmDagtm has no chiral form; basis must be
 * UKQCD for TM_DAG / TM_DAG_TM (as in the reference). */
void orc_tm(double *out, const double *psi, const double *U, const int L[4],
            double kappa, double mu, int twistSign, int basis, int op)
{
    long n = volume(L) * 24;
    if (op == ORC_TM) {
        orc_hop(out, psi, U, L, basis, 0, ORC_ALL);
        orc_twist_add_dslash(out, out, psi, L, kappa, mu, twistSign, basis, 0);
    } else if (op == ORC_TM_DAG) {
        orc_hop(out, psi, U, L, basis, 1, ORC_ALL);
        orc_twist_add_dslash(out, out, psi, L, kappa, mu, twistSign, basis, 1);
    } else if (op == ORC_TM_DAG_TM) {
        double *tmp = (double *)malloc(n * sizeof(double));
        orc_hop(tmp, psi, U, L, basis, 0, ORC_ALL);
        orc_twist_add_dslash(tmp, tmp, psi, L, kappa, mu, twistSign, basis, 0);
        orc_hop(out, tmp, U, L, basis, 1, ORC_ALL);
        orc_twist_add_dslash(out, out, tmp, L, kappa, mu, twistSign, basis, 1);
        free(tmp);
    } else { /* Wilson */
        orc_hop(out, psi, U, L, basis, 0, ORC_ALL);
        orc_delta_add_dslash(out, out, psi, L, kappa);
    }
}

/* tm_xx :2001-2022 and tm_xx_chiral :2037-2058:
 *   EE: tmp = H(psi) on odd sites; this = A^-1 tmp; tmp = H(this) on even sites;
 *       out = -kappa^2 tmp + A psi on even sites (zero on odd).   OO mirrors it.
 * tm_xxDagTm_xx :2024-2035, _chiral :2060-2071 (normal=1):
 *   out = g5 tm_xx(s=-twistSign)[ g5 tm_xx(s)[psi] ] */
static void tm_xx_once(double *out, const double *psi, const double *U, const int L[4],
                       double kappa, double mu, int s, int basis, int parity)
{
    long n = volume(L) * 24;
    double *tmp = (double *)malloc(n * sizeof(double));
    memset(out, 0, n * sizeof(double));
    int first = (parity == ORC_EE) ? ORC_OE : ORC_EO;
    int second = (parity == ORC_EE) ? ORC_EO : ORC_OE;
    orc_hop(tmp, psi, U, L, basis, 0, first);
    orc_twist(out, tmp, L, kappa, mu, s, basis, ORC_TWIST_INV);
    orc_hop(tmp, out, U, L, basis, 0, second);
    orc_twist_add_dslash_pc(out, tmp, psi, L, kappa, mu, s, basis, parity);
    free(tmp);
}

void orc_tm_xx(double *out, const double *psi, const double *U, const int L[4],
               double kappa, double mu, int twistSign, int basis, int parity, int normal)
{
    if (!normal) {
        tm_xx_once(out, psi, U, L, kappa, mu, twistSign, basis, parity);
        return;
    }
    long n = volume(L) * 24;
    double *tmp = (double *)malloc(n * sizeof(double));
    tm_xx_once(tmp, psi, U, L, kappa, mu, twistSign, basis, parity);
    orc_gamma5(out, tmp, L, basis);
    tm_xx_once(tmp, out, U, L, kappa, mu, -twistSign, basis, parity);
    orc_gamma5(out, tmp, L, basis);
    free(tmp);
}

/* reCdotXdagY lib/blas_qkxTM.cpp:4-28: complex accumulation in (iv,mu,ic)
 * order, real part returned. */
double orc_redot(const double *x_, const double *y_, long V)
{
    const cplx *x = (const cplx *)x_, *y = (const cplx *)y_;
    cplx res = 0;
    long n = V * 12;
    for (long k = 0; k < n; k++) res = res + conj(x[k]) * y[k];
    return creal(res);
}
/* X_eq_X_p_aY :31-42 */
void orc_axpy(double *x_, double a, const double *y_, long V)
{
    cplx *x = (cplx *)x_;
    const cplx *y = (const cplx *)y_;
    long n = V * 12;
    for (long k = 0; k < n; k++) x[k] = x[k] + a * y[k];
}
/* X_eq_aX_p_Y :45-56 */
void orc_xpay(double *x_, double a, const double *y_, long V)
{
    cplx *x = (cplx *)x_;
    const cplx *y = (const cplx *)y_;
    long n = V * 12;
    for (long k = 0; k < n; k++) x[k] = a * x[k] + y[k];
}
/* Vector::norm2 lib/lattice_util.cpp:464-477 */
double orc_norm2(const double *x_, long V)
{
    const cplx *x = (const cplx *)x_;
    double s = 0;
    long n = V * 12;
    for (long k = 0; k < n; k++) s += creal(x[k]) * creal(x[k]) + cimag(x[k]) * cimag(x[k]);
    return s;
}

/* CG::operator() lib/solvers.cpp:16-81: CG on D^dag D with r0 = b (x0 assumed
 * zero), stop on the iterated residual r2 <= tol^2 |b|^2 or k == maxIter. */
int orc_cg(double *x, const double *b, const double *U, const int L[4],
           double kappa, double mu, int twistSign, double tol, int maxIter,
           double *r2_final)
{
    long V = volume(L), n = V * 24;
    double *r = (double *)malloc(n * sizeof(double));
    double *p = (double *)malloc(n * sizeof(double));
    double *Ap = (double *)malloc(n * sizeof(double));
    memcpy(r, b, n * sizeof(double));
    memcpy(p, r, n * sizeof(double));
    double r2 = orc_redot(r, r, V);
    double src_norm = orc_redot(b, b, V);
    double stop = src_norm * tol * tol;
    int k = 0;
    while ((r2 > stop) && (k < maxIter)) {
        orc_tm(Ap, p, U, L, kappa, mu, twistSign, ORC_UKQCD, ORC_TM_DAG_TM);
        double pAp = orc_redot(p, Ap, V);
        double alpha = r2 / pAp;
        double r2_old = r2;
        orc_axpy(x, alpha, p, V);
        orc_axpy(r, -alpha, Ap, V);
        r2 = orc_redot(r, r, V);
        double beta = r2 / r2_old;
        orc_xpay(p, beta, r, V);
        k++;
    }
    if (r2_final) *r2_final = r2;
    free(r); free(p); free(Ap);
    return k;
}

/* Gauge::applyBoundaryCondition lib/lattice_util.cpp:708-724: with
 * boundaryT == -1 the t-links on the last time slice change sign. */
void orc_apply_bc(double *U_, const int L[4], int boundaryT)
{
    if (boundaryT != -1) return;
    cplx *U = (cplx *)U_;
    long V = volume(L), V3 = (long)L[0] * L[1] * L[2];
    for (long iv = V3 * (L[3] - 1); iv < V; iv++)
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) U[GL(iv, 3, a, b)] = -U[GL(iv, 3, a, b)];
}
