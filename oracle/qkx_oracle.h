/*
 * qkx_oracle.h -- CPU restatement of the reference twisted-mass operator,
 * its CG and BLAS (FP64, reference AoS layout).
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker: tests/, the smoke()
 * entry and bench.py's cpu_baseline leg may call it; the product library
 * (libqkx_hip.so) never links or calls it.
 *
 * Layout (reference include/qkxTM.h:117-134):
 *   site   iv  = ((t*Lz + z)*Ly + y)*Lx + x                 (LEXIC)
 *   spinor psi[iv][spin 4][colour 3]  complex, interleaved re/im
 *   gauge  U[iv][mu 4][row 3][col 3]  complex, mu: 0=x 1=y 2=z 3=t
 *
 * Parity selectors follow reference include/lattice_util.h:23
 *   enum PARITY{EO=0, OE=1, OO=2, EE=3}
 *   hopping: EO -> even outputs computed (odd zeroed), OE -> odd outputs
 *   twist-PC: OO -> odd outputs, EE -> even outputs
 */
#ifndef QKX_ORACLE_H
#define QKX_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_EO = 0, ORC_OE = 1, ORC_OO = 2, ORC_EE = 3, ORC_ALL = -1 };
enum { ORC_UKQCD = 0, ORC_CHIRAL = 1 };
enum { ORC_TWIST = 0, ORC_TWIST_DAG = 1, ORC_TWIST_INV = 2, ORC_TWIST_DAG_INV = 3 };
enum { ORC_TM = 0, ORC_TM_DAG = 1, ORC_TM_DAG_TM = 2, ORC_WILSON = 3 };

/* Hopping term H (dagger=0) or H^dagger (dagger=1).  parity: ORC_ALL, ORC_EO, ORC_OE. */
void orc_hop(double *out, const double *psi, const double *U, const int L[4],
             int basis, int dagger, int parity);

/* out = -kappa*dslash + psi + i a g5 psi   (dagger: a -> -a), a = 2 kappa mu s */
void orc_twist_add_dslash(double *out, const double *dslash, const double *psi,
                          const int L[4], double kappa, double mu, int twistSign,
                          int basis, int dagger);
/* out = -kappa^2*dslash + (psi + i a g5 psi) on parity (OO/EE), zero elsewhere */
void orc_twist_add_dslash_pc(double *out, const double *dslash, const double *psi,
                             const int L[4], double kappa, double mu, int twistSign,
                             int basis, int parity);
/* out = psi - kappa*dslash */
void orc_delta_add_dslash(double *out, const double *dslash, const double *psi,
                          const int L[4], double kappa);
/* site-local twist family: A, A^dag, A^-1, (A^dag)^-1 */
void orc_twist(double *out, const double *psi, const int L[4], double kappa,
               double mu, int twistSign, int basis, int kind);
void orc_gamma5(double *out, const double *psi, const int L[4], int basis);

/* composites */
void orc_tm(double *out, const double *psi, const double *U, const int L[4],
            double kappa, double mu, int twistSign, int basis, int op);
void orc_tm_xx(double *out, const double *psi, const double *U, const int L[4],
               double kappa, double mu, int twistSign, int basis, int parity, int normal);

/* BLAS over 12*V complex numbers */
double orc_redot(const double *x, const double *y, long V);
void orc_axpy(double *x, double a, const double *y, long V);   /* x = x + a y */
void orc_xpay(double *x, double a, const double *y, long V);   /* x = a x + y */
double orc_norm2(const double *x, long V);

/* Reference CG on D^dag D; returns iterations; r2_final receives the iterated r2 */
int orc_cg(double *x, const double *b, const double *U, const int L[4],
           double kappa, double mu, int twistSign, double tol, int maxIter,
           double *r2_final);

/* Antiperiodic (or periodic) time boundary as a link sign on the last slice */
void orc_apply_bc(double *U, const int L[4], int boundaryT);

/* ---- learned preconditioner (qkx_oracle_ptc.c): gauge-equivariant
 * parallel-transport convolutions.  Fields f[V][C][3] complex; W[9][cout][cin]
 * complex, path 0 = identity, 1+2mu = U_mu(x) f(x+mu), 2+2mu = U_mu(x-mu)^dag
 * f(x-mu).  Links carry the boundary sign. */
void orc_ptc(double *out, const double *in, const double *U, const int L[4], int cin, int cout,
             const double *W);
void orc_ptc_adj(double *out, const double *in, const double *U, const int L[4], int cin, int cout,
                 const double *W);
void orc_ptc_wgrad(double *G, const double *g, const double *in, const double *U, const int L[4], int cin,
                   int cout);
void orc_net_apply(double *out, const double *in, const double *U, const int L[4], int ch, const double *W1,
                   const double *W2);
double orc_net_loss_grad(const double *r, const double *e, const double *U, const int L[4], int ch,
                         const double *W1, const double *W2, double *G1, double *G2);

/* ---- source preparation (qkx_oracle_smear.c) ----
 * SU(3) projection of one 3x3 matrix (18 reals, row-major) in place, and of the
 * spatial links of a field; APE smearing of the spatial links (temporal links
 * copied from in); Gaussian smearing of a spinor; the plaquette.  nsmear = 0
 * leaves out untouched, as the reference does. */
void orc_su3_project(double *m18);
void orc_project_su3_field(double *U, const int L[4]);
void orc_ape_smear(double *out, const double *in, const int L[4], int nsmear, double alpha);
void orc_gauss_smear(double *out, const double *in, const double *U, const int L[4], int nsmear, double alpha);
double orc_plaquette(const double *U, const int L[4]);
/* propagator M[V][4][4][3][3] rotated to the physical basis, sign +1 / -1 */
void orc_prop_rotate(double *M, long V, int sign);

#ifdef __cplusplus
}
#endif
#endif
