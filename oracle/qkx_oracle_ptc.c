/*
 * qkx_oracle_ptc.c -- CPU restatement of the gauge-equivariant
 * parallel-transport convolution (PTC) network used as a learned
 * preconditioner, with its adjoint, weight gradients and a pair loss.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load this file's
 * library.  FP64, one thread, reference AoS layout.
 *
 * The reference has no such component; the algorithm is the published PTC
 * layer of C. Lehner and T. Wettig, "Gauge-equivariant neural networks as
 * preconditioners in lattice QCD", Phys. Rev. D 108, 034503 (2023),
 * arXiv:2302.05419, eq. (5) with paths of length 0 and 1:
 *
 *   out_a(x) = sum_p sum_b W[p][a][b] (T_p in_b)(x)
 *   T_0 = 1,  T_{1+2mu} in(x) = U_mu(x) in(x+mu),
 *             T_{2+2mu} in(x) = U_mu(x-mu)^dag in(x-mu)
 *
 * Fields hold C colour-vector channels per site: f[iv][c][colour] complex.
 * A spinor Complex[V][4][3] is a 4-channel field (channel = spin).  Links are
 * the reference's U[iv][mu][row][col]; the antiperiodic sign is already in
 * the links (orc_apply_bc).  Weights W[p][a][b], p = 0..8, complex.
 */
#include <complex.h>
#include <stdlib.h>
#include <string.h>
#include "qkx_oracle.h"

typedef double complex cplx;

#define GL(iv, mu, a, b) ((((((long)(iv)) * 4 + (mu)) * 3 + (a)) * 3 + (b)))

static long vol(const int L[4]) { return (long)L[0] * L[1] * L[2] * L[3]; }

static long nbr(long iv, const int L[4], int mu, int dir)
{
    int x[4];
    long r = iv;
    for (int d = 0; d < 4; d++) {
        x[d] = (int)(r % L[d]);
        r /= L[d];
    }
    x[mu] = (x[mu] + dir + L[mu]) % L[mu];
    return ((long)(x[3] * L[2] + x[2]) * L[1] + x[1]) * L[0] + x[0];
}

/* (T_p f_c)(x) for one channel c of a C-channel field */
static void transport(cplx t[3], const cplx *f, int C, int c, const cplx *U, const int L[4], long iv, int p)
{
    if (p == 0) {
        for (int k = 0; k < 3; k++) t[k] = f[(iv * C + c) * 3 + k];
        return;
    }
    const int mu = (p - 1) / 2;
    if ((p - 1) % 2 == 0) {   /* forward: U_mu(x) f(x+mu) */
        const long n = nbr(iv, L, mu, +1);
        for (int a = 0; a < 3; a++) {
            cplx s = 0;
            for (int b = 0; b < 3; b++) s += U[GL(iv, mu, a, b)] * f[(n * C + c) * 3 + b];
            t[a] = s;
        }
    } else {                  /* backward: U_mu(x-mu)^dag f(x-mu) */
        const long n = nbr(iv, L, mu, -1);
        for (int a = 0; a < 3; a++) {
            cplx s = 0;
            for (int b = 0; b < 3; b++) s += conj(U[GL(n, mu, b, a)]) * f[(n * C + c) * 3 + b];
            t[a] = s;
        }
    }
}

static int reverse_path(int p) { return p == 0 ? 0 : ((p - 1) % 2 == 0 ? p + 1 : p - 1); }

void orc_ptc(double *out_, const double *in_, const double *U_, const int L[4], int cin, int cout,
             const double *W_)
{
    const cplx *in = (const cplx *)in_, *U = (const cplx *)U_, *W = (const cplx *)W_;
    cplx *out = (cplx *)out_;
    const long V = vol(L);
    for (long iv = 0; iv < V; iv++) {
        for (int a = 0; a < cout; a++)
            for (int k = 0; k < 3; k++) out[(iv * cout + a) * 3 + k] = 0;
        for (int p = 0; p < 9; p++)
            for (int b = 0; b < cin; b++) {
                cplx t[3];
                transport(t, in, cin, b, U, L, iv, p);
                for (int a = 0; a < cout; a++) {
                    const cplx w = W[((long)p * cout + a) * cin + b];
                    for (int k = 0; k < 3; k++) out[(iv * cout + a) * 3 + k] += w * t[k];
                }
            }
    }
}

/* Adjoint of orc_ptc (W of shape [9][cout][cin]): out_b = sum_p sum_a
 * conj(W[p][a][b]) (T_p^dag in_a), with T_p^dag = T_{reverse(p)}. */
void orc_ptc_adj(double *out_, const double *in_, const double *U_, const int L[4], int cin, int cout,
                 const double *W_)
{
    const cplx *in = (const cplx *)in_, *U = (const cplx *)U_, *W = (const cplx *)W_;
    cplx *out = (cplx *)out_;
    const long V = vol(L);
    for (long iv = 0; iv < V; iv++) {
        for (int b = 0; b < cin; b++)
            for (int k = 0; k < 3; k++) out[(iv * cin + b) * 3 + k] = 0;
        for (int p = 0; p < 9; p++)
            for (int a = 0; a < cout; a++) {
                cplx t[3];
                transport(t, in, cout, a, U, L, iv, reverse_path(p));
                for (int b = 0; b < cin; b++) {
                    const cplx w = conj(W[((long)p * cout + a) * cin + b]);
                    for (int k = 0; k < 3; k++) out[(iv * cin + b) * 3 + k] += w * t[k];
                }
            }
    }
}

/* G[p][a][b] += sum_x sum_k g_a(x)_k conj((T_p in_b)(x)_k): the gradient
 * dL/dRe W + i dL/dIm W of a real loss L whose gradient with respect to the
 * layer output is g (same convention). */
void orc_ptc_wgrad(double *G_, const double *g_, const double *in_, const double *U_, const int L[4], int cin,
                   int cout)
{
    const cplx *g = (const cplx *)g_, *in = (const cplx *)in_, *U = (const cplx *)U_;
    cplx *G = (cplx *)G_;
    const long V = vol(L);
    for (int p = 0; p < 9; p++)
        for (int b = 0; b < cin; b++)
            for (long iv = 0; iv < V; iv++) {
                cplx t[3];
                transport(t, in, cin, b, U, L, iv, p);
                for (int a = 0; a < cout; a++) {
                    cplx s = 0;
                    for (int k = 0; k < 3; k++) s += g[(iv * cout + a) * 3 + k] * conj(t[k]);
                    G[((long)p * cout + a) * cin + b] += s;
                }
            }
}

void orc_net_apply(double *out, const double *in, const double *U, const int L[4], int ch, const double *W1,
                   const double *W2)
{
    double *h = malloc(sizeof(double) * 2 * 3 * (size_t)ch * vol(L));
    orc_ptc(h, in, U, L, 4, ch, W1);
    orc_ptc(out, h, U, L, ch, 4, W2);
    free(h);
}

/* loss = |M r - e|^2 for M = L2 L1; G1, G2 (zeroed here) receive its
 * gradient with respect to W1, W2 in the dL/dRe + i dL/dIm convention. */
double orc_net_loss_grad(const double *r, const double *e, const double *U, const int L[4], int ch,
                         const double *W1, const double *W2, double *G1, double *G2)
{
    const long V = vol(L);
    const size_t nh = 2 * 3 * (size_t)ch * V, n4 = 2 * 12 * (size_t)V;
    double *h = malloc(sizeof(double) * nh), *o = malloc(sizeof(double) * n4);
    double *gh = malloc(sizeof(double) * nh), *go = malloc(sizeof(double) * n4);
    orc_ptc(h, r, U, L, 4, ch, W1);
    orc_ptc(o, h, U, L, ch, 4, W2);
    double loss = 0;
    for (size_t i = 0; i < n4; i++) {
        const double d = o[i] - e[i];
        loss += d * d;
        go[i] = 2 * d;
    }
    memset(G1, 0, sizeof(double) * 2 * 9 * 4 * ch);
    memset(G2, 0, sizeof(double) * 2 * 9 * 4 * ch);
    orc_ptc_wgrad(G2, go, h, U, L, ch, 4);
    orc_ptc_adj(gh, go, U, L, ch, 4, W2);
    orc_ptc_wgrad(G1, gh, r, U, L, 4, ch);
    free(h);
    free(o);
    free(gh);
    free(go);
    return loss;
}
