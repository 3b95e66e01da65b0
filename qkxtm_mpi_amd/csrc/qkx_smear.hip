// qkx_smear.hip -- source preparation on the GPU, FP64 like the reference:
// the SU(3) projection, APE smearing of the spatial links, Gaussian smearing
// of spinors (batches of right-hand sides, optionally one time slice) and the
// plaquette.
//
// Reference: qcd_projectSU33d lib/blas.cpp:6-422, Gauge::APE_smearing
// lib/lattice_util.cpp:950-1078, Vector::Gaussian_smearing :1080-1146,
// Gauge::calculatePlaquette :923-946.
//
// Kernels
//   ape_step_kernel    one thread per spatial link (site, mu < 3): the four
//                      staples, alpha-weighted sum with the link, projection,
//                      store.  Reads 13 links per thread, 3 unique per site
//                      from HBM (neighbours come from L2); FP64 ALU-heavy.
//   gauss_step_kernel  one thread per (site, right-hand side): the 3-D
//                      covariant hop on all four spins; the links of a site
//                      are shared by its right-hand sides (adjacent lanes).
//                      HBM-bound: 432 B of links + 384 B per right-hand side.
//   plaq_kernel        grid-stride per-site sum of the six plaquettes into
//                      per-workgroup partials, summed in a fixed order.
#include <cmath>
#include <algorithm>
#include <memory>
#include "qkx_internal.h"
#include "qkx_core.h"

namespace {

using C = Cx<double>;
struct Mat {
    C a[9];   // row-major
};

constexpr int kThreads = 256;
constexpr int kPlaqBlocks = 1024;

__device__ __forceinline__ Mat ldm(const double2 *base, int V, int mu, int s)
{
    Mat m;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        double2 d = ld(base + (size_t)(mu * 9 + k) * V, (uint32_t)s);
        m.a[k] = {d.x, d.y};
    }
    return m;
}
__device__ __forceinline__ void stm(double2 *base, int V, int mu, int s, const Mat &m)
{
#pragma unroll
    for (int k = 0; k < 9; k++) st(base + (size_t)(mu * 9 + k) * V, (uint32_t)s, make_double2(m.a[k].re, m.a[k].im));
}
// a b
__device__ __forceinline__ Mat mm(const Mat &a, const Mat &b)
{
    Mat r;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            r.a[3 * i + j] = cmac(cmac(cmul_f(a.a[3 * i], b.a[j]), a.a[3 * i + 1], b.a[3 + j]), a.a[3 * i + 2], b.a[6 + j]);
    return r;
}
// a b^dag
__device__ __forceinline__ Mat mmd(const Mat &a, const Mat &b)
{
    Mat r;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            r.a[3 * i + j] = cjmac(cjmac(cjmul_f(b.a[3 * j], a.a[3 * i]), b.a[3 * j + 1], a.a[3 * i + 1]),
                                   b.a[3 * j + 2], a.a[3 * i + 2]);
    return r;
}
// a^dag b
__device__ __forceinline__ Mat mdm(const Mat &a, const Mat &b)
{
    Mat r;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            r.a[3 * i + j] = cjmac(cjmac(cjmul_f(a.a[i], b.a[j]), a.a[3 + i], b.a[3 + j]), a.a[6 + i], b.a[6 + j]);
    return r;
}

__device__ __forceinline__ C conjc(C a) { return {a.re, -a.im}; }
__device__ __forceinline__ double nrm2c(C a) { return a.re * a.re + a.im * a.im; }

// Constants of the closed-form eigenvalues, formed on the host as the
// reference forms them (pow(18, 1/3) ...).
struct ProjConst {
    double r18, r12, r23, s3;
};

// SU(3) projection of one matrix: U = M (M^dag M)^{-1/2} exp(-i arg(det M)/3)
// by the reference's closed-form eigensystem of H = M^dag M (Cardano roots of
// the traceless part, eigenvectors from two rows of H - e), then Gram-Schmidt
// of columns 1-2 and column 3 = conj(col1 x col2).  If the off-diagonal moduli
// of H sum to at most 1e-8, row i of M is scaled by 1/sqrt(H_ii) instead.
// The steps follow lib/blas.cpp:32-420 so that results agree to rounding.
__device__ void su3_project(Mat &M, const ProjConst &pc)
{
    const C *m = M.a;
    C det = cmul(m[0], cmul(m[4], m[8])) + cmul(m[1], cmul(m[5], m[6])) + cmul(m[2], cmul(m[3], m[7]));
    det = det - (cmul(m[2], cmul(m[4], m[6])) + cmul(m[0], cmul(m[5], m[7])) + cmul(m[1], cmul(m[3], m[8])));
    const double phase = atan2(det.im, det.re) / 3.0;

    C H[9];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) H[3 * i + j] = cjmul(m[i], m[j]) + cjmul(m[3 + i], m[3 + j]) + cjmul(m[6 + i], m[6 + j]);
    // exact Hermiticity of the off-diagonal pairs
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = i + 1; j < 3; j++) {
            C h = {(H[3 * i + j].re + H[3 * j + i].re) * 0.5, (H[3 * i + j].im - H[3 * j + i].im) * 0.5};
            H[3 * i + j] = h;
            H[3 * j + i] = conjc(h);
        }
    const double n01 = nrm2c(H[1]), n02 = nrm2c(H[2]), n12 = nrm2c(H[5]);
    if (sqrt(n01) + sqrt(n02) + sqrt(n12) <= 1e-08) {
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const double e = 1.0 / sqrt(H[4 * i].re);
#pragma unroll
            for (int j = 0; j < 3; j++) M.a[3 * i + j] = e * M.a[3 * i + j];
        }
        return;
    }
    // eigenvalues: e^3 + a e + b = 0 for the traceless part
    const double tr = (H[0].re + H[4].re + H[8].re) / 3.0;
    const double h00 = H[0].re - tr, h11 = H[4].re - tr, h22 = H[8].re - tr;
    const double a = -(h22 * h22 - h00 * h11 + n01 + n02 + n12);
    const C t1 = cmul(H[1], cmul(H[5], conjc(H[2])));
    const C t2 = cmul(H[2], cmul(conjc(H[1]), conjc(H[5])));
    C b;
    b.re = (-h00 * h11 * h22 + h22 * n01 - t1.re + h11 * n02) + (h00 * n12 - t2.re);
    b.im = -t1.im - t2.im;
    const C bb = cmul(b, b);
    const C disc = {12.0 * a * a * a + 81.0 * bb.re, 81.0 * bb.im};
    // principal square root and cube root through the polar form
    double rr = sqrt(disc.re * disc.re + disc.im * disc.im), th = atan2(disc.im, disc.re);
    double sn, cs;
    sincos(0.5 * th, &sn, &cs);
    const C w = {sqrt(rr) * cs, sqrt(rr) * sn};
    const C z = {-9.0 * b.re + w.re, -9.0 * b.im + w.im};
    rr = sqrt(z.re * z.re + z.im * z.im);
    th = atan2(z.im, z.re);
    sincos(th / 3.0, &sn, &cs);
    const double rc = cbrt(rr);
    const C D = {rc * cs, rc * sn};
    const double dd = D.re * D.re + D.im * D.im;
    double e[3];
    e[0] = D.re / pc.r18 - (a * pc.r23 * D.re) / dd;
    {
        const C Ds = {D.re * pc.r12, D.im * pc.r12};
        e[1] = a * ((Ds.re + pc.s3 * Ds.im) / (Ds.re * Ds.re + Ds.im * Ds.im)) - (D.re + pc.s3 * D.im) / (pc.r18 * 2.0);
    }
    e[2] = -e[0] - e[1];
#pragma unroll
    for (int k = 0; k < 3; k++) e[k] += tr;

    // eigenvectors (stored conjugated, as the reference does) for e0 and e1
    C v[9];
    const double H00 = H[0].re, H11 = H[4].re;
    const C p10_21 = cmul(H[3], H[7]), p20_01 = cmul(H[6], H[1]);
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const double ek = e[k];
        v[3 * k + 0] = {-(ek * H[6].re - H[6].re * H11 + p10_21.re), -(ek * H[6].im - H[6].im * H11 + p10_21.im)};
        v[3 * k + 1] = {-(p20_01.re + ek * H[7].re - H00 * H[7].re), -(p20_01.im + ek * H[7].im - H00 * H[7].im)};
        v[3 * k + 2] = {-ek * ek + ek * H00 + n01 + ek * H11 - H00 * H11, 0.0};
    }
    // Gram-Schmidt of v1 against v0, normalise both; v2 = conj(v0 x v1)
    double nr = nrm2c(v[0]) + nrm2c(v[1]) + nrm2c(v[2]);
    {
        C ov = cmul(v[0], conjc(v[3])) + cmul(v[1], conjc(v[4])) + cmul(v[2], conjc(v[5]));
        ov = (1.0 / nr) * ov;
#pragma unroll
        for (int j = 0; j < 3; j++) v[3 + j] = v[3 + j] - cmul(ov, v[j]);
    }
    nr = 1.0 / sqrt(nr);
#pragma unroll
    for (int j = 0; j < 3; j++) v[j] = nr * v[j];
    nr = 1.0 / sqrt(nrm2c(v[3]) + nrm2c(v[4]) + nrm2c(v[5]));
#pragma unroll
    for (int j = 0; j < 3; j++) v[3 + j] = nr * v[3 + j];
    v[6] = conjc(cmul(v[1], v[5]) - cmul(v[2], v[4]));
    v[7] = conjc(cmul(v[2], v[3]) - cmul(v[0], v[5]));
    v[8] = conjc(cmul(v[0], v[4]) - cmul(v[1], v[3]));

    // U = (M conj(v)^T) diag(exp(-i phase)/sqrt(e_k)) v
    double ps, pcs;
    sincos(phase, &ps, &pcs);
    C vr[9];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const double de = 1.0 / sqrt(e[k]);
        const C s = {de * pcs, -de * ps};
#pragma unroll
        for (int j = 0; j < 3; j++) vr[3 * k + j] = cmul(s, v[3 * k + j]);
    }
    C U[9];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        C hm[3];
#pragma unroll
        for (int k = 0; k < 3; k++)
            hm[k] = cmul(m[3 * i], conjc(v[3 * k])) + cmul(m[3 * i + 1], conjc(v[3 * k + 1])) +
                    cmul(m[3 * i + 2], conjc(v[3 * k + 2]));
#pragma unroll
        for (int l = 0; l < 3; l++) U[3 * i + l] = cmul(hm[0], vr[l]) + cmul(hm[1], vr[3 + l]) + cmul(hm[2], vr[6 + l]);
    }
    // columns: the reference subtracts (sum_i U_i0 conj U_i1)/|col0|^2 col0
    // from col1, normalises both, and sets col3 = conj(col1 x col2)
    nr = nrm2c(U[0]) + nrm2c(U[3]) + nrm2c(U[6]);
    {
        C ov = cmul(U[0], conjc(U[1])) + cmul(U[3], conjc(U[4])) + cmul(U[6], conjc(U[7]));
        ov = (1.0 / nr) * ov;
#pragma unroll
        for (int i = 0; i < 3; i++) U[3 * i + 1] = U[3 * i + 1] - cmul(ov, U[3 * i]);
    }
    nr = 1.0 / sqrt(nr);
#pragma unroll
    for (int i = 0; i < 3; i++) U[3 * i] = nr * U[3 * i];
    nr = 1.0 / sqrt(nrm2c(U[1]) + nrm2c(U[4]) + nrm2c(U[7]));
#pragma unroll
    for (int i = 0; i < 3; i++) U[3 * i + 1] = nr * U[3 * i + 1];
    U[2] = conjc(cmul(U[3], U[7]) - cmul(U[6], U[4]));
    U[5] = conjc(cmul(U[6], U[1]) - cmul(U[0], U[7]));
    U[8] = conjc(cmul(U[0], U[4]) - cmul(U[3], U[1]));
#pragma unroll
    for (int k = 0; k < 9; k++) M.a[k] = U[k];
}

// ---- site coordinates in the parity-blocked layout ------------------------
struct Site {
    int c[4];
};
__device__ __forceinline__ Site coords(const Geom &g, int p, int cb)
{
    Site r;
    const int rest = fdiv(cb, g.dLh0);
    const int xh = cb - rest * g.Lh0;
    const int zt = fdiv(rest, g.dL1);
    r.c[1] = rest - zt * g.L[1];
    r.c[3] = fdiv(zt, g.dL2);
    r.c[2] = zt - r.c[3] * g.L[2];
    r.c[0] = 2 * xh + ((p + r.c[1] + r.c[2] + r.c[3]) & 1);
    return r;
}
__device__ __forceinline__ int blocked(const Geom &g, const Site &x)
{
    const int lex = ((x.c[3] * g.L[2] + x.c[2]) * g.L[1] + x.c[1]) * g.L[0] + x.c[0];
    return ((x.c[0] + x.c[1] + x.c[2] + x.c[3]) & 1) * g.Vh + (lex >> 1);
}
// x + d mu (d = +-1) with periodic wrap; written over constant indices so a
// run-time mu selects instead of indexing a private array in memory
__device__ __forceinline__ Site step(const Geom &g, Site x, int mu, int d)
{
#pragma unroll
    for (int m = 0; m < 4; m++)
        if (m == mu) {
            const int v = x.c[m] + d;
            x.c[m] = v < 0 ? v + g.L[m] : (v >= g.L[m] ? v - g.L[m] : v);
        }
    return x;
}

// ---- traversal ---------------------------------------------------------------
// A launch covers the sites of both parities (time slices t_lo .. t_lo+nt-1)
// with `ups` threads per site (right-hand sides, or link directions; adjacent
// lanes).  Sites of a parity go x fastest, then yb rows of y, then z, then
// the next block of rows, then t: consecutive z of a block of rows are close
// in this order, so the z-neighbours of the sites in flight are still in L2
// when they are used (whole z-planes of a 12-source FP64 field are not).
// Workgroups are remapped so that XCD k (hardware workgroup b runs on XCD
// b % 8) takes the k-th contiguous slab of both parities and sees the
// neighbour reuse of that slab in its own L2.
struct Trav {
    FastDiv dups, dyb, dnyb;
    int ups, yb, nyb, t_lo;
    int nunits;     // units of one parity
    int nblk8;      // workgroups per XCD and parity when remapped, else 0
};

Trav make_trav(const int L[4], int ups, int t, dim3 &grid)
{
    Trav tr;
    tr.ups = ups;
    tr.dups = make_fastdiv(ups);
    tr.yb = 1;   // rows per block: the largest divisor of L[1] up to 8
    for (int d = 2; d <= 8; d++)
        if (L[1] % d == 0) tr.yb = d;
    tr.nyb = L[1] / tr.yb;
    tr.dyb = make_fastdiv(tr.yb);
    tr.dnyb = make_fastdiv(tr.nyb);
    tr.t_lo = t < 0 ? 0 : t;
    tr.nunits = L[0] * L[1] * L[2] / 2 * (t < 0 ? L[3] : 1) * ups;
    int nbh = (tr.nunits + kThreads - 1) / kThreads;
    tr.nblk8 = 0;
    grid = dim3(nbh, 2);
    if (nbh >= 32) {
        nbh = (nbh + 7) / 8 * 8;
        tr.nblk8 = nbh / 8;
        grid = dim3(2 * nbh, 1);
    }
    return tr;
}

// This thread's unit: parity p, site x, its blocked index s and sub-index
// (0 .. ups-1).  False past the end.
__device__ __forceinline__ bool trav_unit(const Geom &g, const Trav &tr, int &p, Site &x, int &s, int &sub)
{
    int lb;
    if (tr.nblk8 > 0) {
        const int b = blockIdx.x, i = b >> 3;
        p = i & 1;
        lb = (b & 7) * tr.nblk8 + (i >> 1);
    } else {
        p = blockIdx.y;
        lb = blockIdx.x;
    }
    const int u = lb * blockDim.x + threadIdx.x;
    if (u >= tr.nunits) return false;
    const int v = fdiv(u, tr.dups);
    sub = u - v * tr.ups;
    const int q1 = fdiv(v, g.dLh0), xh = v - q1 * g.Lh0;
    const int q2 = fdiv(q1, tr.dyb), yi = q1 - q2 * tr.yb;
    const int q3 = fdiv(q2, g.dL2), z = q2 - q3 * g.L[2];
    const int tt = fdiv(q3, tr.dnyb), ybk = q3 - tt * tr.nyb;
    x.c[1] = ybk * tr.yb + yi;
    x.c[2] = z;
    x.c[3] = tr.t_lo + tt;
    x.c[0] = 2 * xh + ((p + x.c[1] + z + x.c[3]) & 1);
    s = p * g.Vh + xh + g.Lh0 * (x.c[1] + g.L[1] * (z + g.L[2] * x.c[3]));
    return true;
}

// ---- APE ---------------------------------------------------------------------
struct ApeArgs {
    const double2 *in;
    double2 *out;
    Geom g;
    Trav tr;        // 3 threads per site: mu = 0, 1, 2
    double alpha;
    ProjConst pc;
};

// out_mu(x) = P[ U_mu(x) + alpha (forward staples over nu != mu in ascending
// order, then the backward ones) ]; the three links of a site are adjacent
// lanes, which share most of their staples' links
__global__ __launch_bounds__(kThreads) void ape_step_kernel(ApeArgs a)
{
    const Geom &g = a.g;
    int p, s, mu;
    Site x;
    if (!trav_unit(g, a.tr, p, x, s, mu)) return;
    const int V = g.V;
    const int xpm = blocked(g, step(g, x, mu, +1));
    const int nus[2] = {mu == 0 ? 1 : 0, mu == 2 ? 1 : 2};
    Mat S;
#pragma unroll
    for (int k = 0; k < 9; k++) S.a[k] = {0.0, 0.0};
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const int nu = nus[j];
        const Mat T = mm(ldm(a.in, V, nu, s), ldm(a.in, V, mu, blocked(g, step(g, x, nu, +1))));
        const Mat F = mmd(T, ldm(a.in, V, nu, xpm));
#pragma unroll
        for (int k = 0; k < 9; k++) S.a[k] = S.a[k] + F.a[k];
    }
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const int nu = nus[j];
        const Site xm = step(g, x, nu, -1);
        const int sxm = blocked(g, xm);
        const Mat T = mm(ldm(a.in, V, mu, sxm), ldm(a.in, V, nu, blocked(g, step(g, xm, mu, +1))));
        const Mat B = mdm(ldm(a.in, V, nu, sxm), T);
#pragma unroll
        for (int k = 0; k < 9; k++) S.a[k] = S.a[k] + B.a[k];
    }
    const Mat U = ldm(a.in, V, mu, s);
#pragma unroll
    for (int k = 0; k < 9; k++) S.a[k] = {a.alpha * S.a[k].re + U.a[k].re, a.alpha * S.a[k].im + U.a[k].im};
    su3_project(S, a.pc);
    stm(a.out, V, mu, s, S);
}

__global__ __launch_bounds__(kThreads) void project_kernel(double2 *U, int V, ProjConst pc)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= V) return;
    const int mu = blockIdx.y;
    Mat M = ldm(U, V, mu, s);
    su3_project(M, pc);
    stm(U, V, mu, s, M);
}

// ---- plaquette -----------------------------------------------------------------
// Re tr[U_mu(x) U_nu(x+mu) (U_nu(x) U_mu(x+nu))^dag] over the six planes.
// The time-boundary sign of a gauge field cancels in every plaquette (both
// temporal links of a (mu, t) plaquette sit on the same time slice), so the
// stored links are used as they are.
__global__ __launch_bounds__(kThreads) void plaq_kernel(const double2 *U, Geom g, double *partial)
{
    double acc = 0.0;
    const int V = g.V;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < V; s += gridDim.x * blockDim.x) {
        const int p = s >= g.Vh ? 1 : 0;
        const Site x = coords(g, p, s - p * g.Vh);
        int sp[4];
#pragma unroll
        for (int m = 0; m < 4; m++) sp[m] = blocked(g, step(g, x, m, +1));
#pragma unroll
        for (int mu = 0; mu < 3; mu++)
#pragma unroll
            for (int nu = mu + 1; nu < 4; nu++) {
                const Mat P1 = mm(ldm(U, V, mu, s), ldm(U, V, nu, sp[mu]));
                const Mat P2 = mm(ldm(U, V, nu, s), ldm(U, V, mu, sp[nu]));
#pragma unroll
                for (int k = 0; k < 9; k++) acc += P1.a[k].re * P2.a[k].re + P1.a[k].im * P2.a[k].im;
            }
    }
    __shared__ double lds[kThreads / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int i = 0; i < (int)(blockDim.x >> 6); i++) t += lds[i];
        partial[blockIdx.x] = t;
    }
}

// ---- Gaussian smearing -------------------------------------------------------
struct GaussArgs {
    SpinorD in, out;
    const double2 *U;
    Geom g;
    Trav tr;        // nr threads per site
    int nr;
    double alpha, norm;
};

// tot += U psi (forward hop) or U^dag psi (backward hop) on the four spins;
// the link elements come from get(k)
template <bool BWD, class GetU>
__device__ __forceinline__ void gauss_hop(C tot[12], GetU get, const C ps[12])
{
#pragma unroll
    for (int i = 0; i < 3; i++) {
        C u0, u1, u2;   // row i of U, or column i of U for U^dag
        if (BWD) { u0 = get(i); u1 = get(3 + i); u2 = get(6 + i); }
        else { u0 = get(3 * i); u1 = get(3 * i + 1); u2 = get(3 * i + 2); }
#pragma unroll
        for (int sp = 0; sp < 4; sp++) {
            const C h = BWD ? cjmac(cjmac(cjmul_f(u0, ps[3 * sp]), u1, ps[3 * sp + 1]), u2, ps[3 * sp + 2])
                            : cmac(cmac(cmul_f(u0, ps[3 * sp]), u1, ps[3 * sp + 1]), u2, ps[3 * sp + 2]);
            tot[3 * sp + i] = tot[3 * sp + i] + h;
        }
    }
}

// The output is not read again in the launch: streamed stores keep its
// 2.3 KB per site (12 FP64 right-hand sides) out of the L2 that holds the
// neighbour spinors.
__device__ __forceinline__ void gauss_finish(const GaussArgs &a, int e, const C tot[12])
{
    typedef __attribute__((ext_vector_type(4))) unsigned int U4;
    C v[12];
    a.in.load(e, v);
#pragma unroll
    for (int k = 0; k < 12; k++) {
        const double2 o = make_double2(a.norm * (v[k].re + a.alpha * tot[k].re), a.norm * (v[k].im + a.alpha * tot[k].im));
        U4 w;
        __builtin_memcpy(&w, &o, sizeof(o));
        __builtin_nontemporal_store(w, reinterpret_cast<U4 *>(a.out.p + (size_t)k * a.out.stride) + (uint32_t)e);
    }
}

// out(x) = norm (in(x) + alpha sum_{mu<3} [U_mu(x) in(x+mu) + U_mu(x-mu)^dag in(x-mu)])
// One thread per (site, right-hand side); the right-hand sides of a site are
// adjacent lanes and read the same links (one memory transaction).  Staging
// the links in LDS, or ordering the six hops with register fences, measured
// no faster.
__global__ __launch_bounds__(kThreads) void gauss_step_kernel(GaussArgs a)
{
    const Geom &g = a.g;
    int p, s, r;
    Site x;
    if (!trav_unit(g, a.tr, p, x, s, r)) return;
    const int V = g.V, nr = a.nr;
    C tot[12];
#pragma unroll
    for (int k = 0; k < 12; k++) tot[k] = {0.0, 0.0};
#pragma unroll
    for (int h = 0; h < 6; h++) {
        const int mu = h >> 1;
        const bool bwd = h & 1;
        const int sn = blocked(g, step(g, x, mu, bwd ? -1 : +1));
        const Mat U = ldm(a.U, V, mu, bwd ? sn : s);
        C ps[12];
        a.in.load(sn * nr + r, ps);
        auto get = [&](int k) { return U.a[k]; };
        if (bwd) gauss_hop<true>(tot, get, ps);
        else gauss_hop<false>(tot, get, ps);
    }
    gauss_finish(a, s * nr + r, tot);
}

ProjConst proj_const()
{
    return ProjConst{std::pow(18., 1. / 3.), std::pow(12., 1. / 3.), std::pow(2. / 3., 1. / 3.), std::sqrt(3.)};
}

bool fp64_links(const qkx_gauge_s *g) { return g && g->prec == QKX_PREC_DOUBLE && g->recon == QKX_RECON_18; }

// Spatial-link workspace of the APE ping-pong, kept between calls (a smearing
// run calls this once per configuration; re-allocating 432 B per site each
// time would cost more than a step).  Grows on demand; work queued on the
// library stream against the old buffer finishes before it is freed.
double2 *g_ape_ws = nullptr;
size_t g_ape_ws_bytes = 0;

double2 *ape_workspace(size_t bytes)
{
    if (bytes <= g_ape_ws_bytes) return g_ape_ws;
    if (g_ape_ws) {
        (void)hipStreamSynchronize(qkx::g_stream);
        (void)hipFree(g_ape_ws);
    }
    g_ape_ws = nullptr;
    g_ape_ws_bytes = 0;
    if (hipMalloc(&g_ape_ws, bytes) != hipSuccess) {
        g_ape_ws = nullptr;
        return nullptr;
    }
    g_ape_ws_bytes = bytes;
    return g_ape_ws;
}

} // namespace

namespace qkx {
void smear_release()
{
    if (g_ape_ws) (void)hipFree(g_ape_ws);
    g_ape_ws = nullptr;
    g_ape_ws_bytes = 0;
}
} // namespace qkx

extern "C" {

int qkx_gauge_project_su3(qkx_gauge *g)
{
    QKX_CHECK_ARG(fp64_links(g), "qkx_gauge_project_su3: the gauge field must be FP64 recon 18");
    hipLaunchKernelGGL(project_kernel, dim3((g->V + kThreads - 1) / kThreads, 3), dim3(kThreads), 0, qkx::g_stream,
                       (double2 *)g->data, g->V, proj_const());
    QKX_HIP(hipGetLastError());
    qkx::gauge_modified(g);
    return QKX_SUCCESS;
}

int qkx_ape_smear(qkx_gauge *out, const qkx_gauge *in, int nsmear, double alpha)
{
    QKX_CHECK_ARG(fp64_links(out) && fp64_links(in), "qkx_ape_smear: both gauge fields must be FP64 recon 18");
    QKX_CHECK_ARG(out != in, "qkx_ape_smear: out and in must be different fields");
    QKX_CHECK_ARG(out->V == in->V && out->L[0] == in->L[0] && out->L[1] == in->L[1] && out->L[2] == in->L[2],
                  "qkx_ape_smear: lattices differ");
    QKX_CHECK_ARG(out->bc == in->bc, "qkx_ape_smear: boundaryT of out and in differ");
    QKX_CHECK_ARG(nsmear >= 0, "qkx_ape_smear: nsmear must be >= 0");
    if (nsmear == 0) return QKX_SUCCESS;   // the reference returns before touching the field
    const int V = in->V;
    double2 *tmp = nullptr;
    if (nsmear > 1 && !(tmp = ape_workspace((size_t)3 * 9 * V * sizeof(double2)))) {
        qkx::set_error("qkx_ape_smear: out of device memory");
        return QKX_ERR_NOMEM;
    }
    ApeArgs a;
    a.g = qkx::make_geom(in->L, in->bc);
    dim3 grid;
    a.tr = make_trav(in->L, 3, -1, grid);
    a.alpha = alpha;
    a.pc = proj_const();
    // ping-pong so that the last step lands in out
    double2 *bufs[2] = {(double2 *)out->data, tmp};
    const double2 *src = (const double2 *)in->data;
    for (int it = 0; it < nsmear; it++) {
        double2 *dst = bufs[(nsmear - 1 - it) & 1];
        a.in = src;
        a.out = dst;
        hipLaunchKernelGGL(ape_step_kernel, grid, dim3(kThreads), 0, qkx::g_stream, a);
        src = dst;
    }
    QKX_HIP(hipGetLastError());
    // temporal links: in's
    QKX_HIP(hipMemcpyAsync((double2 *)out->data + (size_t)3 * 9 * V, (const double2 *)in->data + (size_t)3 * 9 * V,
                           (size_t)9 * V * sizeof(double2), hipMemcpyDeviceToDevice, qkx::g_stream));
    qkx::gauge_modified(out);
    return QKX_SUCCESS;
}

int qkx_plaquette(double *plaq, const qkx_gauge *g)
{
    QKX_CHECK_ARG(plaq && fp64_links(g), "qkx_plaquette: the gauge field must be FP64 recon 18");
    const int nb = std::min(kPlaqBlocks, (g->V + kThreads - 1) / kThreads);
    Geom geo = qkx::make_geom(g->L, g->bc);
    hipLaunchKernelGGL(plaq_kernel, dim3(nb), dim3(kThreads), 0, qkx::g_stream, (const double2 *)g->data, geo,
                       qkx::reduce_scratch());
    QKX_HIP(hipGetLastError());
    int rc = qkx::finish_sum_dev(qkx::reduce_scratch(), nb, qkx::reduce_result());
    if (rc) return rc;
    double sum = 0;
    QKX_HIP(hipMemcpyAsync(&sum, qkx::reduce_result(), sizeof(double), hipMemcpyDeviceToHost, qkx::g_stream));
    QKX_HIP(hipStreamSynchronize(qkx::g_stream));
    *plaq = sum / ((double)g->V * 3.0 * 6.0);
    return QKX_SUCCESS;
}

int qkx_gauss_smear(qkx_spinor *out, const qkx_spinor *in, const qkx_gauge *U, int nsmear, double alpha, int t)
{
    QKX_CHECK_ARG(out && in && fp64_links(U), "qkx_gauss_smear: null field or gauge field not FP64 recon 18");
    QKX_CHECK_ARG(out != in, "qkx_gauss_smear: out and in must be different fields");
    QKX_CHECK_ARG(in->prec == QKX_PREC_DOUBLE && out->prec == QKX_PREC_DOUBLE && in->subset == QKX_SUBSET_FULL &&
                      out->subset == QKX_SUBSET_FULL,
                  "qkx_gauss_smear: spinors must be full FP64 fields");
    QKX_CHECK_ARG(in->V == U->V && out->V == in->V && out->nrhs == in->nrhs &&
                      in->L[0] == U->L[0] && in->L[1] == U->L[1] && in->L[2] == U->L[2],
                  "qkx_gauss_smear: fields differ in lattice or number of right-hand sides");
    QKX_CHECK_ARG(nsmear >= 0, "qkx_gauss_smear: nsmear must be >= 0");
    QKX_CHECK_ARG(t >= -1 && t < in->L[3], "qkx_gauss_smear: t must be -1 (all slices) or a time slice");
    if (nsmear == 0) return QKX_SUCCESS;   // the reference returns before touching the field
    const int *L = in->L;
    const int nr = in->nrhs;
    GaussArgs a;
    a.U = (const double2 *)U->data;
    a.g = qkx::make_geom(L, U->bc);
    a.nr = nr;
    dim3 grid;
    a.tr = make_trav(L, nr, t, grid);
    a.alpha = alpha;
    a.norm = 1.0 / (1 + 6 * alpha);
    int rc;
    if (t >= 0 && (rc = qkx_spinor_zero(out))) return rc;
    std::unique_ptr<qkx::Temp> tmp;
    if (nsmear > 1) {
        tmp.reset(new qkx::Temp(L, QKX_PREC_DOUBLE, QKX_SUBSET_FULL, nr));
        if (!tmp->f) { qkx::set_error("qkx_gauss_smear: out of device memory"); return QKX_ERR_NOMEM; }
        // a pooled temporary may hold anything outside the smeared slice; it
        // is never read there
    }
    qkx_spinor_s *bufs[2] = {out, tmp ? tmp->f : nullptr};
    const qkx_spinor_s *src = in;
    for (int it = 0; it < nsmear; it++) {
        qkx_spinor_s *dst = bufs[(nsmear - 1 - it) & 1];
        a.in = make_spinor<SpinorD>(src);
        a.out = make_spinor<SpinorD>(dst);
        hipLaunchKernelGGL(gauss_step_kernel, grid, dim3(kThreads), 0, qkx::g_stream, a);
        src = dst;
    }
    QKX_HIP(hipGetLastError());
    return QKX_SUCCESS;
}

} // extern "C"
