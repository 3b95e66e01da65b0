// qkx_dslash.hip -- the twisted-mass hopping stencil and its fused site terms.
//
// One thread per output site.  Each thread gathers the eight neighbour
// spinors (only the two spins a projector reads), projects them to half
// spinors, multiplies by the link (or its adjoint), reconstructs, and applies
// a fused epilogue
//     out = cR R + gR i g5 R + cX x + gX i g5 x     (then optionally g5 out)
// which covers every site term of the reference (twist, twist-inverse,
// xpay, Wilson delta, even-odd PC term, gamma5).
//
// Hopping term (reference lib/lattice_util.cpp:1158-1213 and :1643-1698):
//   H psi(x)     = sum_mu (1-g_mu) U_mu(x) psi(x+mu) + (1+g_mu) U_mu(x-mu)^dag psi(x-mu)
//   H^dag psi(x) = the same with the two projectors exchanged.
// Projector/reconstruction tables: include/hopping_term.h (UKQCD),
// include/hopping_term_chiral.h (chiral).  The table entries are the
// reference's conventions and are necessarily identical to it.
#include "qkx_internal.h"
#include "qkx_core.h"

namespace {

enum : int { K0 = 0, KP1, KM1, KPI, KMI, K2 };

struct Proj {
    int sa[2], a[2], sb[2], b[2];   // h_k = a_k psi[sa_k] + b_k psi[sb_k]
    int kq[4], q[4];                // R_s += q_s xi[kq_s]
};

// [basis][mu][0: (1-g_mu), 1: (1+g_mu)]
constexpr Proj kProj[2][4][2] = {
    {   // UKQCD
        {{{0, 1}, {KP1, KP1}, {3, 2}, {KMI, KMI}, {0, 1, 1, 0}, {KP1, KP1, KPI, KPI}},
         {{0, 1}, {KP1, KP1}, {3, 2}, {KPI, KPI}, {0, 1, 1, 0}, {KP1, KP1, KMI, KMI}}},
        {{{0, 1}, {KP1, KP1}, {3, 2}, {KM1, KP1}, {0, 1, 1, 0}, {KP1, KP1, KP1, KM1}},
         {{0, 1}, {KP1, KP1}, {3, 2}, {KP1, KM1}, {0, 1, 1, 0}, {KP1, KP1, KM1, KP1}}},
        {{{0, 1}, {KP1, KP1}, {2, 3}, {KMI, KPI}, {0, 1, 0, 1}, {KP1, KP1, KPI, KMI}},
         {{0, 1}, {KP1, KP1}, {2, 3}, {KPI, KMI}, {0, 1, 0, 1}, {KP1, KP1, KMI, KPI}}},
        {{{2, 3}, {K2, K2}, {0, 0}, {K0, K0}, {0, 0, 0, 1}, {K0, K0, KP1, KP1}},
         {{0, 1}, {K2, K2}, {0, 0}, {K0, K0}, {0, 1, 0, 0}, {KP1, KP1, K0, K0}}},
    },
    {   // chiral
        {{{0, 1}, {KP1, KP1}, {3, 2}, {KPI, KPI}, {0, 1, 1, 0}, {KP1, KP1, KMI, KMI}},
         {{0, 1}, {KP1, KP1}, {3, 2}, {KMI, KMI}, {0, 1, 1, 0}, {KP1, KP1, KPI, KPI}}},
        {{{0, 1}, {KP1, KP1}, {3, 2}, {KP1, KM1}, {0, 1, 1, 0}, {KP1, KP1, KM1, KP1}},
         {{0, 1}, {KP1, KP1}, {3, 2}, {KM1, KP1}, {0, 1, 1, 0}, {KP1, KP1, KP1, KM1}}},
        {{{0, 1}, {KP1, KP1}, {2, 3}, {KPI, KMI}, {0, 1, 0, 1}, {KP1, KP1, KMI, KPI}},
         {{0, 1}, {KP1, KP1}, {2, 3}, {KMI, KPI}, {0, 1, 0, 1}, {KP1, KP1, KPI, KMI}}},
        {{{0, 1}, {KP1, KP1}, {2, 3}, {KP1, KP1}, {0, 1, 0, 1}, {KP1, KP1, KP1, KP1}},
         {{0, 1}, {KP1, KP1}, {2, 3}, {KM1, KM1}, {0, 1, 0, 1}, {KP1, KP1, KM1, KM1}}},
    }};

template <int K, class T> __device__ __forceinline__ Cx<T> umul(Cx<T> z)
{
    if constexpr (K == KP1) return z;
    else if constexpr (K == KM1) return {-z.re, -z.im};
    else if constexpr (K == KPI) return {-z.im, z.re};
    else if constexpr (K == KMI) return {z.im, -z.re};
    else if constexpr (K == K2) return {z.re + z.re, z.im + z.im};
    else return {T(0), T(0)};
}

template <int B, int MU, int KIND> constexpr bool uses_spin_pair(int pair)
{
    constexpr Proj P = kProj[B][MU][KIND];
    bool u = false;
    for (int k = 0; k < 2; k++) {
        if (P.a[k] != K0 && (P.sa[k] >> 1) == pair) u = true;
        if (P.b[k] != K0 && (P.sb[k] >> 1) == pair) u = true;
    }
    return u;
}

template <int B, int MU, int KIND, int k, class T>
__device__ __forceinline__ void project_k(const Cx<T> psi[12], Cx<T> h[6])
{
    constexpr Proj P = kProj[B][MU][KIND];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        Cx<T> v = umul<P.a[k]>(psi[P.sa[k] * 3 + c]);
        if constexpr (P.b[k] != K0) v = v + umul<P.b[k]>(psi[P.sb[k] * 3 + c]);
        h[k * 3 + c] = v;
    }
}

template <int B, int MU, int KIND, int s, class T>
__device__ __forceinline__ void collect_s(Cx<T> R[12], const Cx<T> xi[6])
{
    constexpr Proj P = kProj[B][MU][KIND];
    if constexpr (P.q[s] != K0) {
#pragma unroll
        for (int c = 0; c < 3; c++) R[s * 3 + c] = R[s * 3 + c] + umul<P.q[s]>(xi[P.kq[s] * 3 + c]);
    }
}

// xi = U h (DAGL=false) or U^dag h (DAGL=true), h holds two colour vectors.
template <bool DAGL, class T>
__device__ __forceinline__ void link_mul(const Cx<T> U[9], const Cx<T> h[6], Cx<T> xi[6])
{
#pragma unroll
    for (int k = 0; k < 2; k++)
#pragma unroll
        for (int a = 0; a < 3; a++) {
            Cx<T> acc;
            if constexpr (!DAGL) {
                acc = cmul(U[a * 3 + 0], h[k * 3 + 0]);
                acc = acc + cmul(U[a * 3 + 1], h[k * 3 + 1]);
                acc = acc + cmul(U[a * 3 + 2], h[k * 3 + 2]);
            } else {
                acc = cjmul(U[0 * 3 + a], h[k * 3 + 0]);
                acc = acc + cjmul(U[1 * 3 + a], h[k * 3 + 1]);
                acc = acc + cjmul(U[2 * 3 + a], h[k * 3 + 2]);
            }
            xi[k * 3 + a] = acc;
        }
}

// One direction, one sign: load neighbour spins, project, link, collect.
template <int B, int MU, int KIND, bool DAGL, class SIn, class G, class T>
__device__ __forceinline__ void hop_dir(Cx<T> R[12], const SIn &in, int s_nb, const G &U,
                                        int s_link, T sign)
{
    Cx<T> psi[12];
    if constexpr (uses_spin_pair<B, MU, KIND>(0)) in.template load_half<0>(s_nb, psi);
    if constexpr (uses_spin_pair<B, MU, KIND>(1)) in.template load_half<1>(s_nb, psi);
    Cx<T> h[6];
    project_k<B, MU, KIND, 0>(psi, h);
    project_k<B, MU, KIND, 1>(psi, h);
    Cx<T> Ul[9];
    U.load(MU, s_link, Ul);
    if (sign != T(1)) {
#pragma unroll
        for (int k = 0; k < 6; k++) h[k] = sign * h[k];
    }
    Cx<T> xi[6];
    link_mul<DAGL>(Ul, h, xi);
    collect_s<B, MU, KIND, 0>(R, xi);
    collect_s<B, MU, KIND, 1>(R, xi);
    collect_s<B, MU, KIND, 2>(R, xi);
    collect_s<B, MU, KIND, 3>(R, xi);
}

template <int B, class T> __device__ __forceinline__ void gamma5(const Cx<T> v[12], Cx<T> w[12])
{
#pragma unroll
    for (int c = 0; c < 3; c++) {
        if constexpr (B == QKX_BASIS_UKQCD) {
            w[0 * 3 + c] = v[2 * 3 + c];
            w[1 * 3 + c] = v[3 * 3 + c];
            w[2 * 3 + c] = v[0 * 3 + c];
            w[3 * 3 + c] = v[1 * 3 + c];
        } else {
            w[0 * 3 + c] = v[0 * 3 + c];
            w[1 * 3 + c] = v[1 * 3 + c];
            w[2 * 3 + c] = {-v[2 * 3 + c].re, -v[2 * 3 + c].im};
            w[3 * 3 + c] = {-v[3 * 3 + c].re, -v[3 * 3 + c].im};
        }
    }
}

// Epilogue coefficients.
struct Epi {
    double cR, gR, cX, gX;
    int g5out;
};

template <int B, class T>
__device__ __forceinline__ void epilogue(const Epi &e, const Cx<T> R[12], const Cx<T> *x, Cx<T> out[12])
{
    const T cR = T(e.cR), gR = T(e.gR), cX = T(e.cX), gX = T(e.gX);
    Cx<T> g5R[12], g5x[12];
    if (gR != T(0)) gamma5<B>(R, g5R);
    if (x && gX != T(0)) gamma5<B>(x, g5x);
#pragma unroll
    for (int k = 0; k < 12; k++) {
        Cx<T> v = cR * R[k];
        if (gR != T(0)) v = v + gR * timesi(g5R[k]);
        if (x) {
            v = v + cX * x[k];
            if (gX != T(0)) v = v + gX * timesi(g5x[k]);
        }
        out[k] = v;
    }
    if (e.g5out) {
        Cx<T> w[12];
        gamma5<B>(out, w);
#pragma unroll
        for (int k = 0; k < 12; k++) out[k] = w[k];
    }
}

template <class SIn, class SOut, class SAux, class G>
struct HopArgs {
    Geom g;
    SIn in;
    SOut out;
    SAux aux;
    G U;
    int in_off[2], out_off[2], aux_off[2];
    int par_out;   // -1: all sites; 0/1: that parity only
    int nout;
    int nblk8;     // grid size / 8 when XCD remapping is on, else 0
    const int *skip;   // when non-null and *skip != 0 the launch does nothing
    Epi epi;
};

// Hardware workgroup IDs are dealt round-robin to the 8 XCDs; remap them so
// that each XCD walks one contiguous range of sites and its private L2 sees
// the neighbour reuse of that range.
__device__ __forceinline__ int logical_block(int nblk8)
{
    int b = blockIdx.x;
    if (nblk8 > 0) b = (b & 7) * nblk8 + (b >> 3);
    return b;
}

template <class SIn, class SOut, class SAux, class G, int B, bool DAG, bool AUX>
__global__ __launch_bounds__(256) void hop_kernel(HopArgs<SIn, SOut, SAux, G> a)
{
    using T = typename SOut::Real;
    if (a.skip && *a.skip) return;
    const int o = logical_block(a.nblk8) * blockDim.x + threadIdx.x;
    if (o >= a.nout) return;
    const Geom &g = a.g;
    int p, cb;
    if (a.par_out < 0) {
        p = (o >= g.Vh) ? 1 : 0;
        cb = o - p * g.Vh;
    } else {
        p = a.par_out;
        cb = o;
    }
    // coordinates: cb = xh + Lh0 * rest,  rest = y + Ly*(z + Lz*t)
    const int L0 = g.L[0], L1 = g.L[1], L2 = g.L[2], L3 = g.L[3];
    const int xh = cb % g.Lh0;
    const int rest = cb / g.Lh0;
    const int y = rest % L1;
    const int zt = rest / L1;
    const int z = zt % L2;
    const int t = zt / L2;
    const int x = 2 * xh + ((p + y + z + t) & 1);
    const int q = p ^ 1;
    const int qoff = q * g.Vh, poff = p * g.Vh;
    const int in_q = a.in_off[q];

    const int xp = (x + 1 == L0) ? 0 : x + 1;
    const int xm = (x == 0) ? L0 - 1 : x - 1;
    const int cb_xp = (xp >> 1) + g.Lh0 * rest;
    const int cb_xm = (xm >> 1) + g.Lh0 * rest;
    const int ry = (y + 1 == L1) ? -(L1 - 1) : 1;
    const int rym = (y == 0) ? (L1 - 1) : -1;
    const int cb_yp = cb + g.Lh0 * ry;
    const int cb_ym = cb + g.Lh0 * rym;
    const int LY = L1 * g.Lh0;
    const int cb_zp = cb + ((z + 1 == L2) ? -(L2 - 1) * LY : LY);
    const int cb_zm = cb + ((z == 0) ? (L2 - 1) * LY : -LY);
    const int LZ = L2 * LY;
    const bool tw_p = (t + 1 == L3), tw_m = (t == 0);
    const int cb_tp = cb + (tw_p ? -(L3 - 1) * LZ : LZ);
    const int cb_tm = cb + (tw_m ? (L3 - 1) * LZ : -LZ);
    const T bc = T(g.bc);

    Cx<T> R[12];
#pragma unroll
    for (int k = 0; k < 12; k++) R[k] = {T(0), T(0)};

    constexpr int FW = DAG ? 1 : 0;   // projector of the forward hop
    constexpr int BW = DAG ? 0 : 1;   // projector of the backward hop
    hop_dir<B, 0, FW, false>(R, a.in, in_q + cb_xp, a.U, poff + cb, T(1));
    hop_dir<B, 0, BW, true>(R, a.in, in_q + cb_xm, a.U, qoff + cb_xm, T(1));
    hop_dir<B, 1, FW, false>(R, a.in, in_q + cb_yp, a.U, poff + cb, T(1));
    hop_dir<B, 1, BW, true>(R, a.in, in_q + cb_ym, a.U, qoff + cb_ym, T(1));
    hop_dir<B, 2, FW, false>(R, a.in, in_q + cb_zp, a.U, poff + cb, T(1));
    hop_dir<B, 2, BW, true>(R, a.in, in_q + cb_zm, a.U, qoff + cb_zm, T(1));
    hop_dir<B, 3, FW, false>(R, a.in, in_q + cb_tp, a.U, poff + cb, tw_p ? bc : T(1));
    hop_dir<B, 3, BW, true>(R, a.in, in_q + cb_tm, a.U, qoff + cb_tm, tw_m ? bc : T(1));

    Cx<T> res[12];
    if constexpr (AUX) {
        Cx<T> xv[12];
        a.aux.load(a.aux_off[p] + cb, xv);
        epilogue<B>(a.epi, R, xv, res);
    } else {
        epilogue<B, T>(a.epi, R, nullptr, res);
    }
    a.out.store(a.out_off[p] + cb, res);
}

// Site-local combination: out = cR y + gR i g5 y + cX x + gX i g5 x (g5out),
// optionally restricted to one parity with zeros written on the other.
template <class S, int B>
__global__ __launch_bounds__(256) void site_kernel(S out, S y, S x, int V, int Vh, int has_x,
                                                   int keep_par, Epi e)
{
    using T = typename S::Real;
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= V) return;
    Cx<T> res[12];
    if (keep_par >= 0 && ((s >= Vh) ? 1 : 0) != keep_par) {
#pragma unroll
        for (int k = 0; k < 12; k++) res[k] = {T(0), T(0)};
    } else {
        Cx<T> yv[12];
        y.load(s, yv);
        if (has_x) {
            Cx<T> xv[12];
            x.load(s, xv);
            epilogue<B>(e, yv, xv, res);
        } else {
            epilogue<B, T>(e, yv, nullptr, res);
        }
    }
    out.store(s, res);
}

// ---------------------------------------------------------------------------
// host side
int g_xcd_remap = 1;
int g_block = 256;
const int *g_skip = nullptr;

template <class S, class G, int B>
int launch_hop_t(qkx_spinor_s *out, const qkx_spinor_s *in, const qkx_spinor_s *aux,
                 const qkx_gauge_s *U, int dagger, int par_out, const Epi &e)
{
    HopArgs<S, S, S, G> a;
    a.g = qkx::make_geom(U->L, U->bc);
    a.in = make_spinor<S>(in);
    a.out = make_spinor<S>(out);
    a.aux = aux ? make_spinor<S>(aux) : a.out;
    a.U = make_gauge<G>(U);
    qkx::parity_offsets(in, a.in_off);
    qkx::parity_offsets(out, a.out_off);
    if (aux) qkx::parity_offsets(aux, a.aux_off);
    else a.aux_off[0] = a.aux_off[1] = 0;
    a.par_out = par_out;
    a.nout = par_out < 0 ? a.g.V : a.g.Vh;
    a.epi = e;
    a.skip = g_skip;
    int nb = (a.nout + g_block - 1) / g_block;
    if (g_xcd_remap && nb >= 64) {
        nb = (nb + 7) / 8 * 8;
        a.nblk8 = nb / 8;
    } else {
        a.nblk8 = 0;
    }
    // bytes moved per launch (compulsory traffic) for the instrumentation
    const int nout = a.nout;
    double ws = (double)qkx::spinor_elem_bytes(in->prec) * qkx::spinor_nchunks(in->prec) +
                (in->prec == QKX_PREC_HALF ? 4.0 : 0.0);
    double wg = (U->prec == QKX_PREC_DOUBLE ? 16.0 : U->prec == QKX_PREC_SINGLE ? 8.0 : 4.0) *
                (U->recon == QKX_RECON_18 ? 9 : 6);
    double bytes = par_out < 0 ? (double)nout * (4 * wg + 2 * ws)
                               : (double)nout * (8 * wg + 2 * ws);
    if (aux) bytes += (double)nout * ws;
    qkx::prof_begin();
    if (dagger) {
        if (aux) hipLaunchKernelGGL((hop_kernel<S, S, S, G, B, true, true>), dim3(nb), dim3(g_block), 0, qkx::g_stream, a);
        else     hipLaunchKernelGGL((hop_kernel<S, S, S, G, B, true, false>), dim3(nb), dim3(g_block), 0, qkx::g_stream, a);
    } else {
        if (aux) hipLaunchKernelGGL((hop_kernel<S, S, S, G, B, false, true>), dim3(nb), dim3(g_block), 0, qkx::g_stream, a);
        else     hipLaunchKernelGGL((hop_kernel<S, S, S, G, B, false, false>), dim3(nb), dim3(g_block), 0, qkx::g_stream, a);
    }
    qkx::prof_end(bytes);
    QKX_HIP(hipGetLastError());
    return QKX_SUCCESS;
}

template <class S, int B>
int launch_site_t(qkx_spinor_s *out, const qkx_spinor_s *y, const qkx_spinor_s *x, int keep_par, const Epi &e)
{
    int n = out->nsites;
    int nb = (n + 255) / 256;
    S so = make_spinor<S>(out), sy = make_spinor<S>(y), sx = x ? make_spinor<S>(x) : so;
    hipLaunchKernelGGL((site_kernel<S, B>), dim3(nb), dim3(256), 0, qkx::g_stream, so, sy, sx, n,
                       out->subset == QKX_SUBSET_FULL ? out->Vh : n, x ? 1 : 0, keep_par, e);
    QKX_HIP(hipGetLastError());
    return QKX_SUCCESS;
}

} // namespace

// ---------------------------------------------------------------------------
// internal entry points used by the C API and the solvers
namespace qkx {

void set_hop_skip(const int *flag) { g_skip = flag; }

void set_hop_tuning(int xcd_remap, int block)
{
    g_xcd_remap = xcd_remap;
    g_block = block;
}

// Hop kernel: out(parity par_out or all) = epilogue(H_or_Hdag in, aux)
int hop(qkx_spinor_s *out, const qkx_spinor_s *in, const qkx_spinor_s *aux, const qkx_gauge_s *U,
        int basis, int dagger, int par_out, double cR, double gR, double cX, double gX, int g5out)
{
    Epi e{cR, gR, cX, gX, g5out};
    const int sp = in->prec;
    if (out->prec != sp || (aux && aux->prec != sp)) {
        set_error("hop: spinor precisions differ");
        return QKX_ERR_ARG;
    }
    if (basis == QKX_BASIS_CHIRAL && sp != QKX_PREC_DOUBLE) {
        set_error("hop: chiral basis is FP64 only");
        return QKX_ERR_ARG;
    }
    if (sp == QKX_PREC_DOUBLE) {
        if (U->prec != QKX_PREC_DOUBLE) { set_error("hop: FP64 spinor needs FP64 gauge"); return QKX_ERR_ARG; }
        if (U->recon == QKX_RECON_18) {
            if (basis == QKX_BASIS_UKQCD) return launch_hop_t<SpinorD, GaugeD18, QKX_BASIS_UKQCD>(out, in, aux, U, dagger, par_out, e);
            return launch_hop_t<SpinorD, GaugeD18, QKX_BASIS_CHIRAL>(out, in, aux, U, dagger, par_out, e);
        } else {
            if (basis == QKX_BASIS_UKQCD) return launch_hop_t<SpinorD, GaugeD12, QKX_BASIS_UKQCD>(out, in, aux, U, dagger, par_out, e);
            return launch_hop_t<SpinorD, GaugeD12, QKX_BASIS_CHIRAL>(out, in, aux, U, dagger, par_out, e);
        }
    }
    if (sp == QKX_PREC_SINGLE) {
        if (U->prec == QKX_PREC_SINGLE && U->recon == QKX_RECON_18)
            return launch_hop_t<SpinorF, GaugeF18, QKX_BASIS_UKQCD>(out, in, aux, U, dagger, par_out, e);
        if (U->prec == QKX_PREC_SINGLE && U->recon == QKX_RECON_12)
            return launch_hop_t<SpinorF, GaugeF12, QKX_BASIS_UKQCD>(out, in, aux, U, dagger, par_out, e);
    }
    if (sp == QKX_PREC_HALF) {
        if (U->prec == QKX_PREC_HALF && U->recon == QKX_RECON_12)
            return launch_hop_t<SpinorH, GaugeH12, QKX_BASIS_UKQCD>(out, in, aux, U, dagger, par_out, e);
        if (U->prec == QKX_PREC_SINGLE && U->recon == QKX_RECON_12)
            return launch_hop_t<SpinorH, GaugeF12, QKX_BASIS_UKQCD>(out, in, aux, U, dagger, par_out, e);
    }
    if (sp == QKX_PREC_BF16) {
        if (U->prec == QKX_PREC_BF16 && U->recon == QKX_RECON_12)
            return launch_hop_t<SpinorB, GaugeB12, QKX_BASIS_UKQCD>(out, in, aux, U, dagger, par_out, e);
    }
    set_error("hop: unsupported spinor/gauge precision combination");
    return QKX_ERR_ARG;
}

// Site-local: out = cR y + gR i g5 y + cX x + gX i g5 x (x may be null); keep_par >= 0
// zeroes the other parity of a full field.
int site(qkx_spinor_s *out, const qkx_spinor_s *y, const qkx_spinor_s *x, int basis, int keep_par,
         double cR, double gR, double cX, double gX, int g5out)
{
    Epi e{cR, gR, cX, gX, g5out};
    if (out->prec != y->prec || (x && x->prec != y->prec) || out->nsites != y->nsites ||
        (x && x->nsites != y->nsites)) {
        set_error("site op: fields differ in precision or size");
        return QKX_ERR_ARG;
    }
    switch (out->prec) {
    case QKX_PREC_DOUBLE:
        if (basis == QKX_BASIS_UKQCD) return launch_site_t<SpinorD, QKX_BASIS_UKQCD>(out, y, x, keep_par, e);
        return launch_site_t<SpinorD, QKX_BASIS_CHIRAL>(out, y, x, keep_par, e);
    case QKX_PREC_SINGLE: return launch_site_t<SpinorF, QKX_BASIS_UKQCD>(out, y, x, keep_par, e);
    case QKX_PREC_HALF: return launch_site_t<SpinorH, QKX_BASIS_UKQCD>(out, y, x, keep_par, e);
    case QKX_PREC_BF16: return launch_site_t<SpinorB, QKX_BASIS_UKQCD>(out, y, x, keep_par, e);
    }
    set_error("site op: bad precision");
    return QKX_ERR_ARG;
}

} // namespace qkx
