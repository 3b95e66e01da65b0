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
#include <algorithm>
#include <map>
#include <string>
#include <type_traits>
#include <vector>
#include "qkx_core.h"

// FP64 hop outputs are written with non-temporal stores (1) or plain ones (0)
#ifndef QKX_NT_OUT64
#define QKX_NT_OUT64 1
#endif

namespace {

enum : int { K0 = 0, KP1, KM1, KPI, KMI, K2 };

// Pin the accumulator: the empty asm reads and writes every element of R and
// clobbers memory, so one direction's arithmetic completes before the next
// direction's gathers are issued.  Without it the compiler issues all sixteen
// gathers first (~240 registers of in-flight data, spills for FP64).
template <class T> __device__ __forceinline__ void pin_acc(Cx<T> R[12])
{
    asm volatile("" : "+v"(R[0].re), "+v"(R[0].im), "+v"(R[1].re), "+v"(R[1].im), "+v"(R[2].re),
                      "+v"(R[2].im), "+v"(R[3].re), "+v"(R[3].im), "+v"(R[4].re), "+v"(R[4].im),
                      "+v"(R[5].re), "+v"(R[5].im)::"memory");
    asm volatile("" : "+v"(R[6].re), "+v"(R[6].im), "+v"(R[7].re), "+v"(R[7].im), "+v"(R[8].re),
                      "+v"(R[8].im), "+v"(R[9].re), "+v"(R[9].im), "+v"(R[10].re), "+v"(R[10].im),
                      "+v"(R[11].re), "+v"(R[11].im)::"memory");
}
#define QKX_SCHED_FENCE() pin_acc(R)

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

// acc + n K z for the reconstruction constant K: one FMA per real, so a
// per-site scale n costs nothing on top of the accumulation
template <int K, class T> __device__ __forceinline__ Cx<T> umul_acc(Cx<T> acc, Cx<T> z, T n)
{
    if constexpr (K == KP1) return {fma(n, z.re, acc.re), fma(n, z.im, acc.im)};
    else if constexpr (K == KM1) return {fma(-n, z.re, acc.re), fma(-n, z.im, acc.im)};
    else if constexpr (K == KPI) return {fma(-n, z.im, acc.re), fma(n, z.re, acc.im)};
    else if constexpr (K == KMI) return {fma(n, z.im, acc.re), fma(-n, z.re, acc.im)};
    else if constexpr (K == K2) return {fma(n + n, z.re, acc.re), fma(n + n, z.im, acc.im)};
    else return acc;
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

// R_s += n q_s xi[kq_s]: the reconstruction with a scale folded in
template <int B, int MU, int KIND, int s, class T>
__device__ __forceinline__ void collect_scaled(Cx<T> R[12], const Cx<T> xi[6], T n)
{
    constexpr Proj P = kProj[B][MU][KIND];
    if constexpr (P.q[s] != K0) {
#pragma unroll
        for (int c = 0; c < 3; c++) R[s * 3 + c] = umul_acc<P.q[s]>(R[s * 3 + c], xi[P.kq[s] * 3 + c], n);
    }
}

// xi = U h (DAGL=false) or U^dag h (DAGL=true), h holds two colour vectors.
// FMA: explicit FMA chains -- fewer instructions but less instruction-level
// parallelism; chosen for the 16-bit hops, which the ALU bounds (the FP32 and
// FP64 hops are faster with the independent products).
template <bool DAGL, bool FMA = false, class T>
__device__ __forceinline__ void link_mul(const Cx<T> U[9], const Cx<T> h[6], Cx<T> xi[6])
{
#pragma unroll
    for (int k = 0; k < 2; k++)
#pragma unroll
        for (int a = 0; a < 3; a++) {
            Cx<T> acc;
            if constexpr (FMA && !DAGL) {
                acc = cmul_f(U[a * 3 + 0], h[k * 3 + 0]);
                acc = cmac(acc, U[a * 3 + 1], h[k * 3 + 1]);
                acc = cmac(acc, U[a * 3 + 2], h[k * 3 + 2]);
            } else if constexpr (FMA) {
                acc = cjmul_f(U[0 * 3 + a], h[k * 3 + 0]);
                acc = cjmac(acc, U[1 * 3 + a], h[k * 3 + 1]);
                acc = cjmac(acc, U[2 * 3 + a], h[k * 3 + 2]);
            } else if constexpr (!DAGL) {
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

// 16-bit spinors: the hop is bound by ALU instructions, not memory
template <class S> constexpr bool kFmaChain = std::is_same_v<S, SpinorH> || std::is_same_v<S, SpinorB>;

// Raw operands of one hop (the neighbour spin pairs the projector reads + link).
template <class SIn, class G> struct HopOps {
    typename SIn::Raw psi;
    typename G::Raw link;
};

template <int B, int MU, int KIND> constexpr int spin_mask()
{
    return (uses_spin_pair<B, MU, KIND>(0) ? 1 : 0) | (uses_spin_pair<B, MU, KIND>(1) ? 2 : 0);
}

template <int B, int MU, int KIND, class SIn, class G>
__device__ __forceinline__ HopOps<SIn, G> fetch_dir(const SIn &in, int s_nb, const G &U, int s_link)
{
    HopOps<SIn, G> o;
    o.psi = in.template fetch<spin_mask<B, MU, KIND>()>(s_nb);
    o.link = U.fetch(MU, s_link);
    return o;
}

// ---- FP16 spinors with FP16 links: packed half arithmetic into dot2 --------
// v_dot2_f32_f16 (d = a.lo b.lo + a.hi b.hi + c, FP32 accumulation) takes a
// complex product's real or imaginary part straight from two (re, im) half
// pairs: re(u h) = dot2(u, (h.re, -h.im)), im(u h) = dot2(u, (h.im, h.re)),
// re(conj(u) h) = dot2(u, h), im(conj(u) h) = dot2(u, (h.im, -h.re)).  The
// spinor and the first two link rows are used as stored (no conversions).
// Each needed view of the projected half spinor is formed in half precision
// by one packed add with source modifiers (one rounding, the size of the
// storage rounding); the third link row is reconstructed with dot2 and
// rounded to half.  Written with the compiler builtin, whose hazard tracking
// keeps the required wait states between dependent dot instructions (inline
// assembly would hide them).
typedef _Float16 h2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ h2v h2_of(uint32_t w) { return __builtin_bit_cast(h2v, w); }
__device__ __forceinline__ float dot2(h2v a, h2v b, float c) { return __builtin_amdgcn_fdot2(a, b, c, false); }
__device__ __forceinline__ h2v cj16(h2v z) { return h2v{z.x, -z.y}; }
__device__ __forceinline__ h2v sw16(h2v z) { return h2v{z.y, z.x}; }
__device__ __forceinline__ h2v swn16(h2v z) { return h2v{z.y, -z.x}; }
template <int K> __device__ __forceinline__ h2v kmul16(h2v b)
{
    if constexpr (K == KP1) return b;
    else if constexpr (K == KM1) return h2v{-b.x, -b.y};
    else if constexpr (K == KPI) return h2v{-b.y, b.x};
    else if constexpr (K == KMI) return h2v{b.y, -b.x};
    else return h2v{(_Float16)0, (_Float16)0};
}

// views of projected component k: (h, (h.im, -h.re)) for a backward hop,
// ((h.re, -h.im), (h.im, h.re)) for a forward one, h = psi[sa] + b psi[sb]
// (a = 2 is left to the caller's scale)
template <int B, int MU, int KIND, bool DAGL, int k>
__device__ __forceinline__ void views16(const uint32_t w[12], h2v v0[6], h2v v1[6])
{
    constexpr Proj P = kProj[B][MU][KIND];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const h2v a = h2_of(w[P.sa[k] * 3 + c]);
        if constexpr (P.b[k] == K0) {
            v0[k * 3 + c] = DAGL ? a : cj16(a);
            v1[k * 3 + c] = DAGL ? swn16(a) : sw16(a);
        } else {
            const h2v kb = kmul16<P.b[k]>(h2_of(w[P.sb[k] * 3 + c]));
            if constexpr (DAGL) {
                v0[k * 3 + c] = a + kb;
                v1[k * 3 + c] = swn16(a) + swn16(kb);
            } else {
                v0[k * 3 + c] = cj16(a) + cj16(kb);
                v1[k * 3 + c] = sw16(a) + sw16(kb);
            }
        }
    }
}

// The 9 (re, im) half pairs of a stored FP16 12-real link: rows 0 and 1 as
// stored, row 2 = conj(row0 x row1) formed with dot2 and rounded to half.
__device__ __forceinline__ void link_rows16(const typename Gauge16<false>::Raw &lk, h2v U[9])
{
    U[0] = h2_of(lk.a.x);
    U[1] = h2_of(lk.a.y);
    U[2] = h2_of(lk.a.z);
    U[3] = h2_of(lk.a.w);
    U[4] = h2_of(lk.b.x);
    U[5] = h2_of(lk.b.y);
    // third row conj(x y - z w): views of row 1 and the negated row 0, once
    const h2v c3 = cj16(U[3]), c4 = cj16(U[4]), c5 = cj16(U[5]);
    const h2v s3 = sw16(U[3]), s4 = sw16(U[4]), s5 = sw16(U[5]);
    const h2v n0 = -U[0], n1 = -U[1], n2 = -U[2];
    auto cross = [&](h2v x, h2v nx, h2v yc, h2v ys, h2v nz, h2v z, h2v wc, h2v ws) {
        const float re = dot2(x, yc, dot2(nz, wc, 0.f));
        const float im = dot2(z, ws, dot2(nx, ys, 0.f));
        return h2v{(_Float16)re, (_Float16)im};
    };
    U[6] = cross(U[1], n1, c5, s5, n2, U[2], c4, s4);
    U[7] = cross(U[2], n2, c3, s3, n0, U[0], c5, s5);
    U[8] = cross(U[0], n0, c4, s4, n1, U[1], c3, s3);
}

// One hop of an FP16 neighbour with the link given as its 9 half pairs.
template <int B, int MU, int KIND, bool DAGL, class T>
__device__ __forceinline__ void apply_rows_h16(Cx<T> R[12], const typename SpinorH::Raw &ps, const h2v U[9],
                                               T sign)
{
    constexpr Proj P = kProj[B][MU][KIND];
    static_assert(P.a[0] == P.a[1] && (P.a[0] == KP1 || P.a[0] == K2), "projector shape");
    uint32_t w[12];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        w[4 * j + 0] = ps.c[j].x;
        w[4 * j + 1] = ps.c[j].y;
        w[4 * j + 2] = ps.c[j].z;
        w[4 * j + 3] = ps.c[j].w;
    }
    h2v v0[6], v1[6];
    views16<B, MU, KIND, DAGL, 0>(w, v0, v1);
    views16<B, MU, KIND, DAGL, 1>(w, v0, v1);
    Cx<T> xi[6];
#pragma unroll
    for (int k = 0; k < 2; k++)
#pragma unroll
        for (int a = 0; a < 3; a++) {
            // forward: sum_b U[a][b] h[b]; backward: sum_b conj(U[b][a]) h[b]
            const h2v u0 = DAGL ? U[a] : U[3 * a], u1 = DAGL ? U[3 + a] : U[3 * a + 1],
                      u2 = DAGL ? U[6 + a] : U[3 * a + 2];
            xi[k * 3 + a] = {dot2(u2, v0[k * 3 + 2], dot2(u1, v0[k * 3 + 1], dot2(u0, v0[k * 3], 0.f))),
                             dot2(u2, v1[k * 3 + 2], dot2(u1, v1[k * 3 + 1], dot2(u0, v1[k * 3], 0.f)))};
        }
    T n = ps.n * sign;
    if constexpr (P.a[0] == K2) n = n + n;
    collect_scaled<B, MU, KIND, 0>(R, xi, n);
    collect_scaled<B, MU, KIND, 1>(R, xi, n);
    collect_scaled<B, MU, KIND, 2>(R, xi, n);
    collect_scaled<B, MU, KIND, 3>(R, xi, n);
}

template <int B, int MU, int KIND, bool DAGL, class T>
__device__ __forceinline__ void apply_dir_h16(Cx<T> R[12], const typename SpinorH::Raw &ps,
                                              const typename Gauge16<false>::Raw &lk, T sign)
{
    h2v U[9];
    link_rows16(lk, U);
    apply_rows_h16<B, MU, KIND, DAGL>(R, ps, U, sign);
}

#ifndef QKX_HOP16_DOT2
#define QKX_HOP16_DOT2 1
#endif

// Project, multiply by the link (or its adjoint), reconstruct, accumulate.
template <int B, int MU, int KIND, bool DAGL, class SIn, class G, class T>
__device__ __forceinline__ void apply_dir(Cx<T> R[12], const SIn &in, const G &U, const HopOps<SIn, G> &o,
                                          T sign)
{
    Cx<T> psi[12];
    if constexpr (QKX_HOP16_DOT2 && std::is_same_v<SIn, SpinorH> && std::is_same_v<G, Gauge16<false>>) {
        apply_dir_h16<B, MU, KIND, DAGL>(R, o.psi, o.link, sign);
        return;
    } else if constexpr (std::is_same_v<SIn, SpinorH>) {
        // FP16 neighbours carry a per-site scale: project and multiply the
        // unscaled values and fold scale and boundary sign into the final FMAs
        // (24 multiplies fewer per hop in an ALU-bound kernel)
        in.template unpack_raw<spin_mask<B, MU, KIND>()>(o.psi, psi);
        Cx<T> h[6];
        project_k<B, MU, KIND, 0>(psi, h);
        project_k<B, MU, KIND, 1>(psi, h);
        Cx<T> Ul[9];
        U.unpack(o.link, Ul);
        Cx<T> xi[6];
        link_mul<DAGL, kFmaChain<SIn>>(Ul, h, xi);
        const T n = in.scale(o.psi) * sign;
        collect_scaled<B, MU, KIND, 0>(R, xi, n);
        collect_scaled<B, MU, KIND, 1>(R, xi, n);
        collect_scaled<B, MU, KIND, 2>(R, xi, n);
        collect_scaled<B, MU, KIND, 3>(R, xi, n);
        return;
    }
    in.template unpack<spin_mask<B, MU, KIND>()>(o.psi, psi);
    Cx<T> h[6];
    project_k<B, MU, KIND, 0>(psi, h);
    project_k<B, MU, KIND, 1>(psi, h);
    Cx<T> Ul[9];
    U.unpack(o.link, Ul);
    if (sign != T(1)) {
#pragma unroll
        for (int k = 0; k < 6; k++) h[k] = sign * h[k];
    }
    Cx<T> xi[6];
    link_mul<DAGL, kFmaChain<SIn>>(Ul, h, xi);
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
    // every term is always formed (a zero coefficient adds an exact zero), as
    // explicit FMAs in a fixed order, so that no per-component selects are
    // needed and every kernel that shares this epilogue rounds identically
    const T cR = T(e.cR), gR = T(e.gR), cX = T(e.cX), gX = T(e.gX);
    Cx<T> g5R[12], g5x[12];
    gamma5<B>(R, g5R);
    if (x) gamma5<B>(x, g5x);
#pragma unroll
    for (int k = 0; k < 12; k++) {
        // cR R + gR i g5R (+ cX x + gX i g5x), i (a, b) = (-b, a)
        T re = fma(gR, -g5R[k].im, cR * R[k].re);
        T im = fma(gR, g5R[k].re, cR * R[k].im);
        if (x) {
            re = fma(gX, -g5x[k].im, fma(cX, x[k].re, re));
            im = fma(gX, g5x[k].re, fma(cX, x[k].im, im));
        }
        out[k] = {re, im};
    }
    if (e.g5out) {
        Cx<T> w[12];
        gamma5<B>(out, w);
#pragma unroll
        for (int k = 0; k < 12; k++) out[k] = w[k];
    }
}

// Site order of the hop kernels.  Plain: checkerboard index order (x fastest,
// then y, z, t).  Blocked: x, then yb rows of y, then tb slices of t, then z,
// then the next y block, then the next t block, so a site's +-y, +-z and +-t
// neighbours (and the backward links they carry) are reused after about
// L0 yb tb sites instead of a whole t slice; yb and tb are chosen on the
// host so that this reuse window stays well inside one XCD's L2.
struct HopTrav {
    int on;        // 0: plain order
    int yb, tb, nyb;
    FastDiv dyb, dtb, dnyb;
};

template <class SIn, class SOut, class SAux, class G, int NEPI = QKX_MAX_RHS>
struct HopArgs {
    Geom g;
    SIn in;
    SOut out;
    SAux aux;
    G U;
    int in_off[2], out_off[2], aux_off[2];
    int par_out;   // -1: all sites; 0/1: that parity only
    int nout;      // output entries (sites x right-hand sides)
    int nrhs;      // right-hand sides; thread unit u -> site u / nrhs, rhs u % nrhs
    FastDiv dnr;   // division by nrhs
    int nblk8;     // grid size / 8 when XCD remapping is on, else 0
    HopTrav tr;    // order in which the sites of a parity are walked
    const int *skip;   // when non-null and *skip != 0 the launch does nothing
    double *norm_partial;  // when non-null: per-workgroup (and per-rhs) sum of |out|^2
    Epi epi[NEPI];         // epilogue coefficients per right-hand side
};

// Sum v over the workgroup; thread 0 writes the result to dst.
__device__ __forceinline__ void block_partial(double v, double *dst)
{
    __shared__ double lds[4];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += lds[w];
        *dst = s;
    }
}

// Per right-hand side sums of a workgroup's values, in thread order (fixed,
// whatever order the waves run in): thread t's value belongs to right-hand
// side (u0 + t) mod nr, u0 the unit index of thread 0 modulo nr.  Threads
// 0 .. nr-1 write dst[0 .. nr-1].
__device__ __forceinline__ void block_partial_rhs(double v, int nr, int u0, double *dst)
{
    __shared__ double vals[256];
    vals[threadIdx.x] = v;
    __syncthreads();
    if ((int)threadIdx.x < nr) {
        double s = 0;
        int t = (int)threadIdx.x - u0;   // threadIdx.x and u0 are both below nr
        if (t < 0) t += nr;
        for (; t < (int)blockDim.x; t += nr) s += vals[t];
        dst[threadIdx.x] = s;
    }
}

// Hardware workgroup IDs are dealt round-robin to the 8 XCDs; remap them so
// that each XCD walks one contiguous range of sites and its private L2 sees
// the neighbour reuse of that range.
__device__ __forceinline__ int logical_block(int b, int nblk8)
{
    if (nblk8 > 0) b = (b & 7) * nblk8 + (b >> 3);
    return b;
}

// Parity and unit index of this thread; a unit is one (site, group of
// right-hand sides) of the output parity, `upsite` units per site.  bid is
// the workgroup's index within its launch (or within its instance of a
// multi-instance launch).
__device__ __forceinline__ void hop_unit(const Geom &g, int par_out, int nblk8, int upsite, int &p, int &u, int bid)
{
    if (par_out >= 0) {
        p = par_out;
        u = logical_block(bid, nblk8) * blockDim.x + threadIdx.x;
    } else if (nblk8 > 0) {
        // full lattice: XCD k owns one slab of checkerboard indices and
        // alternates between the two parities, so the links and spinors of a
        // slab (both parities) are fetched into one L2 only
        const int b = bid, i = b >> 3;
        p = i & 1;
        u = ((b & 7) * nblk8 + (i >> 1)) * blockDim.x + threadIdx.x;
    } else {
        const int o = bid * blockDim.x + threadIdx.x;
        const int vhn = g.Vh * upsite;
        p = (o >= vhn) ? 1 : 0;
        u = o - p * vhn;
    }
}

// Checkerboard indices of the 8 neighbours of (p, cb) in the other parity,
// in the order +x -x +y -y +z -z +t -t, and whether the t hops wrap.
struct HopSite {
    int nb[8];
    bool tw_p, tw_m;
};
__device__ __forceinline__ HopSite hop_site_at(const Geom &g, int p, int cb, int xh, int y, int z, int t)
{
    HopSite h;
    const int L0 = g.L[0], L1 = g.L[1], L2 = g.L[2], L3 = g.L[3];
    const int row = cb - xh;   // Lh0 * (y + L1 (z + L2 t))
    const int x = 2 * xh + ((p + y + z + t) & 1);
    const int xp = (x + 1 == L0) ? 0 : x + 1;
    const int xm = (x == 0) ? L0 - 1 : x - 1;
    h.nb[0] = (xp >> 1) + row;
    h.nb[1] = (xm >> 1) + row;
    h.nb[2] = cb + g.Lh0 * ((y + 1 == L1) ? -(L1 - 1) : 1);
    h.nb[3] = cb + g.Lh0 * ((y == 0) ? (L1 - 1) : -1);
    const int LY = L1 * g.Lh0;
    h.nb[4] = cb + ((z + 1 == L2) ? -(L2 - 1) * LY : LY);
    h.nb[5] = cb + ((z == 0) ? (L2 - 1) * LY : -LY);
    const int LZ = L2 * LY;
    h.tw_p = (t + 1 == L3);
    h.tw_m = (t == 0);
    h.nb[6] = cb + (h.tw_p ? -(L3 - 1) * LZ : LZ);
    h.nb[7] = cb + (h.tw_m ? (L3 - 1) * LZ : -LZ);
    return h;
}

__device__ __forceinline__ HopSite hop_site(const Geom &g, int p, int cb)
{
    const int rest = fdiv(cb, g.dLh0);
    const int xh = cb - rest * g.Lh0;
    const int zt = fdiv(rest, g.dL1);
    const int y = rest - zt * g.L[1];
    const int t = fdiv(zt, g.dL2);
    const int z = zt - t * g.L[2];
    return hop_site_at(g, p, cb, xh, y, z, t);
}

// The v-th site of parity p in the walk order of tr: its checkerboard index
// cb and its neighbours.  v < Vh.
__device__ __forceinline__ HopSite hop_site_walk(const Geom &g, const HopTrav &tr, int p, int v, int &cb)
{
    if (!tr.on) {
        cb = v;
        return hop_site(g, p, v);
    }
    const int q1 = fdiv(v, g.dLh0), xh = v - q1 * g.Lh0;
    const int q2 = fdiv(q1, tr.dyb), yi = q1 - q2 * tr.yb;
    const int q3 = fdiv(q2, tr.dtb), ti = q2 - q3 * tr.tb;
    const int q4 = fdiv(q3, g.dL2), z = q3 - q4 * g.L[2];
    const int tbk = fdiv(q4, tr.dnyb), ybk = q4 - tbk * tr.nyb;
    const int y = ybk * tr.yb + yi, t = tbk * tr.tb + ti;
    cb = xh + g.Lh0 * (y + g.L[1] * (z + g.L[2] * t));
    return hop_site_at(g, p, cb, xh, y, z, t);
}

// What a hop thread does with its result.  StoreFin: write it (and, on
// request, per-workgroup |out|^2 partials).  CgFin: the CG update fused into
// the last hop of the normal operator -- with alpha = r2 / p.Ap (p.Ap summed
// from the partials of an earlier hop, in a fixed order), x += alpha p and
// r -= alpha Ap at this site, per-workgroup |r|^2 partials; Ap itself is never
// stored.  x, r, p share the output field's layout.  UPDX false: x is left to
// the direction kernel (cg_dir_kernel), which reads p there anyway, and only
// r is updated here.
struct StoreFin {
    template <class A> __device__ __forceinline__ void begin(const A &) {}
};

template <class SX, class SR, class SP, bool UPDX = true> struct CgFin {
    static constexpr bool kUpdX = UPDX;
    SX x;
    SR r;
    SP p;
    const double *r2cur;      // |r|^2 of the current iteration (device)
    const double *pap_part;   // p.Ap partials
    int npap;
    double *r2_part;          // out: per-workgroup |r|^2 after the update
    double alpha;
    // p.Ap, compensated (double-double), in the same fixed order in every
    // workgroup: strided per-thread sums, an xor butterfly (the same bits in
    // every lane), then the waves in order
    template <class A> __device__ __forceinline__ void begin(const A &)
    {
        __shared__ double lh[4], ll[4];
        DD v{0.0, 0.0};
        for (int i = threadIdx.x; i < npap; i += blockDim.x) v = dd_add(v, pap_part[i]);
        v = dd_wave_sum(v);
        if ((threadIdx.x & 63) == 0) {
            lh[threadIdx.x >> 6] = v.hi;
            ll[threadIdx.x >> 6] = v.lo;
        }
        __syncthreads();
        DD pap{lh[0], ll[0]};
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) pap = dd_add(pap, DD{lh[w], ll[w]});
        alpha = *r2cur / dd_value(pap);
        __syncthreads();
    }
};

// What a batched hop (hop_share_kernel) does with its result in a CG
// iteration: alpha_r = |r_r|^2 / (p.Ap)_r (0 for a frozen right-hand side:
// not active, or |r|^2 at its stopping threshold), x += alpha p and
// r -= alpha Ap at this (site, RHS), per-workgroup per-RHS |r|^2 partials;
// Ap itself is never stored.  pap is one row of per-RHS sums.  UPDX false:
// r only, x is left to the direction kernel (cgb_dir_kernel).
template <class SX, class SR, class SP, bool UPDX = true> struct BCgFin {
    static constexpr bool kUpdX = UPDX;
    SX x;
    SR r;
    SP p;
    const double *r2cur, *stop2;
    const int *active;
    const double *pap;
    double *r2_part;
};
// bytes of one site of a spinor field in each storage (FP16: 48 B and the scale)
template <class S> constexpr double site_bytes_of()
{
    if constexpr (std::is_same_v<S, SpinorD>) return 192.0;
    else if constexpr (std::is_same_v<S, SpinorF>) return 96.0;
    else if constexpr (std::is_same_v<S, SpinorH>) return 52.0;
    else return 48.0;
}

template <class F> struct IsBCg : std::false_type {};
template <class SX, class SR, class SP, bool U> struct IsBCg<BCgFin<SX, SR, SP, U>> : std::true_type {};
// whether a fused CG update also updates x (and so reads p)
template <class F> struct FinUpdX { static constexpr bool v = true; };
template <class SX, class SR, class SP> struct FinUpdX<CgFin<SX, SR, SP, false>> { static constexpr bool v = false; };
template <class SX, class SR, class SP> struct FinUpdX<BCgFin<SX, SR, SP, false>> { static constexpr bool v = false; };

// Minimum waves per SIMD requested from the register allocator for the 16-bit
// hops (1: no request).  At 24^3x48 a half-lattice launch is 1296 workgroups;
// residency of 6 waves per SIMD holds all of them at once.
#ifndef QKX_HOP16_WAVES
#define QKX_HOP16_WAVES 1
#endif
#ifndef QKX_HOP64_WAVES
#define QKX_HOP64_WAVES 1
#endif
template <class S> struct HopWaves { static constexpr int v = 1; };
template <> struct HopWaves<SpinorD> { static constexpr int v = QKX_HOP64_WAVES; };
template <> struct HopWaves<SpinorH> { static constexpr int v = QKX_HOP16_WAVES; };
template <> struct HopWaves<SpinorB> { static constexpr int v = QKX_HOP16_WAVES; };

// The epilogue of output unit (cb, rr) of parity p and what Fin does with it:
// store it (adding |out|^2 to nrm2 when partials are requested) or apply
// the fused CG update (adding |r|^2 to nrm2).
template <int B, bool AUX, class SIn, class SOut, class SAux, class G, int NEPI, class Fin, class T>
__device__ __forceinline__ void hop_finish(const HopArgs<SIn, SOut, SAux, G, NEPI> &a, Fin &fin, int p, int cb,
                                           int rr, const Epi &ep, const Cx<T> R[12], double &nrm2)
{
    constexpr bool CG = !std::is_same_v<Fin, StoreFin>;
    const int nr = a.nrhs;
    Cx<T> res[12];
    if constexpr (AUX) {
        Cx<T> xv[12];
        a.aux.load(a.aux_off[p] + cb * nr + rr, xv);
        epilogue<B>(ep, R, xv, res);
    } else {
        epilogue<B, T>(ep, R, nullptr, res);
    }
    if constexpr (CG) {
        // x += alpha p (unless left to the direction kernel), r -= alpha Ap
        // (single right-hand side)
        const int so = a.out_off[p] + cb;
        using TX = typename decltype(fin.x)::Real;
        using TR = typename decltype(fin.r)::Real;
        using TP = typename decltype(fin.p)::Real;
        if constexpr (Fin::kUpdX) {
            Cx<TX> xv[12];
            Cx<TP> pv[12];
            fin.x.load(so, xv);
            fin.p.load(so, pv);
            const TX ax = TX(fin.alpha);
#pragma unroll
            for (int k = 0; k < 12; k++) xv[k] = {xv[k].re + ax * TX(pv[k].re), xv[k].im + ax * TX(pv[k].im)};
            fin.x.store(so, xv);
        }
        Cx<TR> rv[12];
        fin.r.load(so, rv);
        const TR ar = TR(fin.alpha);
#pragma unroll
        for (int k = 0; k < 12; k++) {
            rv[k] = {rv[k].re - ar * TR(res[k].re), rv[k].im - ar * TR(res[k].im)};
            nrm2 += (double)rv[k].re * (double)rv[k].re + (double)rv[k].im * (double)rv[k].im;
        }
        fin.r.store(so, rv);
    } else {
        if constexpr (std::is_same_v<SOut, SpinorD> && QKX_NT_OUT64)
            a.out.store_nt(a.out_off[p] + cb * nr + rr, res);
        else
            a.out.store(a.out_off[p] + cb * nr + rr, res);
        if (a.norm_partial) {
#pragma unroll
            for (int k = 0; k < 12; k++)
                nrm2 += (double)res[k].re * (double)res[k].re + (double)res[k].im * (double)res[k].im;
        }
    }
}

// The workgroup's norm partials (one per right-hand side) from each
// thread's nrm2; u is the thread's unit.  Every thread of the workgroup
// calls it.
template <class A, class Fin>
__device__ __forceinline__ void hop_partials(const A &a, Fin &fin, int bid, int u, double nrm2)
{
    constexpr bool CG = !std::is_same_v<Fin, StoreFin>;
    const int nr = a.nrhs;
    if constexpr (CG) {
        block_partial(nrm2, fin.r2_part + bid);
    } else if (a.norm_partial) {
        if (nr == 1) {
            block_partial(nrm2, a.norm_partial + bid);
        } else {
            // one partial per right-hand side, summed in thread order
            const int ub = u - (int)threadIdx.x;   // the workgroup's first unit, >= 0
            const int u0 = ub - fdiv(ub, a.dnr) * nr;
            block_partial_rhs(nrm2, nr, u0, a.norm_partial + bid * nr);
        }
    }
}

// The work of one hop workgroup: bid is its index within the launch (or
// within its instance of a multi-instance launch), (p, u) its thread's parity
// and unit (site x rhs) index within it (hop_unit), ep the epilogue of the
// thread's right-hand side.  The caller selects ep from its own argument
// block: a dynamic index into the argument structure here would make the
// compiler copy the whole structure to scratch memory.
template <int B, bool DAG, bool AUX, int DEPTH, class SIn, class SOut, class SAux, class G, int NEPI, class Fin>
__device__ __forceinline__ void hop_body(const HopArgs<SIn, SOut, SAux, G, NEPI> &a, Fin &fin, int bid, int p, int u,
                                         const Epi &ep)
{
    using T = typename SOut::Real;
    const Geom &g = a.g;
    const int nr = a.nrhs;
    const int v = nr == 1 ? u : fdiv(u, a.dnr);   // position of the site in the walk
    const int rr = u - v * nr;
    double nrm2 = 0.0;
    int cb = v;
    if (v < g.Vh) {
    const HopSite hs = hop_site_walk(g, a.tr, p, v, cb);
    const int q = p ^ 1;
    const int qoff = q * g.Vh, poff = p * g.Vh;
    const int in_q = a.in_off[q];
    const int cb_xp = hs.nb[0], cb_xm = hs.nb[1], cb_yp = hs.nb[2], cb_ym = hs.nb[3];
    const int cb_zp = hs.nb[4], cb_zm = hs.nb[5], cb_tp = hs.nb[6], cb_tm = hs.nb[7];
    const bool tw_p = hs.tw_p, tw_m = hs.tw_m;
    const T bc = T(g.bc);

    Cx<T> R[12];
#pragma unroll
    for (int k = 0; k < 12; k++) R[k] = {T(0), T(0)};

    constexpr int FW = DAG ? 1 : 0;   // projector of the forward hop
    constexpr int BW = DAG ? 0 : 1;   // projector of the backward hop
    // Software pipeline over the eight hops.  DEPTH 2: the operands of hop
    // d+1 are in flight while hop d is computed; DEPTH 1: one hop at a time
    // (fewer registers, used where occupancy matters more).  pin_acc()
    // orders the next gathers after the arithmetic of the current hop.
    const SIn &in = a.in;
    const G &U = a.U;
    if constexpr (DEPTH == 2) {
        auto o0 = fetch_dir<B, 0, FW>(in, in_q + cb_xp * nr + rr, U, poff + cb);
        auto o1 = fetch_dir<B, 0, BW>(in, in_q + cb_xm * nr + rr, U, qoff + cb_xm);
        apply_dir<B, 0, FW, false>(R, in, U, o0, T(1));
        QKX_SCHED_FENCE();
        auto o2 = fetch_dir<B, 1, FW>(in, in_q + cb_yp * nr + rr, U, poff + cb);
        apply_dir<B, 0, BW, true>(R, in, U, o1, T(1));
        QKX_SCHED_FENCE();
        auto o3 = fetch_dir<B, 1, BW>(in, in_q + cb_ym * nr + rr, U, qoff + cb_ym);
        apply_dir<B, 1, FW, false>(R, in, U, o2, T(1));
        QKX_SCHED_FENCE();
        auto o4 = fetch_dir<B, 2, FW>(in, in_q + cb_zp * nr + rr, U, poff + cb);
        apply_dir<B, 1, BW, true>(R, in, U, o3, T(1));
        QKX_SCHED_FENCE();
        auto o5 = fetch_dir<B, 2, BW>(in, in_q + cb_zm * nr + rr, U, qoff + cb_zm);
        apply_dir<B, 2, FW, false>(R, in, U, o4, T(1));
        QKX_SCHED_FENCE();
        auto o6 = fetch_dir<B, 3, FW>(in, in_q + cb_tp * nr + rr, U, poff + cb);
        apply_dir<B, 2, BW, true>(R, in, U, o5, T(1));
        QKX_SCHED_FENCE();
        auto o7 = fetch_dir<B, 3, BW>(in, in_q + cb_tm * nr + rr, U, qoff + cb_tm);
        apply_dir<B, 3, FW, false>(R, in, U, o6, tw_p ? bc : T(1));
        QKX_SCHED_FENCE();
        apply_dir<B, 3, BW, true>(R, in, U, o7, tw_m ? bc : T(1));
    } else {
        apply_dir<B, 0, FW, false>(R, in, U, fetch_dir<B, 0, FW>(in, in_q + cb_xp * nr + rr, U, poff + cb), T(1));
        QKX_SCHED_FENCE();
        apply_dir<B, 0, BW, true>(R, in, U, fetch_dir<B, 0, BW>(in, in_q + cb_xm * nr + rr, U, qoff + cb_xm), T(1));
        QKX_SCHED_FENCE();
        apply_dir<B, 1, FW, false>(R, in, U, fetch_dir<B, 1, FW>(in, in_q + cb_yp * nr + rr, U, poff + cb), T(1));
        QKX_SCHED_FENCE();
        apply_dir<B, 1, BW, true>(R, in, U, fetch_dir<B, 1, BW>(in, in_q + cb_ym * nr + rr, U, qoff + cb_ym), T(1));
        QKX_SCHED_FENCE();
        apply_dir<B, 2, FW, false>(R, in, U, fetch_dir<B, 2, FW>(in, in_q + cb_zp * nr + rr, U, poff + cb), T(1));
        QKX_SCHED_FENCE();
        apply_dir<B, 2, BW, true>(R, in, U, fetch_dir<B, 2, BW>(in, in_q + cb_zm * nr + rr, U, qoff + cb_zm), T(1));
        QKX_SCHED_FENCE();
        apply_dir<B, 3, FW, false>(R, in, U, fetch_dir<B, 3, FW>(in, in_q + cb_tp * nr + rr, U, poff + cb),
                                   tw_p ? bc : T(1));
        QKX_SCHED_FENCE();
        apply_dir<B, 3, BW, true>(R, in, U, fetch_dir<B, 3, BW>(in, in_q + cb_tm * nr + rr, U, qoff + cb_tm),
                                  tw_m ? bc : T(1));
    }

    hop_finish<B, AUX>(a, fin, p, cb, rr, ep, R, nrm2);
    }
    hop_partials(a, fin, bid, u, nrm2);
}

template <class SIn, class SOut, class SAux, class G, int B, bool DAG, bool AUX, int DEPTH, class Fin = StoreFin>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(HopWaves<SIn>::v, 8))) void hop_kernel(
    HopArgs<SIn, SOut, SAux, G> a, Fin fin)
{
    if (a.skip && *a.skip) return;   // uniform over the grid
    fin.begin(a);
    int p, u;   // parity and unit (site x rhs) index within it
    hop_unit(a.g, a.par_out, a.nblk8, a.nrhs, p, u, blockIdx.x);
    Epi ep = a.epi[0];                 // uniform (scalar) load for a single rhs
    if (a.nrhs > 1) ep = a.epi[u - fdiv(u, a.dnr) * a.nrhs];
    hop_body<B, DAG, AUX, DEPTH>(a, fin, blockIdx.x, p, u, ep);
}

// Single-RHS FP16 half-lattice hop with the x and y neighbours staged in
// LDS.  A staging domain -- the workgroup (WAVE false) or each wave (WAVE
// true: no workgroup barrier) -- computes the consecutive checkerboard
// indices [db, db + ds); their +-x neighbours lie in their own x rows and
// their +-y neighbours one row away, so all of them, bar the +-y neighbours
// across the periodic y boundary, are input sites of [db - Lh0, db + ds +
// Lh0).  The domain reads that range once, coalesced, into LDS; the z and t
// neighbours, every link and the y-boundary neighbours are gathered from
// global memory as in hop_kernel.  The arithmetic and its order are
// hop_kernel's: the same bits.  The host launches it with Lh0 <= kStageRow,
// blockDim.x <= 256 a multiple of 64, 2 Lh0 <= blockDim.x for workgroup
// domains, and stage_lds_bytes() of dynamic LDS.
constexpr int kStageRow = 64;
__host__ __device__ constexpr int stage_sites(bool wave, int block, int lh0)
{
    return wave ? (block / 64) * (64 + 2 * lh0) : block + 2 * lh0;
}
inline size_t stage_lds_bytes(bool wave, int block, int lh0)
{
    return (size_t)stage_sites(wave, block, lh0) * (3 * sizeof(uint4) + sizeof(float));
}
// minimum waves per SIMD requested from the register allocator
#ifndef QKX_STAGE_WAVES
#define QKX_STAGE_WAVES 4
#endif

template <int B, bool DAG, bool AUX, bool WAVE, class Fin = StoreFin>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(QKX_STAGE_WAVES, 8))) void hop_stage_kernel(
    HopArgs<SpinorH, SpinorH, SpinorH, Gauge16<false>> a, Fin fin)
{
    using T = float;
    using Raw = SpinorH::Raw;
    using LRaw = Gauge16<false>::Raw;
    using Ops = HopOps<SpinorH, Gauge16<false>>;
    extern __shared__ uint4 stage_lds[];
    if (a.skip && *a.skip) return;   // uniform over the grid
    fin.begin(a);
    const Geom &g = a.g;
    int p, u;
    hop_unit(g, a.par_out, a.nblk8, 1, p, u, blockIdx.x);
    const int q = p ^ 1;
    const int qoff = q * g.Vh, poff = p * g.Vh;
    const int in_q = a.in_off[q];
    // this thread's staging domain and its LDS region: sc[3][cap], sn[cap]
    const int cap = stage_sites(WAVE, blockDim.x, g.Lh0);
    const int ds = WAVE ? 64 : (int)blockDim.x;
    const int li = WAVE ? (int)(threadIdx.x & 63) : (int)threadIdx.x;   // index within the domain
    const int reg = WAVE ? (int)(threadIdx.x >> 6) * (64 + 2 * g.Lh0) : 0;
    uint4 *sc = stage_lds + reg;
    float *sn = reinterpret_cast<float *>(stage_lds + 3 * cap) + reg;
    const int db = u - li;
    const int lo = max(db - g.Lh0, 0);
    const int n = min(db + ds + g.Lh0, g.Vh) - lo;   // <= 0 past the last site
    const SpinorH &in = a.in;
    const Gauge16<false> &U = a.U;
    constexpr int FW = DAG ? 1 : 0;
    constexpr int BW = DAG ? 0 : 1;
    const bool live = u < g.Vh;
    HopSite hs{};
    if (live) hs = hop_site(g, p, u);
    // the x links go out first and stay in flight across the staging
    LRaw lxp, lxm;
    if (live) {
        lxp = U.fetch(0, poff + u);
        lxm = U.fetch(0, qoff + hs.nb[1]);
    }
    {
        // n <= ds + 2 Lh0: at most 2 rows of ds for a workgroup, 3 for a wave
        constexpr int K = WAVE ? 3 : 2;
        Raw sr[K];
#pragma unroll
        for (int k = 0; k < K; k++)
            if (li + k * ds < n) sr[k] = in.fetch<3>(in_q + lo + li + k * ds);
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int i = li + k * ds;
            if (i < n) {
                sc[i] = sr[k].c[0];
                sc[cap + i] = sr[k].c[1];
                sc[2 * cap + i] = sr[k].c[2];
                sn[i] = sr[k].n;
            }
        }
    }
    // a wave's LDS accesses complete in program order
    if constexpr (WAVE) __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    else __syncthreads();
    double nrm2 = 0.0;
    if (live) {
        auto staged = [&](int nb) {
            const int i = nb - lo;
            Raw r;
            r.c[0] = sc[i];
            r.c[1] = sc[cap + i];
            r.c[2] = sc[2 * cap + i];
            r.n = sn[i];
            return r;
        };
        // a +-y neighbour across the y boundary can lie outside the staged
        // range: gathered one hop ahead, like the z and t operands
        auto y_ops = [&](int nb, int s_link, auto mask) {
            Ops o;
            o.link = U.fetch(1, s_link);
            if ((unsigned)(nb - lo) >= (unsigned)n) o.psi = in.fetch<decltype(mask)::value>(in_q + nb);
            return o;
        };
        auto y_psi = [&](Ops &o, int nb) {
            if ((unsigned)(nb - lo) < (unsigned)n) o.psi = staged(nb);
        };
        const T bc = T(g.bc);
        Cx<T> R[12];
#pragma unroll
        for (int k = 0; k < 12; k++) R[k] = {T(0), T(0)};
        Ops oyp = y_ops(hs.nb[2], poff + u, std::integral_constant<int, spin_mask<B, 1, FW>()>{});
        apply_dir<B, 0, FW, false>(R, in, U, Ops{staged(hs.nb[0]), lxp}, T(1));
        QKX_SCHED_FENCE();
        Ops oym = y_ops(hs.nb[3], qoff + hs.nb[3], std::integral_constant<int, spin_mask<B, 1, BW>()>{});
        apply_dir<B, 0, BW, true>(R, in, U, Ops{staged(hs.nb[1]), lxm}, T(1));
        QKX_SCHED_FENCE();
        const Ops ozp = fetch_dir<B, 2, FW>(in, in_q + hs.nb[4], U, poff + u);
        y_psi(oyp, hs.nb[2]);
        apply_dir<B, 1, FW, false>(R, in, U, oyp, T(1));
        QKX_SCHED_FENCE();
        const Ops ozm = fetch_dir<B, 2, BW>(in, in_q + hs.nb[5], U, qoff + hs.nb[5]);
        y_psi(oym, hs.nb[3]);
        apply_dir<B, 1, BW, true>(R, in, U, oym, T(1));
        QKX_SCHED_FENCE();
        const Ops otp = fetch_dir<B, 3, FW>(in, in_q + hs.nb[6], U, poff + u);
        apply_dir<B, 2, FW, false>(R, in, U, ozp, T(1));
        QKX_SCHED_FENCE();
        const Ops otm = fetch_dir<B, 3, BW>(in, in_q + hs.nb[7], U, qoff + hs.nb[7]);
        apply_dir<B, 2, BW, true>(R, in, U, ozm, T(1));
        QKX_SCHED_FENCE();
        apply_dir<B, 3, FW, false>(R, in, U, otp, hs.tw_p ? bc : T(1));
        QKX_SCHED_FENCE();
        apply_dir<B, 3, BW, true>(R, in, U, otm, hs.tw_m ? bc : T(1));
        hop_finish<B, AUX>(a, fin, p, u, 0, a.epi[0], R, nrm2);
    }
    hop_partials(a, fin, blockIdx.x, u, nrm2);
}

// Several independent hops -- each on its own lattice, gauge field and
// fields, one right-hand side each -- in one launch: workgroups
// [blk0[i], blk0[i + 1]) do instance i (each instance's range is a whole
// number of 8-workgroup rounds, so the XCD remapping of an instance works as
// in its own launch).  Used by the multi-instance solver.
constexpr int kMaxInst = 8;
template <class S, class G> struct MultiHop {
    int n;
    int blk0[kMaxInst + 1];
    HopArgs<S, S, S, G, 1> inst[kMaxInst];
};

template <class S, class G, int B, bool DAG, bool AUX, int DEPTH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(HopWaves<S>::v, 8))) void hop_multi_kernel(
    MultiHop<S, G> m)
{
    int i = 0;
    while (i + 1 < m.n && (int)blockIdx.x >= m.blk0[i + 1]) i++;
    // The instance's arguments are read in place from the kernel argument
    // segment, where the launch put the bytes of m (the only argument, at
    // offset 0): indexing the by-value parameter with i would make the
    // compiler copy all of it to scratch memory first.
    const auto *mk = (const MultiHop<S, G> *)__builtin_amdgcn_kernarg_segment_ptr();
    const HopArgs<S, S, S, G, 1> &a = mk->inst[i];
    if (a.skip && *a.skip) return;   // uniform over the instance
    StoreFin fin;
    const int bid = (int)blockIdx.x - m.blk0[i];
    int p, u;
    hop_unit(a.g, a.par_out, a.nblk8, 1, p, u, bid);
    hop_body<B, DAG, AUX, DEPTH>(a, fin, bid, p, u, a.epi[0]);
}

// Batches of FP16 spinors on FP16 12-real links, 8 to 16 right-hand sides:
// a workgroup holds whole sites (nr * (sites per workgroup) threads, the RHS
// of a site on consecutive threads).  The 8 links of each of its sites are
// fetched and reconstructed once per workgroup, by one thread per (site,
// direction), into LDS; every (site, RHS) thread then reads them from there.
// In hop_kernel each thread fetches and reconstructs its site's links itself,
// and a wave repeats that work for the ~5 sites it covers 12 times over.
// Sites per workgroup at most: 32 for 8-16 RHS (12 KB of LDS), 85 for 3-7 RHS
// (29 KB: still 5 workgroups of 4 waves per CU, the register-limited
// occupancy).  Two RHS would need 128 sites and 43 KB, which costs occupancy.
constexpr int kShareSites = 32;
constexpr int kShareSitesFew = 85;
constexpr int kShareSitesF32 = 64;   // FP32 links in LDS: 64 sites = 41 KB (3-7 RHS)
constexpr int kShareMinRhs = 3;
template <class F> __device__ __forceinline__ double *fin_r2_part(const F &f)
{
    if constexpr (IsBCg<F>::value) return f.r2_part;
    else return nullptr;
}

template <int B, bool DAG, bool AUX, class Fin = StoreFin, int SITES = kShareSites>
__global__ __launch_bounds__(256) void hop_share_kernel(HopArgs<SpinorH, SpinorH, SpinorH, Gauge16<false>> a, Fin fin)
{
    using T = float;
    constexpr bool CG = IsBCg<Fin>::value;
    if (a.skip && *a.skip) return;   // uniform over the grid
    __shared__ int lsite[SITES][8];          // link site of each (site, direction); -1: no site
    __shared__ uint32_t lrow[SITES][8][9];   // the link's 9 (re, im) half pairs
    __shared__ double alpha_s[QKX_MAX_RHS];        // CG step of each right-hand side
    if constexpr (CG) {
        const int t = threadIdx.x;
        if (t < a.nrhs)
            alpha_s[t] = (!fin.active[t] || fin.r2cur[t] <= fin.stop2[t]) ? 0.0 : fin.r2cur[t] / fin.pap[t];
    }
    const Geom &g = a.g;
    const int nr = a.nrhs;
    int p, u;
    hop_unit(g, a.par_out, a.nblk8, nr, p, u, blockIdx.x);
    const int v = fdiv(u, a.dnr);
    const int rr = u - v * nr;
    const int j = fdiv((int)threadIdx.x, a.dnr);   // site within the workgroup
    int cb = v;
    HopSite hs{};
    const bool valid = v < g.Vh;
    if (valid) hs = hop_site_walk(g, a.tr, p, v, cb);
    const int q = p ^ 1;
    const int qoff = q * g.Vh, poff = p * g.Vh;
    if (rr == 0) {
#pragma unroll
        for (int d = 0; d < 8; d++)
            lsite[j][d] = !valid ? -1 : (d & 1) ? qoff + hs.nb[d] : poff + cb;
    }
    __syncthreads();
    // (site, direction) tasks, dense over the workgroup's threads
    const int nsites = fdiv((int)blockDim.x, a.dnr);
    for (int t = threadIdx.x; t < nsites * 8; t += blockDim.x) {
        const int js = t >> 3, d = t & 7;
        const int sl = lsite[js][d];
        if (sl >= 0) {
            h2v Ul[9];
            link_rows16(a.U.fetch(d >> 1, sl), Ul);
#pragma unroll
            for (int k = 0; k < 9; k++) lrow[js][d][k] = __builtin_bit_cast(uint32_t, Ul[k]);
        }
    }
    __syncthreads();
    double nrm2 = 0.0;
    if (valid) {
        const int in_q = a.in_off[q];
        const T bc = T(g.bc);
        Cx<T> R[12];
#pragma unroll
        for (int k = 0; k < 12; k++) R[k] = {T(0), T(0)};
        constexpr int FW = DAG ? 1 : 0;
        constexpr int BW = DAG ? 0 : 1;
        const SpinorH &in = a.in;
        auto rows = [&](int d, h2v U[9]) {
#pragma unroll
            for (int k = 0; k < 9; k++) U[k] = h2_of(lrow[j][d][k]);
        };
        // neighbour spinors two hops ahead of the arithmetic
        auto f0 = in.template fetch<spin_mask<B, 0, FW>()>(in_q + hs.nb[0] * nr + rr);
        auto f1 = in.template fetch<spin_mask<B, 0, BW>()>(in_q + hs.nb[1] * nr + rr);
        h2v U[9];
        rows(0, U);
        apply_rows_h16<B, 0, FW, false>(R, f0, U, T(1));
        QKX_SCHED_FENCE();
        auto f2 = in.template fetch<spin_mask<B, 1, FW>()>(in_q + hs.nb[2] * nr + rr);
        rows(1, U);
        apply_rows_h16<B, 0, BW, true>(R, f1, U, T(1));
        QKX_SCHED_FENCE();
        auto f3 = in.template fetch<spin_mask<B, 1, BW>()>(in_q + hs.nb[3] * nr + rr);
        rows(2, U);
        apply_rows_h16<B, 1, FW, false>(R, f2, U, T(1));
        QKX_SCHED_FENCE();
        auto f4 = in.template fetch<spin_mask<B, 2, FW>()>(in_q + hs.nb[4] * nr + rr);
        rows(3, U);
        apply_rows_h16<B, 1, BW, true>(R, f3, U, T(1));
        QKX_SCHED_FENCE();
        auto f5 = in.template fetch<spin_mask<B, 2, BW>()>(in_q + hs.nb[5] * nr + rr);
        rows(4, U);
        apply_rows_h16<B, 2, FW, false>(R, f4, U, T(1));
        QKX_SCHED_FENCE();
        auto f6 = in.template fetch<spin_mask<B, 3, FW>()>(in_q + hs.nb[6] * nr + rr);
        rows(5, U);
        apply_rows_h16<B, 2, BW, true>(R, f5, U, T(1));
        QKX_SCHED_FENCE();
        auto f7 = in.template fetch<spin_mask<B, 3, BW>()>(in_q + hs.nb[7] * nr + rr);
        rows(6, U);
        apply_rows_h16<B, 3, FW, false>(R, f6, U, hs.tw_p ? bc : T(1));
        QKX_SCHED_FENCE();
        rows(7, U);
        apply_rows_h16<B, 3, BW, true>(R, f7, U, hs.tw_m ? bc : T(1));

        Cx<T> res[12];
        const Epi ep = a.epi[rr];
        if constexpr (AUX) {
            Cx<T> xv[12];
            a.aux.load(a.aux_off[p] + cb * nr + rr, xv);
            epilogue<B>(ep, R, xv, res);
        } else {
            epilogue<B, T>(ep, R, nullptr, res);
        }
        if constexpr (CG) {
            // x += alpha p (unless left to the direction kernel), r -= alpha Ap
            // for this (site, RHS)
            const int so = a.out_off[p] + cb * nr + rr;
            using TX = typename decltype(fin.x)::Real;
            using TR = typename decltype(fin.r)::Real;
            if constexpr (Fin::kUpdX) {
                Cx<TX> xv[12];
                Cx<typename decltype(fin.p)::Real> pv[12];
                fin.x.load(so, xv);
                fin.p.load(so, pv);
                const TX ax = TX(alpha_s[rr]);
#pragma unroll
                for (int k = 0; k < 12; k++) xv[k] = {xv[k].re + ax * TX(pv[k].re), xv[k].im + ax * TX(pv[k].im)};
                fin.x.store(so, xv);
            }
            Cx<TR> rv[12];
            fin.r.load(so, rv);
            const TR ar = TR(alpha_s[rr]);
#pragma unroll
            for (int k = 0; k < 12; k++) {
                rv[k] = {rv[k].re - ar * TR(res[k].re), rv[k].im - ar * TR(res[k].im)};
                nrm2 += (double)rv[k].re * (double)rv[k].re + (double)rv[k].im * (double)rv[k].im;
            }
            fin.r.store(so, rv);
        } else {
            a.out.store(a.out_off[p] + cb * nr + rr, res);
            if (a.norm_partial) {
#pragma unroll
                for (int k = 0; k < 12; k++)
                    nrm2 += (double)res[k].re * (double)res[k].re + (double)res[k].im * (double)res[k].im;
            }
        }
    }
    if (CG || a.norm_partial) {
        // one partial per right-hand side, summed in thread order
        const int ub = u - (int)threadIdx.x;   // the workgroup's first unit, >= 0
        const int u0 = ub - fdiv(ub, a.dnr) * nr;
        double *dst = CG ? fin_r2_part(fin) : a.norm_partial;
        block_partial_rhs(nrm2, nr, u0, dst + blockIdx.x * nr);
    }
}

// Batches of FP32 or BF16 spinors on 12-real links (FP32 or BF16), 3-16
// right-hand sides: as hop_share_kernel, a workgroup holds whole sites, each
// site's 8 links are fetched and their third rows rebuilt once per
// workgroup, here into LDS as 9 FP32 complex numbers (72 B per link, the
// boundary sign of the wrapping t hops already applied), and every (site,
// RHS) thread multiplies its projected neighbour by them.  The per-thread
// arithmetic is that of hop_kernel; the compiler may contract its products
// and sums into FMAs differently in the two kernels, so results agree to
// the last FP32 bit (one rounding of the storage format).
template <class S, class G, int B, bool DAG, bool AUX, int SITES>
__global__ __launch_bounds__(256) void hop_share_f_kernel(HopArgs<S, S, S, G> a)
{
    using T = float;
    if (a.skip && *a.skip) return;   // uniform over the grid
    __shared__ int lsite[SITES][8];               // link site of each (site, direction); -1: no site
    __shared__ float lsign[SITES][8];             // boundary sign of the hop
    __shared__ Cx<float> lU[SITES][8][9];         // the rebuilt links
    const Geom &g = a.g;
    const int nr = a.nrhs;
    int p, u;
    hop_unit(g, a.par_out, a.nblk8, nr, p, u, blockIdx.x);
    const int v = fdiv(u, a.dnr);
    const int rr = u - v * nr;
    const int j = fdiv((int)threadIdx.x, a.dnr);
    int cb = v;
    HopSite hs{};
    const bool valid = v < g.Vh;
    if (valid) hs = hop_site_walk(g, a.tr, p, v, cb);
    const int q = p ^ 1;
    const int qoff = q * g.Vh, poff = p * g.Vh;
    const int in_q = a.in_off[q];
    const S &in = a.in;
    constexpr int FW = DAG ? 1 : 0;
    constexpr int BW = DAG ? 0 : 1;
    // neighbour spinors one hop ahead of the arithmetic; the first two are
    // requested before the link phase (their latency overlaps it)
    typename S::Raw f0{}, f1{};
    if (valid) {
        f0 = in.template fetch<spin_mask<B, 0, FW>()>(in_q + hs.nb[0] * nr + rr);
        f1 = in.template fetch<spin_mask<B, 0, BW>()>(in_q + hs.nb[1] * nr + rr);
    }
    if (rr == 0) {
#pragma unroll
        for (int d = 0; d < 8; d++) {
            lsite[j][d] = !valid ? -1 : (d & 1) ? qoff + hs.nb[d] : poff + cb;
            lsign[j][d] = (d == 6 && hs.tw_p) || (d == 7 && hs.tw_m) ? (float)g.bc : 1.0f;
        }
    }
    __syncthreads();
    const int nsites = fdiv((int)blockDim.x, a.dnr);
    for (int t = threadIdx.x; t < nsites * 8; t += blockDim.x) {
        const int js = t >> 3, d = t & 7;
        const int sl = lsite[js][d];
        if (sl >= 0) {
            Cx<float> Ul[9];
            a.U.unpack(a.U.fetch(d >> 1, sl), Ul);
            const float sg = lsign[js][d];
#pragma unroll
            for (int k = 0; k < 9; k++) lU[js][d][k] = {sg * Ul[k].re, sg * Ul[k].im};
        }
    }
    __syncthreads();
    double nrm2 = 0.0;
    if (valid) {
        Cx<T> R[12];
#pragma unroll
        for (int k = 0; k < 12; k++) R[k] = {T(0), T(0)};
        // one hop with the link from LDS: project, multiply, reconstruct
        auto hop = [&](auto MUC, auto KINDC, auto DAGLC, const typename S::Raw &raw, int d) {
            constexpr int MU = decltype(MUC)::value, KIND = decltype(KINDC)::value;
            constexpr bool DAGL = decltype(DAGLC)::value;
            Cx<T> psi[12];
            in.template unpack<spin_mask<B, MU, KIND>()>(raw, psi);
            Cx<T> h[6];
            project_k<B, MU, KIND, 0>(psi, h);
            project_k<B, MU, KIND, 1>(psi, h);
            Cx<T> Ul[9];
#pragma unroll
            for (int k = 0; k < 9; k++) Ul[k] = lU[j][d][k];
            Cx<T> xi[6];
            link_mul<DAGL, kFmaChain<S>>(Ul, h, xi);
            collect_s<B, MU, KIND, 0>(R, xi);
            collect_s<B, MU, KIND, 1>(R, xi);
            collect_s<B, MU, KIND, 2>(R, xi);
            collect_s<B, MU, KIND, 3>(R, xi);
        };
        using I0 = std::integral_constant<int, 0>;
        using I1 = std::integral_constant<int, 1>;
        using I2 = std::integral_constant<int, 2>;
        using I3 = std::integral_constant<int, 3>;
        using IF = std::integral_constant<int, FW>;
        using IB = std::integral_constant<int, BW>;
        using BT = std::true_type;
        using BF = std::false_type;
        hop(I0{}, IF{}, BF{}, f0, 0);
        QKX_SCHED_FENCE();
        auto f2 = in.template fetch<spin_mask<B, 1, FW>()>(in_q + hs.nb[2] * nr + rr);
        hop(I0{}, IB{}, BT{}, f1, 1);
        QKX_SCHED_FENCE();
        auto f3 = in.template fetch<spin_mask<B, 1, BW>()>(in_q + hs.nb[3] * nr + rr);
        hop(I1{}, IF{}, BF{}, f2, 2);
        QKX_SCHED_FENCE();
        auto f4 = in.template fetch<spin_mask<B, 2, FW>()>(in_q + hs.nb[4] * nr + rr);
        hop(I1{}, IB{}, BT{}, f3, 3);
        QKX_SCHED_FENCE();
        auto f5 = in.template fetch<spin_mask<B, 2, BW>()>(in_q + hs.nb[5] * nr + rr);
        hop(I2{}, IF{}, BF{}, f4, 4);
        QKX_SCHED_FENCE();
        auto f6 = in.template fetch<spin_mask<B, 3, FW>()>(in_q + hs.nb[6] * nr + rr);
        hop(I2{}, IB{}, BT{}, f5, 5);
        QKX_SCHED_FENCE();
        auto f7 = in.template fetch<spin_mask<B, 3, BW>()>(in_q + hs.nb[7] * nr + rr);
        hop(I3{}, IF{}, BF{}, f6, 6);
        QKX_SCHED_FENCE();
        hop(I3{}, IB{}, BT{}, f7, 7);

        Cx<T> res[12];
        const Epi ep = a.epi[rr];
        if constexpr (AUX) {
            Cx<T> xv[12];
            a.aux.load(a.aux_off[p] + cb * nr + rr, xv);
            epilogue<B>(ep, R, xv, res);
        } else {
            epilogue<B, T>(ep, R, nullptr, res);
        }
        a.out.store(a.out_off[p] + cb * nr + rr, res);
        if (a.norm_partial) {
#pragma unroll
            for (int k = 0; k < 12; k++)
                nrm2 += (double)res[k].re * (double)res[k].re + (double)res[k].im * (double)res[k].im;
        }
    }
    if (a.norm_partial) {
        const int ub = u - (int)threadIdx.x;   // the workgroup's first unit, >= 0
        const int u0 = ub - fdiv(ub, a.dnr) * nr;
        block_partial_rhs(nrm2, nr, u0, a.norm_partial + blockIdx.x * nr);
    }
}

// One hop for RT right-hand sides of one site: the link is fetched and
// unpacked once, then applied to the RT neighbour spinors (consecutive in
// the batch layout).
template <int B, int MU, int KIND, bool DAGL, int RT, class SIn, class G, class T>
__device__ __forceinline__ void tile_dir(Cx<T> (&R)[RT][12], const SIn &in, const G &U, int s_nb, int s_link,
                                         T sign)
{
    const typename G::Raw lr = U.fetch(MU, s_link);
    typename SIn::Raw pr[RT];
#pragma unroll
    for (int j = 0; j < RT; j++) pr[j] = in.template fetch<spin_mask<B, MU, KIND>()>(s_nb + j);
    Cx<T> Ul[9];
    U.unpack(lr, Ul);
#pragma unroll
    for (int j = 0; j < RT; j++) {
        Cx<T> psi[12];
        in.template unpack<spin_mask<B, MU, KIND>()>(pr[j], psi);
        Cx<T> h[6];
        project_k<B, MU, KIND, 0>(psi, h);
        project_k<B, MU, KIND, 1>(psi, h);
        if (sign != T(1)) {
#pragma unroll
            for (int k = 0; k < 6; k++) h[k] = sign * h[k];
        }
        Cx<T> xi[6];
        link_mul<DAGL, kFmaChain<SIn>>(Ul, h, xi);
        collect_s<B, MU, KIND, 0>(R[j], xi);
        collect_s<B, MU, KIND, 1>(R[j], xi);
        collect_s<B, MU, KIND, 2>(R[j], xi);
        collect_s<B, MU, KIND, 3>(R[j], xi);
    }
}

template <int RT, class T> __device__ __forceinline__ void pin_tile(Cx<T> (&R)[RT][12])
{
#pragma unroll
    for (int j = 0; j < RT; j++) pin_acc(R[j]);
}

// Hop kernel for batches: one thread per (site, RT consecutive right-hand
// sides), so each link is fetched once per RT spinors instead of once per
// spinor.  Same epilogue, boundary and mapping rules as hop_kernel.
template <class SIn, class SOut, class SAux, class G, int B, bool DAG, bool AUX, int RT>
__global__ __launch_bounds__(256) void hop_tile_kernel(HopArgs<SIn, SOut, SAux, G> a)
{
    using T = typename SOut::Real;
    if (a.skip && *a.skip) return;   // uniform over the grid
    const Geom &g = a.g;
    const int nr = a.nrhs, ng = nr / RT;
    int p, u;
    hop_unit(g, a.par_out, a.nblk8, ng, p, u, blockIdx.x);
    const int cb = u / ng;
    const int r0 = (u - cb * ng) * RT;
    double nrm2[RT];
#pragma unroll
    for (int j = 0; j < RT; j++) nrm2[j] = 0.0;
    if (cb < g.Vh) {
        const HopSite hs = hop_site(g, p, cb);
        const int poff = p * g.Vh, qoff = (p ^ 1) * g.Vh;
        const int in_q = a.in_off[p ^ 1];
        const T bc = T(g.bc);
        Cx<T> R[RT][12];
#pragma unroll
        for (int j = 0; j < RT; j++)
#pragma unroll
            for (int k = 0; k < 12; k++) R[j][k] = {T(0), T(0)};
        constexpr int FW = DAG ? 1 : 0;
        constexpr int BW = DAG ? 0 : 1;
        const SIn &in = a.in;
        const G &U = a.U;
        tile_dir<B, 0, FW, false, RT>(R, in, U, in_q + hs.nb[0] * nr + r0, poff + cb, T(1));
        pin_tile<RT>(R);
        tile_dir<B, 0, BW, true, RT>(R, in, U, in_q + hs.nb[1] * nr + r0, qoff + hs.nb[1], T(1));
        pin_tile<RT>(R);
        tile_dir<B, 1, FW, false, RT>(R, in, U, in_q + hs.nb[2] * nr + r0, poff + cb, T(1));
        pin_tile<RT>(R);
        tile_dir<B, 1, BW, true, RT>(R, in, U, in_q + hs.nb[3] * nr + r0, qoff + hs.nb[3], T(1));
        pin_tile<RT>(R);
        tile_dir<B, 2, FW, false, RT>(R, in, U, in_q + hs.nb[4] * nr + r0, poff + cb, T(1));
        pin_tile<RT>(R);
        tile_dir<B, 2, BW, true, RT>(R, in, U, in_q + hs.nb[5] * nr + r0, qoff + hs.nb[5], T(1));
        pin_tile<RT>(R);
        tile_dir<B, 3, FW, false, RT>(R, in, U, in_q + hs.nb[6] * nr + r0, poff + cb, hs.tw_p ? bc : T(1));
        pin_tile<RT>(R);
        tile_dir<B, 3, BW, true, RT>(R, in, U, in_q + hs.nb[7] * nr + r0, qoff + hs.nb[7], hs.tw_m ? bc : T(1));
#pragma unroll
        for (int j = 0; j < RT; j++) {
            const Epi ep = a.epi[r0 + j];
            Cx<T> res[12];
            if constexpr (AUX) {
                Cx<T> xv[12];
                a.aux.load(a.aux_off[p] + cb * nr + r0 + j, xv);
                epilogue<B>(ep, R[j], xv, res);
            } else {
                epilogue<B, T>(ep, R[j], nullptr, res);
            }
            a.out.store(a.out_off[p] + cb * nr + r0 + j, res);
            if (a.norm_partial) {
#pragma unroll
                for (int k = 0; k < 12; k++)
                    nrm2[j] += (double)res[k].re * (double)res[k].re + (double)res[k].im * (double)res[k].im;
            }
        }
    }
    if (a.norm_partial) {
        // one partial per right-hand side, summed in thread order: thread t
        // holds right-hand sides ((u0 + t) mod ng) RT + j
        __shared__ double vals[256 * RT];
#pragma unroll
        for (int j = 0; j < RT; j++) vals[threadIdx.x * RT + j] = nrm2[j];
        __syncthreads();
        if ((int)threadIdx.x < nr) {
            const int gi = threadIdx.x / RT, j = threadIdx.x - gi * RT;
            const int u0 = (((u - (int)threadIdx.x) % ng) + ng) % ng;
            double sum = 0;
            for (int t = (gi - u0 + ng) % ng; t < (int)blockDim.x; t += ng) sum += vals[t * RT + j];
            a.norm_partial[blockIdx.x * nr + threadIdx.x] = sum;
        }
    }
}

// Site-local combination: out = cR y + gR i g5 y + cX x + gX i g5 x (g5out),
// optionally restricted to one parity with zeros written on the other.
struct EpiSet {
    Epi e[QKX_MAX_RHS];
};

template <class S, int B>
__global__ __launch_bounds__(256) void site_kernel(S out, S y, S x, int V, int Vh, int has_x,
                                                   int keep_par, int nrhs, EpiSet es)
{
    using T = typename S::Real;
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= V) return;
    Epi e = es.e[0];
    if (nrhs > 1) e = es.e[s % nrhs];
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
int g_depth64 = 2;   // pipeline depth of the FP64 hop (2 measured faster than 1 at 24^3x48)
int g_depth32 = 2;   // pipeline depth of the FP32/FP16/BF16 hops
int g_rhs_tile = 0;  // right-hand sides per thread in batched hops (0 or 1: one)
int g_share_links = 1;  // FP16 batches of >= 3 RHS share links through LDS: 0 off, 1 automatic, 2 always
// Automatic sharing needs at least 4 waves per SIMD of (site, RHS) units (MI355X:
// 1024 SIMDs): in smaller launches the extra barrier phase costs more than it saves.
constexpr long kShareMinUnits = 4L * 1024 * 64;
bool share_links_for(int nrhs, long units)
{
    if (!g_share_links || nrhs < kShareMinRhs || g_block / nrhs > kShareSitesFew) return false;
    return g_share_links == 2 || units >= kShareMinUnits;
}
#ifndef QKX_SITE_ORDER_DEFAULT
#define QKX_SITE_ORDER_DEFAULT 2
#endif
// 0: checkerboard index order; 1: blocked walk (HopTrav) for every hop;
// 2: blocked walk for the full-lattice FP64 hops only
int g_site_order = QKX_SITE_ORDER_DEFAULT;
#ifndef QKX_STAGE_DEFAULT
#define QKX_STAGE_DEFAULT 0
#endif
// single-RHS FP16 half-lattice hops stage their x and y neighbours in LDS
// (hop_stage_kernel): 1 on, 0 off
int g_nb_stage = QKX_STAGE_DEFAULT;
const int *g_skip = nullptr;

// Launch shape of one hop: workgroup size, links shared through LDS by the
// (site, RHS) threads of a workgroup, and the software-pipeline depth of the
// per-thread kernel.  Every shape computes each output site with the same
// arithmetic: the same bits for FP16 and FP64, and for FP32/BF16 with shared
// links up to the compiler's FMA contraction (last bit); the per-workgroup
// norm partials group the sites differently.
struct HopShape {
    int block;
    int share;
    int depth;
};
// The autotuner: on the first launch of a kind (spinor and link storage,
// right-hand sides, output sites, full or half lattice, dagger, xpay input,
// basis) it times the candidate shapes on scratch fields of that kind and
// caches the fastest for every later launch of the kind.
struct TuneKey {
    int sp, gp, gr, nrhs, sites, full, dag, aux, basis;
    bool operator<(const TuneKey &o) const
    {
        const int x[9] = {sp, gp, gr, nrhs, sites, full, dag, aux, basis};
        const int y[9] = {o.sp, o.gp, o.gr, o.nrhs, o.sites, o.full, o.dag, o.aux, o.basis};
        for (int i = 0; i < 9; i++)
            if (x[i] != y[i]) return x[i] < y[i];
        return false;
    }
};
struct TuneEntry {
    HopShape shape;
    double us;            // its measured time per launch
    double default_us;    // the time of the untuned default shape
};
std::map<TuneKey, TuneEntry> g_tuned;
int g_autotune = 1;                 // 0 off, 1 batched launches (nrhs > 1), 2 every launch
const HopShape *g_force = nullptr;  // set while the tuner times a candidate
bool g_tuning = false;
thread_local int g_no_tune = 0;     // callers that must not allocate (scratch fields) set this

// Blocked walk for a launch moving site_bytes per output site: rows of y in
// blocks of up to 8, and as many t slices per block (up to 8, dividing L3)
// as keep one z plane of a block -- the traffic between a site's use as a
// +z neighbour and its use as a -z neighbour -- under 1.5 MB, well inside
// the 4 MB L2 of one XCD.
constexpr double kTravWindow = 1.5e6;
constexpr double kTravAutoBytes = 700.0;   // FP64 full-lattice hops move 768-1152 B per site
HopTrav hop_trav(const Geom &g, int par_out, double site_bytes)
{
    HopTrav tr{};
    if (!g_site_order) return tr;
    if (g_site_order == 2 && !(par_out < 0 && site_bytes >= kTravAutoBytes)) return tr;
    int yb = 1;
    for (int d = 2; d <= 8; d++)
        if (g.L[1] % d == 0) yb = d;
    const double row = (par_out < 0 ? g.L[0] : g.Lh0) * site_bytes;
    int tb = 1;
    for (int d = 2; d <= 8; d++)
        if (g.L[3] % d == 0 && row * yb * d <= kTravWindow) tb = d;
    tr.yb = yb;
    tr.tb = tb;
    tr.nyb = g.L[1] / yb;
    tr.dyb = make_fastdiv(yb);
    tr.dtb = make_fastdiv(tb);
    tr.dnyb = make_fastdiv(tr.nyb);
    tr.on = (yb > 1 || tb > 1) ? 1 : 0;
    return tr;
}

template <class S, class G, int B, class Fin = StoreFin>
int launch_hop_t(qkx_spinor_s *out, const qkx_spinor_s *in, const qkx_spinor_s *aux,
                 const qkx_gauge_s *U, int dagger, int par_out, const Epi *e, double *norm_partial,
                 int *nblocks_out, const Fin &fin = Fin{});

// Time the candidate shapes of one launch kind on scratch fields of that
// kind (same precision, parity subset and right-hand-side count; links as
// given) and record the fastest in g_tuned.  Nothing is recorded when the
// stream is being captured into a graph or when instrumentation is on (the
// launch keeps its default shape and a later launch tunes); when the scratch
// fields cannot be had, the default shape is recorded.
template <class S, class G, int B>
void tune_hop(const TuneKey &key, const qkx_spinor_s *out, const qkx_spinor_s *in, const qkx_spinor_s *aux,
              const qkx_gauge_s *U, int dagger, int par_out, const Epi *e, const HopShape &dflt, bool shareable)
{
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(qkx::g_stream, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone) return;
    if (qkx::prof_active()) return;
    std::vector<HopShape> cands;
    const bool sh_any = shareable && g_share_links != 0, sh_only = shareable && g_share_links == 2;
    for (int block : {256, 128}) {
        if (sh_any) cands.push_back({block, 1, dflt.depth});
        if (!sh_only)
            for (int depth : {2, 1}) cands.push_back({block, 0, depth});
    }
    qkx_spinor_s *so = qkx::pool_get_batch(out->L, out->prec, out->subset, out->nrhs);
    qkx_spinor_s *si = qkx::pool_get_batch(in->L, in->prec, in->subset, in->nrhs);
    qkx_spinor_s *sa = aux ? qkx::pool_get_batch(aux->L, aux->prec, aux->subset, aux->nrhs) : nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool ok = so && si && (!aux || sa) && hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess;
    ok = ok && qkx_spinor_zero(si) == QKX_SUCCESS && (!sa || qkx_spinor_zero(sa) == QKX_SUCCESS);
    g_tuning = true;
    double best = 1e300, dflt_us = -1;
    HopShape pick = dflt;
    for (const HopShape &c : cands) {
        if (!ok) break;
        g_force = &c;
        constexpr int kWarm = 2, kReps = 5;
        int rc = 0;
        for (int i = 0; i < kWarm && !rc; i++) rc = launch_hop_t<S, G, B>(so, si, sa, U, dagger, par_out, e, nullptr, nullptr);
        (void)hipEventRecord(e0, qkx::g_stream);
        for (int i = 0; i < kReps && !rc; i++) rc = launch_hop_t<S, G, B>(so, si, sa, U, dagger, par_out, e, nullptr, nullptr);
        (void)hipEventRecord(e1, qkx::g_stream);
        float ms = 0;
        if (rc || hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
            ok = false;
            break;
        }
        const double us = 1e3 * ms / kReps;
        if (c.block == dflt.block && c.share == dflt.share && (c.share || c.depth == dflt.depth)) dflt_us = us;
        if (us < best) {
            best = us;
            pick = c;
        }
    }
    g_force = nullptr;
    g_tuning = false;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    // the scratch fields are freed, not kept in the pool: at the production
    // volume they are gigabytes each
    if (so) qkx::pool_release(so);
    if (si) qkx::pool_release(si);
    if (sa) qkx::pool_release(sa);
    (void)hipGetLastError();
    // a kind that could not be timed (no memory for its scratch fields) keeps
    // the default shape for good rather than retrying at every launch
    if (ok && best < 1e300) g_tuned[key] = TuneEntry{pick, best, dflt_us};
    else g_tuned[key] = TuneEntry{dflt, -1.0, -1.0};
}

template <class S, class G, int B, class Fin>
int launch_hop_t(qkx_spinor_s *out, const qkx_spinor_s *in, const qkx_spinor_s *aux,
                 const qkx_gauge_s *U, int dagger, int par_out, const Epi *e, double *norm_partial,
                 int *nblocks_out, const Fin &fin)
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
    a.nrhs = in->nrhs;
    a.dnr = make_fastdiv(a.nrhs);
    a.nout = (par_out < 0 ? a.g.V : a.g.Vh) * a.nrhs;
    for (int r = 0; r < a.nrhs; r++) a.epi[r] = e[r];
    a.skip = g_skip;
    a.norm_partial = norm_partial;
    // right-hand sides per thread for batches (links fetched once per tile)
    // (automatic: 1 -- at 24^3x48 with 12 right-hand sides, tiles of 2, 3, 4
    // lose more to lower occupancy than they save in link loads)
    int rt = 1;
    if (a.nrhs > 1 && g_rhs_tile > 1 && g_rhs_tile <= 4 && a.nrhs % g_rhs_tile == 0) rt = g_rhs_tile;
    const int upsite = a.nrhs / rt;   // threads per output site
    // FP16 batches of 8 or more RHS on FP16 links: each workgroup fetches and
    // reconstructs its sites' links once into LDS (hop_share_kernel); its
    // size is a whole number of sites
    constexpr bool bcg = IsBCg<Fin>::value;   // batched CG update fused (shared links only)
    constexpr bool store = std::is_same_v<Fin, StoreFin>;
    constexpr bool share16 = std::is_same_v<S, SpinorH> && std::is_same_v<G, Gauge16<false>> &&
                             (store || bcg) && QKX_HOP16_DOT2;
    // FP32 and BF16 batches on 12-real links share FP32 links (hop_share_f_kernel)
    constexpr bool share32 = store && ((std::is_same_v<S, SpinorF> && std::is_same_v<G, GaugeF12>) ||
                                       (std::is_same_v<S, SpinorB> && std::is_same_v<G, Gauge16<true>>));
    constexpr bool share_ok = share16 || share32;
    const int sites_out = par_out < 0 ? a.g.V : a.g.Vh;
    HopShape sh{g_block, share_ok && rt == 1 && share_links_for(a.nrhs, (long)sites_out * a.nrhs) ? 1 : 0,
                sizeof(typename S::Real) == 8 ? g_depth64 : g_depth32};
    if constexpr (store) {
        if (g_force) {
            sh = *g_force;
        } else if (rt == 1 && !g_tuning && !g_no_tune && (g_autotune == 2 || (g_autotune == 1 && a.nrhs > 1))) {
            const TuneKey key{in->prec, U->prec, U->recon, a.nrhs, sites_out, par_out < 0 ? 1 : 0, dagger,
                              aux ? 1 : 0, B};
            auto it = g_tuned.find(key);
            if (it == g_tuned.end()) {
                tune_hop<S, G, B>(key, out, in, aux, U, dagger, par_out, e, sh, share_ok && a.nrhs >= kShareMinRhs);
                it = g_tuned.find(key);
            }
            if (it != g_tuned.end()) {
                const int dshare = sh.share;
                sh = it->second.shape;
                if (g_share_links != 1) sh.share = dshare;   // an explicit setting wins
            }
        }
    }
    const bool few = a.nrhs < 8;   // 3-7 RHS: the larger site capacity
    // sites per workgroup of the shared-link kernels: as many as the
    // workgroup holds, up to the LDS capacity of the kernel instance
    const int share_cap = few ? (share32 ? kShareSitesF32 : kShareSitesFew) : kShareSites;
    const bool share = share_ok && rt == 1 && sh.share && a.nrhs >= kShareMinRhs && a.nrhs <= QKX_MAX_RHS &&
                       (share32 || sh.block / a.nrhs <= share_cap);
    const int bs = share ? a.nrhs * std::min(share_cap, sh.block / a.nrhs) : sh.block;   // threads per workgroup
    int nb = (sites_out * upsite + bs - 1) / bs;
    a.nblk8 = 0;
    if (g_xcd_remap && nb >= 64) {
        if (par_out >= 0) {
            nb = (nb + 7) / 8 * 8;
            a.nblk8 = nb / 8;
        } else {
            int nbh = ((a.g.Vh * upsite + bs - 1) / bs + 7) / 8 * 8;   // blocks per parity
            a.nblk8 = nbh / 8;
            nb = 2 * nbh;
        }
    }
    if (nblocks_out) *nblocks_out = nb;
    // bytes moved per launch (compulsory traffic) for the instrumentation:
    // links once per site, spinors once per (site, rhs)
    const int nout = a.nout;
    double ws = (double)qkx::spinor_elem_bytes(in->prec) * qkx::spinor_nchunks(in->prec) +
                (in->prec == QKX_PREC_HALF ? 4.0 : 0.0);
    double wg = (U->prec == QKX_PREC_DOUBLE ? 16.0 : U->prec == QKX_PREC_SINGLE ? 8.0 : 4.0) *
                (U->recon == QKX_RECON_18 ? 9 : 6);
    double bytes = (double)sites_out * (par_out < 0 ? 4 : 8) * wg + (double)nout * 2 * ws;
    if (aux && aux != in) bytes += (double)nout * ws;
    a.tr = hop_trav(a.g, par_out, bytes / sites_out);
    // a fused CG update also reads and writes r at each output (site, rhs),
    // and reads p, reads and writes x unless x is left to the direction kernel
    if constexpr (!store) {
        bytes += (double)nout * 2 * site_bytes_of<decltype(fin.r)>();
        if constexpr (FinUpdX<Fin>::v)
            bytes += (double)nout * (2 * site_bytes_of<decltype(fin.x)>() + site_bytes_of<decltype(fin.p)>());
    }
    // single-RHS FP16 half-lattice hops in checkerboard order: x and y
    // neighbours staged in LDS when switched on and the x row fits
    constexpr bool stage16 = std::is_same_v<S, SpinorH> && std::is_same_v<G, Gauge16<false>> && QKX_HOP16_DOT2;
    const bool stage_wave = g_nb_stage == 2;
    const bool stage = stage16 && g_nb_stage && !share && rt == 1 && a.nrhs == 1 && par_out >= 0 && !a.tr.on &&
                       a.g.Lh0 <= kStageRow && bs <= 256 && bs % 64 == 0 && (stage_wave || 2 * a.g.Lh0 <= bs);
    const size_t stage_bytes = stage ? stage_lds_bytes(stage_wave, bs, a.g.Lh0) : 0;
    qkx::prof_begin();
    const int depth = sh.depth;
    const dim3 grid(nb), blk(bs);
    hipStream_t st = qkx::g_stream;
#define QKX_LAUNCH_HOP(DG, AX, DP) \
    hipLaunchKernelGGL((hop_kernel<S, S, S, G, B, DG, AX, DP, Fin>), grid, blk, 0, st, a, fin)
#define QKX_LAUNCH_TILE(DG, AX, RT) hipLaunchKernelGGL((hop_tile_kernel<S, S, S, G, B, DG, AX, RT>), grid, blk, 0, st, a)
#define QKX_LAUNCH_TILES(RT)                                                              \
    do {                                                                                  \
        if (dagger) { if (aux) QKX_LAUNCH_TILE(true, true, RT); else QKX_LAUNCH_TILE(true, false, RT); } \
        else        { if (aux) QKX_LAUNCH_TILE(false, true, RT); else QKX_LAUNCH_TILE(false, false, RT); } \
    } while (0)
    // the fused CG hop always has an aux field; RHS tiles exist for stored
    // UKQCD-basis batches only (chiral batches use one RHS per thread)
    constexpr bool tiles = store && B == QKX_BASIS_UKQCD;
    if (!tiles) rt = 1;
    if constexpr (bcg) {
        // the fused batched update exists in the shared-link kernel only, on an aux field
        if (!share || !aux) {
            qkx::set_error("hop: the batched CG update needs shared links (FP16, 3-16 RHS, a large enough launch) and an aux field");
            return QKX_ERR_ARG;
        }
        if (few) {
            if (dagger) hipLaunchKernelGGL((hop_share_kernel<B, true, true, Fin, kShareSitesFew>), grid, blk, 0, st, a, fin);
            else hipLaunchKernelGGL((hop_share_kernel<B, false, true, Fin, kShareSitesFew>), grid, blk, 0, st, a, fin);
        } else {
            if (dagger) hipLaunchKernelGGL((hop_share_kernel<B, true, true, Fin>), grid, blk, 0, st, a, fin);
            else hipLaunchKernelGGL((hop_share_kernel<B, false, true, Fin>), grid, blk, 0, st, a, fin);
        }
    } else if (share && share32) {
        if constexpr (share32) {
#define QKX_LAUNCH_SHARE(DG, AX, NS) hipLaunchKernelGGL((hop_share_f_kernel<S, G, B, DG, AX, NS>), grid, blk, 0, st, a)
            if (few) {
                if (dagger) { if (aux) QKX_LAUNCH_SHARE(true, true, kShareSitesF32); else QKX_LAUNCH_SHARE(true, false, kShareSitesF32); }
                else        { if (aux) QKX_LAUNCH_SHARE(false, true, kShareSitesF32); else QKX_LAUNCH_SHARE(false, false, kShareSitesF32); }
            } else {
                if (dagger) { if (aux) QKX_LAUNCH_SHARE(true, true, kShareSites); else QKX_LAUNCH_SHARE(true, false, kShareSites); }
                else        { if (aux) QKX_LAUNCH_SHARE(false, true, kShareSites); else QKX_LAUNCH_SHARE(false, false, kShareSites); }
            }
#undef QKX_LAUNCH_SHARE
        }
    } else if (share) {
        if constexpr (share16) {
#define QKX_LAUNCH_SHARE(DG, AX, NS) \
    hipLaunchKernelGGL((hop_share_kernel<B, DG, AX, StoreFin, NS>), grid, blk, 0, st, a, fin)
            if (few) {
                if (dagger) { if (aux) QKX_LAUNCH_SHARE(true, true, kShareSitesFew); else QKX_LAUNCH_SHARE(true, false, kShareSitesFew); }
                else        { if (aux) QKX_LAUNCH_SHARE(false, true, kShareSitesFew); else QKX_LAUNCH_SHARE(false, false, kShareSitesFew); }
            } else {
                if (dagger) { if (aux) QKX_LAUNCH_SHARE(true, true, kShareSites); else QKX_LAUNCH_SHARE(true, false, kShareSites); }
                else        { if (aux) QKX_LAUNCH_SHARE(false, true, kShareSites); else QKX_LAUNCH_SHARE(false, false, kShareSites); }
            }
#undef QKX_LAUNCH_SHARE
        }
    } else if (stage) {
        if constexpr (stage16 && !bcg) {
#define QKX_LAUNCH_STAGE(DG, AX, WV) \
    hipLaunchKernelGGL((hop_stage_kernel<B, DG, AX, WV, Fin>), grid, blk, stage_bytes, st, a, fin)
#define QKX_LAUNCH_STAGES(WV)                                                                                 \
    do {                                                                                                      \
        if (dagger) { if (aux) QKX_LAUNCH_STAGE(true, true, WV); else if constexpr (store) QKX_LAUNCH_STAGE(true, false, WV); } \
        else        { if (aux) QKX_LAUNCH_STAGE(false, true, WV); else if constexpr (store) QKX_LAUNCH_STAGE(false, false, WV); } \
    } while (0)
            if (stage_wave) QKX_LAUNCH_STAGES(true);
            else QKX_LAUNCH_STAGES(false);
#undef QKX_LAUNCH_STAGES
#undef QKX_LAUNCH_STAGE
        }
    } else if (rt == 4) {
        if constexpr (tiles) QKX_LAUNCH_TILES(4);
    } else if (rt == 3) {
        if constexpr (tiles) QKX_LAUNCH_TILES(3);
    } else if (rt == 2) {
        if constexpr (tiles) QKX_LAUNCH_TILES(2);
    } else if constexpr (!bcg) {
        if (depth == 2) {
            if (dagger) { if (aux) QKX_LAUNCH_HOP(true, true, 2); else if constexpr (store) QKX_LAUNCH_HOP(true, false, 2); }
            else        { if (aux) QKX_LAUNCH_HOP(false, true, 2); else if constexpr (store) QKX_LAUNCH_HOP(false, false, 2); }
        } else {
            if (dagger) { if (aux) QKX_LAUNCH_HOP(true, true, 1); else if constexpr (store) QKX_LAUNCH_HOP(true, false, 1); }
            else        { if (aux) QKX_LAUNCH_HOP(false, true, 1); else if constexpr (store) QKX_LAUNCH_HOP(false, false, 1); }
        }
    }
#undef QKX_LAUNCH_HOP
#undef QKX_LAUNCH_TILE
#undef QKX_LAUNCH_TILES
    qkx::prof_end(bytes, in->prec);
    QKX_HIP(hipGetLastError());
    return QKX_SUCCESS;
}

// n independent single-RHS hops in one launch (hop_multi_kernel), UKQCD
// basis, each instance on its own lattice: the fixed default launch shape
// (256 threads, the precision's pipeline depth, XCD remapping per instance).
template <class S, class G>
int launch_hop_multi_t(int n, qkx_spinor_s *const *out, const qkx_spinor_s *const *in,
                       const qkx_spinor_s *const *aux, const qkx_gauge_s *const *U, int dagger, int par_out,
                       const Epi *e, double *const *norm_partial, int *nblocks_out, const int *const *skip)
{
    MultiHop<S, G> m{};
    m.n = n;
    const int bs = 256;
    int total = 0;
    double bytes = 0;
    for (int i = 0; i < n; i++) {
        HopArgs<S, S, S, G, 1> &a = m.inst[i];
        a.g = qkx::make_geom(U[i]->L, U[i]->bc);
        a.in = make_spinor<S>(in[i]);
        a.out = make_spinor<S>(out[i]);
        a.aux = aux && aux[i] ? make_spinor<S>(aux[i]) : a.out;
        a.U = make_gauge<G>(U[i]);
        qkx::parity_offsets(in[i], a.in_off);
        qkx::parity_offsets(out[i], a.out_off);
        if (aux && aux[i]) qkx::parity_offsets(aux[i], a.aux_off);
        else a.aux_off[0] = a.aux_off[1] = 0;
        a.par_out = par_out;
        a.nrhs = 1;
        a.dnr = make_fastdiv(1);
        const int sites_out = par_out < 0 ? a.g.V : a.g.Vh;
        a.nout = sites_out;
        a.epi[0] = e[i];
        a.skip = skip ? skip[i] : nullptr;
        a.norm_partial = norm_partial ? norm_partial[i] : nullptr;
        a.tr = HopTrav{};
        int nb = (sites_out + bs - 1) / bs;
        a.nblk8 = 0;
        if (g_xcd_remap && nb >= 64) {
            if (par_out >= 0) {
                nb = (nb + 7) / 8 * 8;
                a.nblk8 = nb / 8;
            } else {
                const int nbh = ((a.g.Vh + bs - 1) / bs + 7) / 8 * 8;
                a.nblk8 = nbh / 8;
                nb = 2 * nbh;
            }
        } else {
            nb = (nb + 7) / 8 * 8;   // keep the next instance's range aligned to the XCD rounds
        }
        if (nblocks_out) nblocks_out[i] = nb;
        m.blk0[i] = total;
        total += nb;
        const double ws = (double)qkx::spinor_elem_bytes(in[i]->prec) * qkx::spinor_nchunks(in[i]->prec) +
                          (in[i]->prec == QKX_PREC_HALF ? 4.0 : 0.0);
        const double wg = (U[i]->prec == QKX_PREC_DOUBLE ? 16.0 : U[i]->prec == QKX_PREC_SINGLE ? 8.0 : 4.0) *
                          (U[i]->recon == QKX_RECON_18 ? 9 : 6);
        bytes += (double)sites_out * ((par_out < 0 ? 4 : 8) * wg + 2 * ws) +
                 (aux && aux[i] && aux[i] != in[i] ? (double)sites_out * ws : 0.0);
    }
    m.blk0[n] = total;
    qkx::prof_begin();
    const int depth = sizeof(typename S::Real) == 8 ? g_depth64 : g_depth32;
    const dim3 grid(total), blk(bs);
    hipStream_t st = qkx::g_stream;
    constexpr int U0 = QKX_BASIS_UKQCD;
#define QKX_LAUNCH_MULTI(DG, AX, DP) hipLaunchKernelGGL((hop_multi_kernel<S, G, U0, DG, AX, DP>), grid, blk, 0, st, m)
    const bool ax = aux != nullptr;
    if (depth == 2) {
        if (dagger) { if (ax) QKX_LAUNCH_MULTI(true, true, 2); else QKX_LAUNCH_MULTI(true, false, 2); }
        else        { if (ax) QKX_LAUNCH_MULTI(false, true, 2); else QKX_LAUNCH_MULTI(false, false, 2); }
    } else {
        if (dagger) { if (ax) QKX_LAUNCH_MULTI(true, true, 1); else QKX_LAUNCH_MULTI(true, false, 1); }
        else        { if (ax) QKX_LAUNCH_MULTI(false, true, 1); else QKX_LAUNCH_MULTI(false, false, 1); }
    }
#undef QKX_LAUNCH_MULTI
    qkx::prof_end(bytes, in[0]->prec);
    QKX_HIP(hipGetLastError());
    return QKX_SUCCESS;
}

template <class S, int B>
int launch_site_t(qkx_spinor_s *out, const qkx_spinor_s *y, const qkx_spinor_s *x, int keep_par, const EpiSet &e)
{
    int n = entries(out);
    int nb = (n + 255) / 256;
    S so = make_spinor<S>(out), sy = make_spinor<S>(y), sx = x ? make_spinor<S>(x) : so;
    hipLaunchKernelGGL((site_kernel<S, B>), dim3(nb), dim3(256), 0, qkx::g_stream, so, sy, sx, n,
                       out->subset == QKX_SUBSET_FULL ? out->Vh * out->nrhs : n, x ? 1 : 0, keep_par, out->nrhs,
                       e);
    QKX_HIP(hipGetLastError());
    return QKX_SUCCESS;
}

} // namespace

// ---------------------------------------------------------------------------
// internal entry points used by the C API and the solvers
namespace qkx {

void set_hop_skip(const int *flag) { g_skip = flag; }

void set_rhs_tile(int rt) { g_rhs_tile = rt; }

void set_site_order(int blocked) { g_site_order = blocked; }

void set_neighbour_staging(int on) { g_nb_stage = on; }

void set_batch_link_sharing(int on) { g_share_links = on; }

void set_autotune(int mode)
{
    g_autotune = mode;
    g_tuned.clear();
}

void set_no_tune(int on) { g_no_tune = on; }

// "kind shape time default_time" lines of the tuned launch kinds
std::string tune_report()
{
    std::string r;
    for (const auto &kv : g_tuned) {
        const TuneKey &k = kv.first;
        const TuneEntry &t = kv.second;
        char line[256];
        snprintf(line, sizeof line,
                 "spinor_prec=%d gauge_prec=%d recon=%d nrhs=%d sites=%d full=%d dagger=%d xpay=%d basis=%d : "
                 "block=%d share=%d depth=%d %.2f us (default %.2f us)\n",
                 k.sp, k.gp, k.gr, k.nrhs, k.sites, k.full, k.dag, k.aux, k.basis, t.shape.block, t.shape.share,
                 t.shape.depth, t.us, t.default_us);
        r += line;
    }
    return r;
}

void set_hop_tuning(int xcd_remap, int block, int depth64, int depth32)
{
    g_xcd_remap = xcd_remap;
    g_block = block;
    g_depth64 = depth64;
    g_depth32 = depth32;
}

// Hop kernel: out(parity par_out or all) = epilogue(H_or_Hdag in, aux)
int hop(qkx_spinor_s *out, const qkx_spinor_s *in, const qkx_spinor_s *aux, const qkx_gauge_s *U,
        int basis, int dagger, int par_out, double cR, double gR, double cX, double gX, int g5out,
        double *norm_partial, int *nblocks_out)
{
    double coef[QKX_MAX_RHS][4];
    for (int r = 0; r < QKX_MAX_RHS; r++) {
        coef[r][0] = cR;
        coef[r][1] = gR;
        coef[r][2] = cX;
        coef[r][3] = gX;
    }
    return hop_batch(out, in, aux, U, basis, dagger, par_out, coef, g5out, norm_partial, nblocks_out);
}

int hop_batch(qkx_spinor_s *out, const qkx_spinor_s *in, const qkx_spinor_s *aux, const qkx_gauge_s *U,
              int basis, int dagger, int par_out, const double coef[][4], int g5out, double *norm_partial,
              int *nblocks_out)
{
    if (out->nrhs != in->nrhs || (aux && aux->nrhs != in->nrhs)) {
        set_error("hop: fields hold different numbers of right-hand sides");
        return QKX_ERR_ARG;
    }
    Epi e[QKX_MAX_RHS];
    for (int r = 0; r < in->nrhs; r++) e[r] = Epi{coef[r][0], coef[r][1], coef[r][2], coef[r][3], g5out};
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
            if (basis == QKX_BASIS_UKQCD) return launch_hop_t<SpinorD, GaugeD18, QKX_BASIS_UKQCD>(out, in, aux, U, dagger, par_out, e, norm_partial, nblocks_out);
            return launch_hop_t<SpinorD, GaugeD18, QKX_BASIS_CHIRAL>(out, in, aux, U, dagger, par_out, e, norm_partial, nblocks_out);
        } else {
            if (basis == QKX_BASIS_UKQCD) return launch_hop_t<SpinorD, GaugeD12, QKX_BASIS_UKQCD>(out, in, aux, U, dagger, par_out, e, norm_partial, nblocks_out);
            return launch_hop_t<SpinorD, GaugeD12, QKX_BASIS_CHIRAL>(out, in, aux, U, dagger, par_out, e, norm_partial, nblocks_out);
        }
    }
    if (sp == QKX_PREC_SINGLE) {
        if (U->prec == QKX_PREC_SINGLE && U->recon == QKX_RECON_18)
            return launch_hop_t<SpinorF, GaugeF18, QKX_BASIS_UKQCD>(out, in, aux, U, dagger, par_out, e, norm_partial, nblocks_out);
        if (U->prec == QKX_PREC_SINGLE && U->recon == QKX_RECON_12)
            return launch_hop_t<SpinorF, GaugeF12, QKX_BASIS_UKQCD>(out, in, aux, U, dagger, par_out, e, norm_partial, nblocks_out);
    }
    if (sp == QKX_PREC_HALF) {
        if (U->prec == QKX_PREC_HALF && U->recon == QKX_RECON_12)
            return launch_hop_t<SpinorH, GaugeH12, QKX_BASIS_UKQCD>(out, in, aux, U, dagger, par_out, e, norm_partial, nblocks_out);
        if (U->prec == QKX_PREC_SINGLE && U->recon == QKX_RECON_12)
            return launch_hop_t<SpinorH, GaugeF12, QKX_BASIS_UKQCD>(out, in, aux, U, dagger, par_out, e, norm_partial, nblocks_out);
    }
    if (sp == QKX_PREC_BF16) {
        if (U->prec == QKX_PREC_BF16 && U->recon == QKX_RECON_12)
            return launch_hop_t<SpinorB, GaugeB12, QKX_BASIS_UKQCD>(out, in, aux, U, dagger, par_out, e, norm_partial, nblocks_out);
    }
    set_error("hop: unsupported spinor/gauge precision combination");
    return QKX_ERR_ARG;
}

// The last hop of a CG normal operator with the CG update fused in (UKQCD
// basis, single right-hand side).  in/aux/p share one precision; x and r are
// the solver's accumulator and residual.
template <class S, class G>
int launch_hop_cg(const qkx_spinor_s *in, const qkx_spinor_s *aux, const qkx_gauge_s *U, int dagger, int par_out,
                  const Epi &e, const CgFuse &f, int *nblocks_out)
{
    auto go = [&](auto X, auto R) {
        CgFin<decltype(X), decltype(R), S> fin{X, R, make_spinor<S>(f.p), f.r2cur, f.pap_part, f.npap, f.r2_part,
                                               0.0};
        return launch_hop_t<S, G, QKX_BASIS_UKQCD>(f.p, in, aux, U, dagger, par_out, &e, nullptr, nblocks_out, fin);
    };
    const int xp = f.x->prec, rp = f.r->prec;
    if (f.x_in_dir) {
        // r only: x (FP32) is updated by the direction kernel
        auto go_r = [&](auto R) {
            CgFin<SpinorF, decltype(R), S, false> fin{make_spinor<SpinorF>(f.x), R, make_spinor<S>(f.p), f.r2cur,
                                                      f.pap_part, f.npap, f.r2_part, 0.0};
            return launch_hop_t<S, G, QKX_BASIS_UKQCD>(f.p, in, aux, U, dagger, par_out, &e, nullptr, nblocks_out,
                                                       fin);
        };
        if constexpr (!std::is_same_v<S, SpinorD>) {
            if (xp == QKX_PREC_SINGLE && rp == QKX_PREC_SINGLE) return go_r(make_spinor<SpinorF>(f.r));
            if constexpr (std::is_same_v<S, SpinorH>) {
                if (xp == QKX_PREC_SINGLE && rp == QKX_PREC_HALF) return go_r(make_spinor<SpinorH>(f.r));
            }
        }
        set_error("hop_cg: x updated in the direction kernel needs an FP32 x and a 32- or 16-bit p");
        return QKX_ERR_ARG;
    }
    if constexpr (std::is_same_v<S, SpinorD>) {
        if (xp == QKX_PREC_DOUBLE && rp == QKX_PREC_DOUBLE) return go(make_spinor<SpinorD>(f.x), make_spinor<SpinorD>(f.r));
    } else {
        if (xp == QKX_PREC_SINGLE && rp == QKX_PREC_SINGLE) return go(make_spinor<SpinorF>(f.x), make_spinor<SpinorF>(f.r));
        if constexpr (std::is_same_v<S, SpinorH>) {
            if (xp == QKX_PREC_SINGLE && rp == QKX_PREC_HALF) return go(make_spinor<SpinorF>(f.x), make_spinor<SpinorH>(f.r));
        }
    }
    set_error("hop_cg: unsupported x/r precision combination");
    return QKX_ERR_ARG;
}

bool hop_cg_batch_available(const qkx_spinor_s *in, const qkx_gauge_s *U, const qkx_spinor_s *x,
                            const qkx_spinor_s *r)
{
    // the fused update is an even-odd hop: Vh output sites
    return QKX_HOP16_DOT2 && in->prec == QKX_PREC_HALF && U->prec == QKX_PREC_HALF && U->recon == QKX_RECON_12 &&
           in->nrhs <= QKX_MAX_RHS && share_links_for(in->nrhs, (long)in->Vh * in->nrhs) &&
           x->prec == QKX_PREC_SINGLE && r->prec == QKX_PREC_HALF;
}

int hop_cg_batch(const qkx_spinor_s *in, const qkx_spinor_s *aux, const qkx_gauge_s *U, int dagger, int par_out,
                 const double coef[][4], int g5out, const CgBatchFuse &f, int *nblocks_out)
{
    if (!hop_cg_batch_available(in, U, f.x, f.r) || !aux || f.p->prec != QKX_PREC_HALF ||
        aux->prec != QKX_PREC_HALF || f.p->nrhs != in->nrhs || f.x->nrhs != in->nrhs || f.r->nrhs != in->nrhs) {
        set_error("hop_cg_batch: FP16 fields of 3-16 RHS on FP16 links (shared links in use), FP32 x, FP16 r and p");
        return QKX_ERR_ARG;
    }
    Epi e[QKX_MAX_RHS];
    for (int r = 0; r < in->nrhs; r++) e[r] = Epi{coef[r][0], coef[r][1], coef[r][2], coef[r][3], g5out};
    if (f.x_in_dir) {
        BCgFin<SpinorF, SpinorH, SpinorH, false> fin{make_spinor<SpinorF>(f.x), make_spinor<SpinorH>(f.r),
                                                     make_spinor<SpinorH>(f.p), f.r2cur, f.stop2, f.active, f.pap,
                                                     f.r2_part};
        return launch_hop_t<SpinorH, GaugeH12, QKX_BASIS_UKQCD>(f.p, in, aux, U, dagger, par_out, e, nullptr,
                                                                nblocks_out, fin);
    }
    BCgFin<SpinorF, SpinorH, SpinorH> fin{make_spinor<SpinorF>(f.x), make_spinor<SpinorH>(f.r),
                                          make_spinor<SpinorH>(f.p), f.r2cur, f.stop2, f.active, f.pap, f.r2_part};
    return launch_hop_t<SpinorH, GaugeH12, QKX_BASIS_UKQCD>(f.p, in, aux, U, dagger, par_out, e, nullptr, nblocks_out,
                                                            fin);
}

int hop_cg(const qkx_spinor_s *in, const qkx_spinor_s *aux, const qkx_gauge_s *U, int dagger, int par_out,
           double cR, double gR, double cX, double gX, int g5out, const CgFuse &f, int *nblocks_out)
{
    if (in->nrhs != 1 || f.p->prec != in->prec || (aux && aux->prec != in->prec)) {
        set_error("hop_cg: single right-hand side, one precision for in, aux and p");
        return QKX_ERR_ARG;
    }
    const Epi e{cR, gR, cX, gX, g5out};
    const int sp = in->prec;
    if (sp == QKX_PREC_DOUBLE && U->prec == QKX_PREC_DOUBLE)
        return U->recon == QKX_RECON_18 ? launch_hop_cg<SpinorD, GaugeD18>(in, aux, U, dagger, par_out, e, f, nblocks_out)
                                         : launch_hop_cg<SpinorD, GaugeD12>(in, aux, U, dagger, par_out, e, f, nblocks_out);
    if (sp == QKX_PREC_SINGLE && U->prec == QKX_PREC_SINGLE)
        return U->recon == QKX_RECON_18 ? launch_hop_cg<SpinorF, GaugeF18>(in, aux, U, dagger, par_out, e, f, nblocks_out)
                                         : launch_hop_cg<SpinorF, GaugeF12>(in, aux, U, dagger, par_out, e, f, nblocks_out);
    if (sp == QKX_PREC_HALF && U->prec == QKX_PREC_HALF && U->recon == QKX_RECON_12)
        return launch_hop_cg<SpinorH, GaugeH12>(in, aux, U, dagger, par_out, e, f, nblocks_out);
    if (sp == QKX_PREC_HALF && U->prec == QKX_PREC_SINGLE && U->recon == QKX_RECON_12)
        return launch_hop_cg<SpinorH, GaugeF12>(in, aux, U, dagger, par_out, e, f, nblocks_out);
    if (sp == QKX_PREC_BF16 && U->prec == QKX_PREC_BF16 && U->recon == QKX_RECON_12)
        return launch_hop_cg<SpinorB, GaugeB12>(in, aux, U, dagger, par_out, e, f, nblocks_out);
    set_error("hop_cg: unsupported spinor/gauge precision combination");
    return QKX_ERR_ARG;
}

// n independent single-RHS hops, each on its own lattice and links, in one
// launch; see qkx_core.h
int hop_multi(int n, qkx_spinor_s *const *out, const qkx_spinor_s *const *in, const qkx_spinor_s *const *aux,
              const qkx_gauge_s *const *U, int dagger, int par_out, const double (*coef)[4], int g5out,
              double *const *norm_partial, int *nblocks_out, const int *const *skip)
{
    if (n < 1 || n > kMaxInst) {
        set_error("hop_multi: 1 to 8 instances");
        return QKX_ERR_ARG;
    }
    Epi e[kMaxInst];
    const int sp = in[0]->prec, gp = U[0]->prec, gr = U[0]->recon;
    for (int i = 0; i < n; i++) {
        if (in[i]->nrhs != 1 || out[i]->nrhs != 1 || in[i]->prec != sp || out[i]->prec != sp ||
            (aux && (!aux[i] || aux[i]->prec != sp || aux[i]->nrhs != 1)) || U[i]->prec != gp || U[i]->recon != gr) {
            set_error("hop_multi: single right-hand-side fields of one precision, links of one storage");
            return QKX_ERR_ARG;
        }
        e[i] = Epi{coef[i][0], coef[i][1], coef[i][2], coef[i][3], g5out};
    }
    if (sp == QKX_PREC_DOUBLE && gp == QKX_PREC_DOUBLE)
        return gr == QKX_RECON_18 ? launch_hop_multi_t<SpinorD, GaugeD18>(n, out, in, aux, U, dagger, par_out, e, norm_partial, nblocks_out, skip)
                                  : launch_hop_multi_t<SpinorD, GaugeD12>(n, out, in, aux, U, dagger, par_out, e, norm_partial, nblocks_out, skip);
    if (sp == QKX_PREC_SINGLE && gp == QKX_PREC_SINGLE && gr == QKX_RECON_12)
        return launch_hop_multi_t<SpinorF, GaugeF12>(n, out, in, aux, U, dagger, par_out, e, norm_partial, nblocks_out, skip);
    if (sp == QKX_PREC_HALF && gp == QKX_PREC_HALF && gr == QKX_RECON_12)
        return launch_hop_multi_t<SpinorH, GaugeH12>(n, out, in, aux, U, dagger, par_out, e, norm_partial, nblocks_out, skip);
    if (sp == QKX_PREC_BF16 && gp == QKX_PREC_BF16 && gr == QKX_RECON_12)
        return launch_hop_multi_t<SpinorB, GaugeB12>(n, out, in, aux, U, dagger, par_out, e, norm_partial, nblocks_out, skip);
    set_error("hop_multi: unsupported spinor/gauge precision combination");
    return QKX_ERR_ARG;
}

// Site-local: out = cR y + gR i g5 y + cX x + gX i g5 x (x may be null); keep_par >= 0
// zeroes the other parity of a full field.
int site(qkx_spinor_s *out, const qkx_spinor_s *y, const qkx_spinor_s *x, int basis, int keep_par,
         double cR, double gR, double cX, double gX, int g5out)
{
    double coef[QKX_MAX_RHS][4];
    for (int r = 0; r < QKX_MAX_RHS; r++) {
        coef[r][0] = cR;
        coef[r][1] = gR;
        coef[r][2] = cX;
        coef[r][3] = gX;
    }
    return site_batch(out, y, x, basis, keep_par, coef, g5out);
}

int site_batch(qkx_spinor_s *out, const qkx_spinor_s *y, const qkx_spinor_s *x, int basis, int keep_par,
               const double coef[][4], int g5out)
{
    EpiSet e;
    for (int r = 0; r < QKX_MAX_RHS; r++) e.e[r] = Epi{coef[r][0], coef[r][1], coef[r][2], coef[r][3], g5out};
    if (out->prec != y->prec || (x && x->prec != y->prec) || entries(out) != entries(y) ||
        (x && entries(x) != entries(y)) || out->nrhs != y->nrhs || (x && x->nrhs != y->nrhs) ||
        out->subset != y->subset || (x && x->subset != y->subset)) {
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
