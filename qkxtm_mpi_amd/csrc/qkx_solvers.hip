// qkx_solvers.hip -- CG on the normal equations, on the device.
//
//  * qkx_cg: the reference CG (lib/solvers.cpp:16-81) -- full lattice, FP64,
//    D^dag D x = b, r0 = b, stop on the iterated |r|^2 <= tol^2 |b|^2.
//  * qkx_solve: D x = b to a TRUE relative residual, on the full lattice or on
//    the even-odd Schur system  M = A_ee - k^2 H_eo A_oo^-1 H_oe  (the operator
//    of Vector::tm_xx, lib/lattice_util.cpp:2001-2022), with the inner CG in
//    FP64, FP32, FP16 or BF16 storage and FP64 reliable updates: the FP64
//    solution accumulates the low-precision increments, and the residual is
//    recomputed in FP64 whenever the iterated residual drops by 'delta'.
//
// One CG iteration is six launches and no host synchronisation:
//   Op p (1-2 hops, the last one also sums |Op p|^2 = p.A p per workgroup),
//   Op^dag (1-2 hops), update (x += a p, r -= a Ap, per-workgroup |r|^2),
//   direction (p = r + b p).
// Reductions never use device-wide fences: a producer writes one partial per
// workgroup and every workgroup of the consuming launch adds the partials in
// the same fixed order, so all workgroups see bit-identical scalars and the
// sums are reproducible.  Per-iteration scalars live in two slots used
// alternately (slot j&1 read, slot (j+1)&1 written by iteration j); only the
// direction launch writes a slot, and only the slot no one reads in that
// launch.  The host enqueues a batch of iterations and reads the slot back
// once per batch; a raised flag (converged / reliable update due / breakdown)
// makes every later launch of the batch return at entry.
#include "qkx_internal.h"
#include "qkx_core.h"
#include <chrono>
#include <algorithm>
#include <memory>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace {

constexpr int kThreads = 256;

struct CGSlot {
    double r2;      // iterated |r|^2 after the iteration
    double r2_old;  // |r|^2 before it
    double pAp;
    double maxr2;   // |r|^2 at the last reliable update
    int flag;       // 0 run, 1 reliable update due, 2 iterated convergence, 3 breakdown
    int k;          // iterations done
};

int g_cg_fuse = 2;               // fuse the CG update into the operator's last hop (2: x in cg_dir for FP16)
int g_cg_graph = 0;              // replay pairs of CG iterations from captured graphs: bit 0 single-RHS
                                 // solves, bit 1 batched solves (both off: measured no faster)
CGSlot *g_slots = nullptr;       // device, two slots
CGSlot *g_slots_host = nullptr;  // pinned mirror
double *g_part = nullptr;        // partials: [0, cap) p.Ap, [cap, 2 cap) |r|^2
int g_part_cap = 0;
double *g_red = nullptr;         // [0] p.Ap, [1] |r|^2 pre-summed when the partials are many
// Above this many partials a single-workgroup kernel sums them first, so that
// each consuming workgroup reads one value instead of all of them.
#ifndef QKX_DIRECT_PARTIALS
#define QKX_DIRECT_PARTIALS 2048
#endif
constexpr int kDirectPartials = QKX_DIRECT_PARTIALS;

int ensure_state(int nblocks)
{
    if (!g_slots) {
        QKX_HIP(hipMalloc(&g_slots, 2 * sizeof(CGSlot)));
        QKX_HIP(hipHostMalloc((void **)&g_slots_host, 2 * sizeof(CGSlot), hipHostMallocDefault));
        QKX_HIP(hipMalloc(&g_red, 2 * sizeof(double)));
    }
    if (nblocks > g_part_cap) {
        if (g_part) (void)hipFree(g_part);
        g_part = nullptr;
        QKX_HIP(hipMalloc(&g_part, sizeof(double) * 2 * (size_t)nblocks));
        g_part_cap = nblocks;
    }
    return QKX_SUCCESS;
}

// Sum of v[0..n) seen by every thread of the workgroup, compensated
// (double-double), in an order fixed by n and the workgroup size only, so all
// workgroups of a launch agree bit for bit.  The xor butterfly leaves the same
// bits in every lane (dd_add is symmetric in its operands).
__device__ double block_sum_array(const double *v, int n)
{
    __shared__ double lh[kThreads / 64], ll[kThreads / 64];
    DD s{0.0, 0.0};
    for (int i = threadIdx.x; i < n; i += blockDim.x) s = dd_add(s, v[i]);
    s = dd_wave_sum(s);
    if ((threadIdx.x & 63) == 0) {
        lh[threadIdx.x >> 6] = s.hi;
        ll[threadIdx.x >> 6] = s.lo;
    }
    __syncthreads();
    DD r{lh[0], ll[0]};
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) r = dd_add(r, DD{lh[w], ll[w]});
    __syncthreads();   // lh, ll are reused by the next call
    return dd_value(r);
}

// out[0] = sum(part[0..n)), compensated, in a fixed order, by one workgroup
__global__ __launch_bounds__(1024) void sum_partials_kernel(const double *part, int n, double *out)
{
    __shared__ double lh[1024], ll[1024];
    DD v{0.0, 0.0};
    for (int i = threadIdx.x; i < n; i += blockDim.x) v = dd_add(v, part[i]);
    lh[threadIdx.x] = v.hi;
    ll[threadIdx.x] = v.lo;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            const DD c = dd_add(DD{lh[threadIdx.x], ll[threadIdx.x]}, DD{lh[threadIdx.x + st], ll[threadIdx.x + st]});
            lh[threadIdx.x] = c.hi;
            ll[threadIdx.x] = c.lo;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = dd_value(DD{lh[0], ll[0]});
}

// partials (part, n) as the consumers should read them: pre-summed into
// dst when there are many
void reduce_if_many(const double *&part, int &n, double *dst)
{
    if (n <= kDirectPartials) return;
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(1024), 0, qkx::g_stream, part, n, dst);
    part = dst;
    n = 1;
}

__device__ void block_partial(double v, double *dst)
{
    __shared__ double lds[kThreads / 64];
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

// alpha = r2/pAp; x += alpha p; r -= alpha Ap; per-workgroup |r|^2.
// x (SX) and r (SR) may be kept in a wider type than p and Ap (S): 16-bit
// storage of p/Ap with FP32 accumulation of the solution and residual.
template <class SX, class SR, class S>
__global__ __launch_bounds__(kThreads) void cg_update_kernel(SX x, SR r, S p, S Ap, int n, const CGSlot *cur,
                                                             const double *pap_part, int npap, double *r2_part)
{
    if (cur->flag) return;
    using T = typename SX::Real;
    using TR = typename SR::Real;
    using TP = typename S::Real;
    const double pap = block_sum_array(pap_part, npap);
    const double alpha = cur->r2 / pap;
    const T a = T(alpha);
    const TR ar = TR(alpha);
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    double acc = 0;
    if (s < n) {
        Cx<T> xv[12];
        Cx<TR> rv[12];
        Cx<TP> pv[12], av[12];
        x.load(s, xv);
        p.load(s, pv);
#pragma unroll
        for (int k = 0; k < 12; k++) xv[k] = {xv[k].re + a * T(pv[k].re), xv[k].im + a * T(pv[k].im)};
        x.store(s, xv);
        r.load(s, rv);
        Ap.load(s, av);
#pragma unroll
        for (int k = 0; k < 12; k++) {
            rv[k] = {rv[k].re - ar * TR(av[k].re), rv[k].im - ar * TR(av[k].im)};
            acc += (double)rv[k].re * (double)rv[k].re + (double)rv[k].im * (double)rv[k].im;
        }
        r.store(s, rv);
    }
    block_partial(acc, r2_part + blockIdx.x);
}

// Direction update.  mode 0: iteration j -- decide the flag from the summed
// |r|^2, write the next slot (workgroup 0) and, when no flag is raised,
// p = r + beta p.  mode 1: restart after a reliable update -- |r|^2 from the
// partials, beta = |r|^2 / cur->r2_old, always update p, clear the flag.
// HASX: in mode 0 also x += alpha p with the p of the iteration (before its
// update), alpha = cur->r2 / p.Ap summed as the fused hop sums it, so x gets
// the same bits as when the hop updates it.
template <class S, class SR, class SX, bool HASX>
__global__ __launch_bounds__(kThreads) void cg_dir_kernel(S p, SR r, SX x, int n, const CGSlot *cur, CGSlot *next,
                                                          const double *r2_part, int nr2, const double *pap_part,
                                                          int npap, double stop2, double delta2, int mode)
{
    if (mode == 0 && cur->flag) {
        if (blockIdx.x == 0 && threadIdx.x == 0) *next = *cur;
        return;
    }
    // this site's p and r (and x) are fetched before the partial sums are
    // reduced, so their HBM latency overlaps the reduction
    using T = typename SR::Real;
    using TX = typename SX::Real;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    Cx<typename S::Real> pv[12];
    Cx<T> rv[12];
    Cx<TX> xv[12];
    if (s < n) {
        p.load(s, pv);
        r.load(s, rv);
        if (HASX && mode == 0) x.load(s, xv);
    }
    const double r2n = block_sum_array(r2_part, nr2);
    double beta;
    int flag = 0;
    if (mode == 0) {
        const double pap = block_sum_array(pap_part, npap);
        if constexpr (HASX) {
            if (s < n) {
                const TX ax = TX(cur->r2 / pap);
#pragma unroll
                for (int k = 0; k < 12; k++)
                    xv[k] = {xv[k].re + ax * TX(pv[k].re), xv[k].im + ax * TX(pv[k].im)};
                x.store(s, xv);
            }
        }
        if (!(pap > 0.0) || !isfinite(r2n)) flag = 3;
        else if (r2n <= stop2) flag = 2;
        else if (r2n < delta2 * cur->maxr2) flag = 1;
        beta = r2n / cur->r2;
        if (blockIdx.x == 0 && threadIdx.x == 0)
            *next = CGSlot{r2n, cur->r2, pap, cur->maxr2, flag, cur->k + 1};
    } else {
        beta = r2n / cur->r2_old;
        if (blockIdx.x == 0 && threadIdx.x == 0)
            *next = CGSlot{r2n, cur->r2_old, cur->pAp, r2n, isfinite(r2n) ? 0 : 3, cur->k};
    }
    if (flag) return;
    if (s >= n) return;
    const T b = T(beta);
    Cx<typename S::Real> out[12];
#pragma unroll
    for (int k = 0; k < 12; k++)
        out[k] = {(typename S::Real)(rv[k].re + b * T(pv[k].re)), (typename S::Real)(rv[k].im + b * T(pv[k].im))};
    p.store(s, out);
}

// every workgroup sums v[0..n) with block_sum_array and writes its result
__global__ __launch_bounds__(kThreads) void block_sum_probe_kernel(const double *v, int n, double *out)
{
    const double r = block_sum_array(v, n);
    if (threadIdx.x == 0) out[blockIdx.x] = r;
}

// per-workgroup |r|^2
template <class S>
__global__ __launch_bounds__(kThreads) void norm_partial_kernel(S r, int n, double *part)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    double acc = 0;
    if (s < n) {
        Cx<typename S::Real> a[12];
        r.load(s, a);
#pragma unroll
        for (int k = 0; k < 12; k++) acc += (double)a[k].re * (double)a[k].re + (double)a[k].im * (double)a[k].im;
    }
    block_partial(acc, part + blockIdx.x);
}

// xh += xl (FP64 += lower precision), then xl = 0
template <class SL>
__global__ __launch_bounds__(256) void accum_kernel(SpinorD xh, SL xl, int n)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    Cx<double> a[12];
    Cx<typename SL::Real> b[12];
    xh.load(s, a);
    xl.load(s, b);
#pragma unroll
    for (int k = 0; k < 12; k++) a[k] = {a[k].re + (double)b[k].re, a[k].im + (double)b[k].im};
    xh.store(s, a);
    // xl restarts from zero (stored as a zero field stores it)
#pragma unroll
    for (int k = 0; k < 12; k++) b[k] = {0, 0};
    xl.store(s, b);
}

int nblk(int n) { return (n + kThreads - 1) / kThreads; }

// ---------------------------------------------------------------------------
// Normal operator A = Op^dag Op in one precision; Op = D (full) or the Schur M.
struct NormalOp {
    const qkx_gauge_s *U;
    double kappa, a, nrm;
    bool eo;
    qkx_spinor_s *tmp_full;  // full-lattice temp (full solve)
    qkx_spinor_s *tmp_odd;   // odd temp (eo)
    qkx_spinor_s *tmp_even;  // even temp (eo)
    // out = A in; per-workgroup |Op in|^2 into part (count in *np)
    int apply(qkx_spinor_s *out, const qkx_spinor_s *in, double *part = nullptr, int *np = nullptr) const
    {
        if (!eo) {
            int rc = qkx::hop(tmp_full, in, in, U, QKX_BASIS_UKQCD, 0, -1, -kappa, 0, 1.0, a, 0, part, np);
            if (rc) return rc;
            return qkx::hop(out, tmp_full, tmp_full, U, QKX_BASIS_UKQCD, 1, -1, -kappa, 0, 1.0, -a, 0);
        }
        // M^dag M = g5 M(-s) g5 M(s)  (tm_xxDagTm_xx, lib/lattice_util.cpp:2024-2035)
        int rc = qkx::schur(tmp_even, in, U, kappa, a, nrm, QKX_BASIS_UKQCD, 0, 1, tmp_odd, part, np);
        if (rc) return rc;
        return qkx::schur(out, tmp_even, U, kappa, -a, nrm, QKX_BASIS_UKQCD, 0, 1, tmp_odd);
    }
    // The same with the CG update fused into the last hop (no Ap field): the
    // first half leaves the |Op p|^2 partials, which f then points at
    // (pre-summed when many); the last hop updates x and r and leaves |r|^2
    // partials in f.r2_part, *nr2 of them.
    int apply_cg(const qkx_spinor_s *in, double *part, qkx::CgFuse &f, int *nr2) const
    {
        int np = 0, rc;
        if (!eo) {
            if ((rc = qkx::hop(tmp_full, in, in, U, QKX_BASIS_UKQCD, 0, -1, -kappa, 0, 1.0, a, 0, part, &np)))
                return rc;
        } else if ((rc = qkx::schur(tmp_even, in, U, kappa, a, nrm, QKX_BASIS_UKQCD, 0, 1, tmp_odd, part, &np))) {
            return rc;
        }
        const double *pp = part;
        reduce_if_many(pp, np, g_red);
        f.pap_part = pp;
        f.npap = np;
        if (!eo) return qkx::hop_cg(tmp_full, tmp_full, U, 1, -1, -kappa, 0, 1.0, -a, 0, f, nr2);
        // second Schur half with the twist sign flipped: hop to the odd sites, then the fused hop back
        if ((rc = qkx::hop(tmp_odd, tmp_even, nullptr, U, QKX_BASIS_UKQCD, 0, 1, 1.0 / nrm, a / nrm, 0.0, 0.0, 0)))
            return rc;
        return qkx::hop_cg(tmp_odd, tmp_even, U, 0, 0, -kappa * kappa, 0.0, 1.0, -a, 1, f, nr2);
    }
    int apply_op(qkx_spinor_s *out, const qkx_spinor_s *in) const
    {
        if (!eo) return qkx::hop(out, in, in, U, QKX_BASIS_UKQCD, 0, -1, -kappa, 0, 1.0, a, 0);
        return qkx::schur(out, in, U, kappa, a, nrm, QKX_BASIS_UKQCD, 0, 0, tmp_odd);
    }
    int apply_dag(qkx_spinor_s *out, const qkx_spinor_s *in) const
    {
        if (!eo) return qkx::hop(out, in, in, U, QKX_BASIS_UKQCD, 1, -1, -kappa, 0, 1.0, -a, 0);
        int rc = qkx::site(tmp_even, in, nullptr, QKX_BASIS_UKQCD, -1, 1.0, 0, 0, 0, 1);
        if (rc) return rc;
        return qkx::schur(out, tmp_even, U, kappa, -a, nrm, QKX_BASIS_UKQCD, 0, 1, tmp_odd);
    }
};

template <class F> int with_spinor(const qkx_spinor_s *f, F fn)
{
    return dispatch_prec(f->prec, [&](auto tag) { return fn(make_spinor<decltype(tag)>(f)); });
}

// CG engine state on the host
struct Engine {
    int j = 0;            // iterations enqueued (selects the slot parity)
    double stop2 = 0;     // iterated stopping threshold
    double delta2 = 0;    // reliable-update trigger
    CGSlot *cur() const { return g_slots + (j & 1); }
    CGSlot *next() const { return g_slots + ((j + 1) & 1); }
};

template <class SX, class SR, class S>
int enqueue_blas(Engine &E, qkx_spinor_s *x, qkx_spinor_s *r, qkx_spinor_s *p, qkx_spinor_s *Ap, int npap)
{
    int n = entries(p);
    SX X = make_spinor<SX>(x);
    SR R = make_spinor<SR>(r);
    S P = make_spinor<S>(p), AP = make_spinor<S>(Ap);
    const double *pap_part = g_part;
    double *r2_part = g_part + g_part_cap;
    reduce_if_many(pap_part, npap, g_red);
    hipLaunchKernelGGL((cg_update_kernel<SX, SR, S>), dim3(nblk(n)), dim3(kThreads), 0, qkx::g_stream, X, R, P, AP, n,
                       E.cur(), pap_part, npap, r2_part);
    const double *r2_src = r2_part;
    int nr2 = nblk(n);
    reduce_if_many(r2_src, nr2, g_red + 1);
    hipLaunchKernelGGL((cg_dir_kernel<S, SR, S, false>), dim3(nblk(n)), dim3(kThreads), 0, qkx::g_stream, P, R, P, n,
                       E.cur(), E.next(), r2_src, nr2, pap_part, npap, E.stop2, E.delta2, 0);
    QKX_HIP(hipGetLastError());
    return QKX_SUCCESS;
}

// direction step alone (the update was fused into the operator's last hop;
// with f.x_in_dir the x part of it is done here, on an FP32 x)
template <class S, class SR>
int enqueue_dir(Engine &E, qkx_spinor_s *r, qkx_spinor_s *p, const qkx::CgFuse &f, int nr2)
{
    const int n = entries(p);
    const double *r2_src = f.r2_part;
    reduce_if_many(r2_src, nr2, g_red + 1);
    const S P = make_spinor<S>(p);
    if constexpr (!std::is_same_v<S, SpinorD>) {
        if (f.x_in_dir) {
            hipLaunchKernelGGL((cg_dir_kernel<S, SR, SpinorF, true>), dim3(nblk(n)), dim3(kThreads), 0, qkx::g_stream,
                               P, make_spinor<SR>(r), make_spinor<SpinorF>(f.x), n, E.cur(), E.next(), r2_src, nr2,
                               f.pap_part, f.npap, E.stop2, E.delta2, 0);
            QKX_HIP(hipGetLastError());
            return QKX_SUCCESS;
        }
    }
    hipLaunchKernelGGL((cg_dir_kernel<S, SR, S, false>), dim3(nblk(n)), dim3(kThreads), 0, qkx::g_stream, P,
                       make_spinor<SR>(r), P, n, E.cur(), E.next(), r2_src, nr2, f.pap_part, f.npap, E.stop2, E.delta2,
                       0);
    QKX_HIP(hipGetLastError());
    return QKX_SUCCESS;
}

int enqueue_iteration(Engine &E, const NormalOp &A, qkx_spinor_s *x, qkx_spinor_s *r, qkx_spinor_s *p,
                      qkx_spinor_s *Ap)
{
    const int px = x->prec, pr = r->prec, pp = p->prec;
    if (g_cg_fuse) {
        qkx::set_hop_skip(&E.cur()->flag);
        qkx::CgFuse f{x, r, p, &E.cur()->r2, nullptr, 0, g_part + g_part_cap};
        // mode 2: an FP32 x with a 32- or 16-bit p (the mixed solves' inner CG)
        // is updated by the direction kernel, which reads p anyway, so the
        // fused hop does not read p
        f.x_in_dir = g_cg_fuse == 2 && px == QKX_PREC_SINGLE && pp != QKX_PREC_DOUBLE;
        int nr2 = 0;
        int rc = A.apply_cg(p, g_part, f, &nr2);
        qkx::set_hop_skip(nullptr);
        if (rc) return rc;
        if (pp == QKX_PREC_DOUBLE) rc = enqueue_dir<SpinorD, SpinorD>(E, r, p, f, nr2);
        else if (pp == QKX_PREC_SINGLE) rc = enqueue_dir<SpinorF, SpinorF>(E, r, p, f, nr2);
        else if (pp == QKX_PREC_HALF) rc = enqueue_dir<SpinorH, SpinorH>(E, r, p, f, nr2);
        else rc = enqueue_dir<SpinorB, SpinorF>(E, r, p, f, nr2);
        E.j++;
        return rc;
    }
    qkx::set_hop_skip(&E.cur()->flag);
    int npap = 0;
    int rc = A.apply(Ap, p, g_part, &npap);
    qkx::set_hop_skip(nullptr);
    if (rc) return rc;
    if (px == QKX_PREC_DOUBLE && pr == QKX_PREC_DOUBLE && pp == QKX_PREC_DOUBLE)
        rc = enqueue_blas<SpinorD, SpinorD, SpinorD>(E, x, r, p, Ap, npap);
    else if (px == QKX_PREC_SINGLE && pr == QKX_PREC_SINGLE && pp == QKX_PREC_SINGLE)
        rc = enqueue_blas<SpinorF, SpinorF, SpinorF>(E, x, r, p, Ap, npap);
    else if (px == QKX_PREC_SINGLE && pr == QKX_PREC_HALF && pp == QKX_PREC_HALF)
        rc = enqueue_blas<SpinorF, SpinorH, SpinorH>(E, x, r, p, Ap, npap);
    else if (px == QKX_PREC_SINGLE && pr == QKX_PREC_SINGLE && pp == QKX_PREC_BF16)
        rc = enqueue_blas<SpinorF, SpinorF, SpinorB>(E, x, r, p, Ap, npap);
    else {
        qkx::set_error("cg: unsupported precision combination");
        return QKX_ERR_ARG;
    }
    E.j++;
    return rc;
}

// Pairs of CG iterations replayed from a captured graph.  After two
// iterations the slot parity is back where it started, so one executable per
// starting parity serves a whole solve; a replay costs one launch instead of
// the ten or so kernel launches of two iterations.  On MI355X the eager
// launches already run ahead of the GPU inside an iteration batch, so replay
// saves no GPU time and the capture adds about 0.2 ms per solve: off by
// default (qkx_set_cg_graph), kept for hosts whose launch cost is higher.  An executable bakes
// in every kernel argument: it is captured again when one changes (the
// stopping threshold after an early iterated stop) and freed with the solve.
// The captured work is kernels only (no memset or copy nodes).  Captures that
// fail (or instrumentation, whose events cannot be captured) fall back to
// launching the iterations one by one.
struct IterGraphs {
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    double stop2[2] = {0.0, 0.0};
    bool broken = false;
    IterGraphs() = default;
    IterGraphs(const IterGraphs &) = delete;
    IterGraphs &operator=(const IterGraphs &) = delete;
    ~IterGraphs()
    {
        for (auto &e : exec)
            if (e) (void)hipGraphExecDestroy(e);
    }
    // capture two iterations from parity par; false if capture is unavailable
    bool capture(Engine &E, const NormalOp &A, qkx_spinor_s *x, qkx_spinor_s *r, qkx_spinor_s *p, qkx_spinor_s *Ap,
                 int par)
    {
        if (exec[par]) {
            (void)hipGraphExecDestroy(exec[par]);
            exec[par] = nullptr;
        }
        if (hipStreamBeginCapture(qkx::g_stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        const int j0 = E.j;
        int rc = enqueue_iteration(E, A, x, r, p, Ap);
        if (!rc) rc = enqueue_iteration(E, A, x, r, p, Ap);
        E.j = j0;   // nothing ran yet: the launch of the graph runs the pair
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(qkx::g_stream, &g);
        bool ok = !rc && ec == hipSuccess && g;
        if (ok && hipGraphInstantiate(&exec[par], g, nullptr, nullptr, 0) != hipSuccess) {
            exec[par] = nullptr;
            ok = false;
        }
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        stop2[par] = E.stop2;
        return ok;
    }
    int run(Engine &E, const NormalOp &A, qkx_spinor_s *x, qkx_spinor_s *r, qkx_spinor_s *p, qkx_spinor_s *Ap,
            int n)
    {
        int rc;
        const bool use = (g_cg_graph & 1) && !qkx::prof_active();
        while (n > 0) {
            const int par = E.j & 1;
            if (use && !broken && n >= 2) {
                if ((!exec[par] || stop2[par] != E.stop2) && !capture(E, A, x, r, p, Ap, par)) broken = true;
                if (!broken) {
                    QKX_HIP(hipGraphLaunch(exec[par], qkx::g_stream));
                    E.j += 2;
                    n -= 2;
                    continue;
                }
            }
            if ((rc = enqueue_iteration(E, A, x, r, p, Ap))) return rc;
            n--;
        }
        return QKX_SUCCESS;
    }
};

// Start (or restart after a reliable update) the recurrence from residual r:
// mode 0 -> p = r, slot = {|r|^2, flag 0, k}; mode 1 -> p = r + beta p with
// beta = |r|^2 / r2_old of the current slot.
int restart(Engine &E, qkx_spinor_s *p, qkx_spinor_s *r, int mode)
{
    int n = entries(r);
    int rc = with_spinor(r, [&](auto R) {
        hipLaunchKernelGGL((norm_partial_kernel<decltype(R)>), dim3(nblk(n)), dim3(kThreads), 0, qkx::g_stream, R, n,
                           g_part + g_part_cap);
        QKX_HIP(hipGetLastError());
        return QKX_SUCCESS;
    });
    if (rc) return rc;
    if (mode == 0) {
        if ((rc = qkx_spinor_copy(p, r))) return rc;
    }
    const double *r2_src = g_part + g_part_cap;
    int nr2 = nblk(n);
    reduce_if_many(r2_src, nr2, g_red + 1);
    rc = with_spinor(p, [&](auto P) {
        return with_spinor(r, [&](auto R) {
            hipLaunchKernelGGL((cg_dir_kernel<decltype(P), decltype(R), decltype(P), false>), dim3(nblk(n)),
                               dim3(kThreads), 0, qkx::g_stream, P, R, P, mode == 0 ? 0 : n, E.cur(), E.next(),
                               r2_src, nr2, g_part, 0, 0.0, 0.0, 1);
            QKX_HIP(hipGetLastError());
            return QKX_SUCCESS;
        });
    });
    E.j++;
    return rc;
}

int read_slot(const Engine &E, CGSlot &h)
{
    QKX_HIP(hipMemcpyAsync(g_slots_host, E.cur(), sizeof(CGSlot), hipMemcpyDeviceToHost, qkx::g_stream));
    QKX_HIP(hipStreamSynchronize(qkx::g_stream));
    h = *g_slots_host;
    return QKX_SUCCESS;
}

int write_slot(const Engine &E, const CGSlot &h)
{
    *g_slots_host = h;
    QKX_HIP(hipMemcpyAsync(E.cur(), g_slots_host, sizeof(CGSlot), hipMemcpyHostToDevice, qkx::g_stream));
    QKX_HIP(hipStreamSynchronize(qkx::g_stream));
    return QKX_SUCCESS;
}

// xh += xl, then xl = 0, in one pass
int launch_accum(qkx_spinor_s *xh, qkx_spinor_s *xl)
{
    int n = entries(xh);
    return with_spinor(xl, [&](auto S) {
        hipLaunchKernelGGL((accum_kernel<decltype(S)>), dim3((n + 255) / 256), dim3(256), 0, qkx::g_stream,
                           make_spinor<SpinorD>(xh), S, n);
        QKX_HIP(hipGetLastError());
        return QKX_SUCCESS;
    });
}

double host_norm2(const qkx_spinor_s *f)
{
    double v = 0;
    qkx_norm2(&v, f);
    return v;
}

// ---------------------------------------------------------------------------
// Several independent systems -- each on its own lattice, links, twisted mass
// and kappa -- solved in lockstep: every CG step is one launch for all of them
// (hop_multi for the stencil, the kernels below for the update and the
// direction), each workgroup working for one instance.  Each instance keeps its
// own slot (scalars, flags, iteration count); an instance whose flag is raised
// skips its share of every launch until the host has dealt with it.
constexpr int kMaxInst = 8;
constexpr int kDone = 4;   // slot flag of an instance that has met its tolerance

template <class SX, class SR, class S> struct MultiCg {
    int n;
    int blk0[kMaxInst + 1];        // workgroups [blk0[i], blk0[i + 1]) serve instance i
    int nsites[kMaxInst];
    SX x[kMaxInst];
    SR r[kMaxInst];
    S p[kMaxInst], Ap[kMaxInst];
    const CGSlot *cur[kMaxInst];
    CGSlot *next[kMaxInst];
    const double *pap_part[kMaxInst];
    int npap[kMaxInst];
    double *r2_part[kMaxInst];     // written by the update, read by the direction step
    int nr2[kMaxInst];
    double stop2[kMaxInst];
    double delta2;
};

__device__ __forceinline__ int multi_instance(const int *blk0, int n)
{
    int i = 0;
    while (i + 1 < n && (int)blockIdx.x >= blk0[i + 1]) i++;
    return i;
}

template <class SX, class SR, class S>
__global__ __launch_bounds__(kThreads) void cgm_update_kernel(MultiCg<SX, SR, S> m)
{
    const int i = multi_instance(m.blk0, m.n);
    const CGSlot *cur = m.cur[i];
    if (cur->flag) return;
    using T = typename SX::Real;
    using TR = typename SR::Real;
    using TP = typename S::Real;
    const double pap = block_sum_array(m.pap_part[i], m.npap[i]);
    const double alpha = cur->r2 / pap;
    const T a = T(alpha);
    const TR ar = TR(alpha);
    const int bid = (int)blockIdx.x - m.blk0[i];
    const int s = bid * blockDim.x + threadIdx.x;
    double acc = 0;
    if (s < m.nsites[i]) {
        Cx<T> xv[12];
        Cx<TR> rv[12];
        Cx<TP> pv[12], av[12];
        m.x[i].load(s, xv);
        m.p[i].load(s, pv);
#pragma unroll
        for (int k = 0; k < 12; k++) xv[k] = {xv[k].re + a * T(pv[k].re), xv[k].im + a * T(pv[k].im)};
        m.x[i].store(s, xv);
        m.r[i].load(s, rv);
        m.Ap[i].load(s, av);
#pragma unroll
        for (int k = 0; k < 12; k++) {
            rv[k] = {rv[k].re - ar * TR(av[k].re), rv[k].im - ar * TR(av[k].im)};
            acc += (double)rv[k].re * (double)rv[k].re + (double)rv[k].im * (double)rv[k].im;
        }
        m.r[i].store(s, rv);
    }
    block_partial(acc, m.r2_part[i] + bid);
}

template <class SX, class SR, class S>
__global__ __launch_bounds__(kThreads) void cgm_dir_kernel(MultiCg<SX, SR, S> m)
{
    const int i = multi_instance(m.blk0, m.n);
    const CGSlot *cur = m.cur[i];
    CGSlot *next = m.next[i];
    const int bid = (int)blockIdx.x - m.blk0[i];
    if (cur->flag) {
        if (bid == 0 && threadIdx.x == 0) *next = *cur;
        return;
    }
    using T = typename SR::Real;
    const int s = bid * blockDim.x + threadIdx.x;
    const bool on = s < m.nsites[i];
    Cx<typename S::Real> pv[12];
    Cx<T> rv[12];
    if (on) {
        m.p[i].load(s, pv);
        m.r[i].load(s, rv);
    }
    const double r2n = block_sum_array(m.r2_part[i], m.nr2[i]);
    const double pap = block_sum_array(m.pap_part[i], m.npap[i]);
    int flag = 0;
    if (!(pap > 0.0) || !isfinite(r2n)) flag = 3;
    else if (r2n <= m.stop2[i]) flag = 2;
    else if (r2n < m.delta2 * cur->maxr2) flag = 1;
    const double beta = r2n / cur->r2;
    if (bid == 0 && threadIdx.x == 0) *next = CGSlot{r2n, cur->r2, pap, cur->maxr2, flag, cur->k + 1};
    if (flag || !on) return;
    const T b = T(beta);
    Cx<typename S::Real> out[12];
#pragma unroll
    for (int k = 0; k < 12; k++)
        out[k] = {(typename S::Real)(rv[k].re + b * T(pv[k].re)), (typename S::Real)(rv[k].im + b * T(pv[k].im))};
    m.p[i].store(s, out);
}

// per-instance pre-sums of many partials: workgroup i sums part[i][0..n[i])
// into out[i] (compensated, fixed order)
struct MultiSum {
    const double *part[kMaxInst];
    int cnt[kMaxInst];
    double *out[kMaxInst];
};
__global__ __launch_bounds__(1024) void sum_partials_multi_kernel(MultiSum m)
{
    __shared__ double lh[1024], ll[1024];
    const int i = blockIdx.x;
    DD v{0.0, 0.0};
    for (int k = threadIdx.x; k < m.cnt[i]; k += blockDim.x) v = dd_add(v, m.part[i][k]);
    lh[threadIdx.x] = v.hi;
    ll[threadIdx.x] = v.lo;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            const DD c = dd_add(DD{lh[threadIdx.x], ll[threadIdx.x]}, DD{lh[threadIdx.x + st], ll[threadIdx.x + st]});
            lh[threadIdx.x] = c.hi;
            ll[threadIdx.x] = c.lo;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) m.out[i][0] = dd_value(DD{lh[0], ll[0]});
}

// Multi-instance solver state on the device: two rows of kMaxInst slots (the
// slot parity of iteration j is j & 1), the per-instance partials of p.Ap and
// |r|^2, and their pre-sums.
struct MultiState {
    CGSlot *slots = nullptr;        // [2][kMaxInst]
    CGSlot *slots_host = nullptr;   // pinned [kMaxInst]
    double *part = nullptr;         // per instance: [cap] p.Ap, then [cap] |r|^2
    double *red = nullptr;          // [2 kMaxInst] pre-sums
    size_t cap_total = 0;
};
MultiState g_ms;

int ensure_multi(size_t part_doubles)
{
    if (!g_ms.slots) {
        QKX_HIP(hipMalloc(&g_ms.slots, 2 * kMaxInst * sizeof(CGSlot)));
        QKX_HIP(hipHostMalloc((void **)&g_ms.slots_host, kMaxInst * sizeof(CGSlot), hipHostMallocDefault));
        QKX_HIP(hipMalloc(&g_ms.red, 2 * kMaxInst * sizeof(double)));
    }
    if (part_doubles > g_ms.cap_total) {
        if (g_ms.part) (void)hipFree(g_ms.part);
        g_ms.part = nullptr;
        QKX_HIP(hipMalloc(&g_ms.part, part_doubles * sizeof(double)));
        g_ms.cap_total = part_doubles;
    }
    return QKX_SUCCESS;
}

// One instance of a multi-instance solve: its operators, fields and scalars.
struct MInst {
    const qkx_gauge_s *Uop = nullptr, *Ui = nullptr, *Um = nullptr;
    NormalOp Ah{}, Al{}, Am{};
    std::vector<std::unique_ptr<qkx::Temp>> temps;
    qkx_spinor_s *ho, *he, *lo, *le, *mo, *me, *sm, *rm, *bh, *bn, *xs, *s, *bo, *to, *r, *pp, *Ap, *xl;
    double kappa = 0, a = 0, nrm = 1, b2 = 0, stop2 = 0;
    double *pap = nullptr, *r2 = nullptr;   // partial regions
    int npap = 0, nr2 = 0, nsites = 0, nrel = 0;
    int state = 0;   // 0 running, 1 converged, 2 breakdown / non-finite
    qkx_spinor_s *T(const int L[4], int prec, int subset, bool need = true)
    {
        if (!need) return nullptr;
        temps.emplace_back(new qkx::Temp(L, prec, subset));
        return temps.back()->f;
    }
};

CGSlot *mslot(int j, int i) { return g_ms.slots + (j & 1) * kMaxInst + i; }

template <class SX, class SR, class S>
int enqueue_multi_iteration(int n, int j, MInst *I, double delta2)
{
    qkx_spinor_s *pp[kMaxInst], *Ap[kMaxInst], *lo[kMaxInst], *le[kMaxInst];
    const qkx_gauge_s *Ui[kMaxInst];
    const int *skip[kMaxInst];
    double *pap[kMaxInst];
    double c1p[kMaxInst][4], c2p[kMaxInst][4], c1m[kMaxInst][4], c2m[kMaxInst][4];
    for (int i = 0; i < n; i++) {
        pp[i] = I[i].pp;
        Ap[i] = I[i].Ap;
        lo[i] = I[i].lo;
        le[i] = I[i].le;
        Ui[i] = I[i].Ui;
        skip[i] = &mslot(j, i)->flag;
        pap[i] = I[i].pap;
        const double k2 = I[i].kappa * I[i].kappa, a = I[i].a, nr = I[i].nrm;
        const double a1[4] = {1.0 / nr, -a / nr, 0, 0}, a2[4] = {-k2, 0, 1.0, a};
        const double b1[4] = {1.0 / nr, a / nr, 0, 0}, b2[4] = {-k2, 0, 1.0, -a};
        for (int k = 0; k < 4; k++) {
            c1p[i][k] = a1[k];
            c2p[i][k] = a2[k];
            c1m[i][k] = b1[k];
            c2m[i][k] = b2[k];
        }
    }
    // M^dag M = g5 M(-s) g5 M(s), every instance in each launch; the second
    // hop of M(s) leaves each instance's |M p|^2 = p.Ap partials
    int npap[kMaxInst];
    int rc;
    if ((rc = qkx::hop_multi(n, lo, pp, nullptr, Ui, 0, 1, c1p, 0, nullptr, nullptr, skip))) return rc;
    if ((rc = qkx::hop_multi(n, le, lo, pp, Ui, 0, 0, c2p, 1, pap, npap, skip))) return rc;
    if ((rc = qkx::hop_multi(n, lo, le, nullptr, Ui, 0, 1, c1m, 0, nullptr, nullptr, skip))) return rc;
    if ((rc = qkx::hop_multi(n, Ap, lo, le, Ui, 0, 0, c2m, 1, nullptr, nullptr, skip))) return rc;
    MultiCg<SX, SR, S> m{};
    m.n = n;
    m.delta2 = delta2;
    MultiSum ps{}, rs{};
    bool pre_pap = false, pre_r2 = false;
    int tot = 0;
    for (int i = 0; i < n; i++) {
        m.blk0[i] = tot;
        m.nsites[i] = I[i].nsites;
        tot += nblk(I[i].nsites);
        m.x[i] = make_spinor<SX>(I[i].xl ? I[i].xl : I[i].xs);
        m.r[i] = make_spinor<SR>(I[i].r);
        m.p[i] = make_spinor<S>(I[i].pp);
        m.Ap[i] = make_spinor<S>(I[i].Ap);
        m.cur[i] = mslot(j, i);
        m.next[i] = mslot(j + 1, i);
        m.pap_part[i] = I[i].pap;
        m.npap[i] = npap[i];
        m.r2_part[i] = I[i].r2;
        m.nr2[i] = nblk(I[i].nsites);
        m.stop2[i] = I[i].stop2;
        ps.part[i] = I[i].pap;
        ps.cnt[i] = npap[i];
        ps.out[i] = g_ms.red + i;
        rs.part[i] = I[i].r2;
        rs.cnt[i] = m.nr2[i];
        rs.out[i] = g_ms.red + kMaxInst + i;
        if (npap[i] > kDirectPartials) pre_pap = true;
        if (m.nr2[i] > kDirectPartials) pre_r2 = true;
    }
    m.blk0[n] = tot;
    if (pre_pap) {
        hipLaunchKernelGGL(sum_partials_multi_kernel, dim3(n), dim3(1024), 0, qkx::g_stream, ps);
        for (int i = 0; i < n; i++) {
            m.pap_part[i] = g_ms.red + i;
            m.npap[i] = 1;
        }
    }
    hipLaunchKernelGGL((cgm_update_kernel<SX, SR, S>), dim3(tot), dim3(kThreads), 0, qkx::g_stream, m);
    if (pre_r2) {
        hipLaunchKernelGGL(sum_partials_multi_kernel, dim3(n), dim3(1024), 0, qkx::g_stream, rs);
        for (int i = 0; i < n; i++) {
            m.r2_part[i] = g_ms.red + kMaxInst + i;
            m.nr2[i] = 1;
        }
    }
    hipLaunchKernelGGL((cgm_dir_kernel<SX, SR, S>), dim3(tot), dim3(kThreads), 0, qkx::g_stream, m);
    QKX_HIP(hipGetLastError());
    return QKX_SUCCESS;
}

int enqueue_multi(int lp, int n, int j, MInst *I, double delta2)
{
    switch (lp) {
    case QKX_PREC_DOUBLE: return enqueue_multi_iteration<SpinorD, SpinorD, SpinorD>(n, j, I, delta2);
    case QKX_PREC_SINGLE: return enqueue_multi_iteration<SpinorF, SpinorF, SpinorF>(n, j, I, delta2);
    case QKX_PREC_HALF: return enqueue_multi_iteration<SpinorF, SpinorH, SpinorH>(n, j, I, delta2);
    case QKX_PREC_BF16: return enqueue_multi_iteration<SpinorF, SpinorF, SpinorB>(n, j, I, delta2);
    }
    qkx::set_error("qkx_solve_multi: bad inner precision");
    return QKX_ERR_ARG;
}

// (Re)start instance i from its residual r at iteration j: mode 0 p = r and
// the slot from scratch, mode 1 p = r + beta p (beta = |r|^2 / r2_old);
// writes the instance's slot of parity j + 1.
int restart_inst(MInst &I, int i, int j, int mode)
{
    const int n = I.nsites;
    int rc = with_spinor(I.r, [&](auto R) {
        hipLaunchKernelGGL((norm_partial_kernel<decltype(R)>), dim3(nblk(n)), dim3(kThreads), 0, qkx::g_stream, R, n,
                           I.r2);
        QKX_HIP(hipGetLastError());
        return QKX_SUCCESS;
    });
    if (rc) return rc;
    if (mode == 0 && (rc = qkx_spinor_copy(I.pp, I.r))) return rc;
    const double *r2_src = I.r2;
    int nr2 = nblk(n);
    reduce_if_many(r2_src, nr2, g_ms.red + kMaxInst + i);
    return with_spinor(I.pp, [&](auto P) {
        return with_spinor(I.r, [&](auto R) {
            hipLaunchKernelGGL((cg_dir_kernel<decltype(P), decltype(R), decltype(P), false>), dim3(nblk(n)),
                               dim3(kThreads), 0, qkx::g_stream, P, R, P, mode == 0 ? 0 : n, mslot(j, i),
                               mslot(j + 1, i), r2_src, nr2, I.pap, 0, 0.0, 0.0, 1);
            QKX_HIP(hipGetLastError());
            return QKX_SUCCESS;
        });
    });
}

int read_mslots(int j, int n, CGSlot *h)
{
    QKX_HIP(hipMemcpyAsync(g_ms.slots_host, mslot(j, 0), n * sizeof(CGSlot), hipMemcpyDeviceToHost, qkx::g_stream));
    QKX_HIP(hipStreamSynchronize(qkx::g_stream));
    for (int i = 0; i < n; i++) h[i] = g_ms.slots_host[i];
    return QKX_SUCCESS;
}

int write_mslots(int j, int n, const CGSlot *h)
{
    QKX_HIP(hipStreamSynchronize(qkx::g_stream));   // the pinned buffer may still feed an earlier copy
    for (int i = 0; i < n; i++) g_ms.slots_host[i] = h[i];
    QKX_HIP(hipMemcpyAsync(mslot(j, 0), g_ms.slots_host, n * sizeof(CGSlot), hipMemcpyHostToDevice, qkx::g_stream));
    return QKX_SUCCESS;
}

} // namespace

namespace qkx {
int cg_batch_size = 16;
int cg_fuse_on() { return g_cg_fuse; }
bool cg_graph_batched() { return (g_cg_graph & 2) != 0; }
}

extern "C" {

int qkx_set_cg_fuse(int on)
{
    QKX_CHECK_ARG(on >= 0 && on <= 2, "qkx_set_cg_fuse: 0, 1 or 2");
    g_cg_fuse = on;
    return QKX_SUCCESS;
}

int qkx_set_cg_graph(int mode)
{
    QKX_CHECK_ARG(mode >= 0 && mode <= 3, "qkx_set_cg_graph: mode 0-3");
    g_cg_graph = mode;
    return QKX_SUCCESS;
}


// Reference CG (lib/solvers.cpp:16-81), FP64, full lattice.  p.Ap is
// evaluated as |D p|^2, which equals <p, D^dag D p>.  With nstore > 0 the
// iterate x is copied into store[i] after every `every` iterations.
static int cg_impl(qkx_spinor *x, const qkx_spinor *b, const qkx_gauge *U, const qkx_params *p,
                   qkx_solver_stats *st, int every, qkx_spinor **store, int nstore, int *nstored)
{
    QKX_CHECK_ARG(x && b && U && p, "qkx_cg: null argument");
    QKX_CHECK_ARG(x->prec == QKX_PREC_DOUBLE && b->prec == QKX_PREC_DOUBLE && U->prec == QKX_PREC_DOUBLE,
                  "qkx_cg: FP64 fields required");
    QKX_CHECK_ARG(x->subset == QKX_SUBSET_FULL && b->subset == QKX_SUBSET_FULL, "qkx_cg: full fields required");
    if (int rc = ensure_state(b->V / kThreads + 64)) return rc;
    auto t0 = std::chrono::steady_clock::now();
    NormalOp A{U, p->kappa, 2.0 * p->kappa * p->mu * p->twistSign, 1.0 + 4.0 * p->kappa * p->kappa * p->mu * p->mu,
               false, nullptr, nullptr, nullptr};
    qkx::Temp tf(b->L, QKX_PREC_DOUBLE, QKX_SUBSET_FULL), r(b->L, QKX_PREC_DOUBLE, QKX_SUBSET_FULL),
        pp(b->L, QKX_PREC_DOUBLE, QKX_SUBSET_FULL), Ap(b->L, QKX_PREC_DOUBLE, QKX_SUBSET_FULL);
    if (!tf.f || !r.f || !pp.f || !Ap.f) return QKX_ERR_NOMEM;
    A.tmp_full = tf.f;
    int rc;
    if ((rc = qkx_spinor_copy(r.f, b))) return rc;
    double src2 = host_norm2(b);
    Engine E;
    E.stop2 = src2 * p->tol * p->tol;
    E.delta2 = 0.0;
    if ((rc = write_slot(E, CGSlot{0, 0, 0, 0, 0, 0}))) return rc;
    if ((rc = restart(E, pp.f, r.f, 0))) return rc;
    CGSlot h;
    if ((rc = read_slot(E, h))) return rc;
    int status = QKX_SUCCESS;
    const int maxIter = p->maxIter;
    if (h.r2 > E.stop2) {
        int saved = 0;
        IterGraphs graphs;
        while (true) {
            int batch = std::min(nstore > 0 ? every : qkx::cg_batch_size, maxIter - h.k);
            if (batch <= 0) break;
            if ((rc = graphs.run(E, A, x, r.f, pp.f, Ap.f, batch))) return rc;
            if ((rc = read_slot(E, h))) return rc;
            if (h.flag == 3) { status = QKX_ERR_BREAKDOWN; break; }
            if (h.flag == 2) break;
            if (saved < nstore) {
                if ((rc = qkx_spinor_copy(store[saved], x))) return rc;
                saved++;
            }
        }
        if (nstored) *nstored = saved;
        if (h.flag != 2 && status == QKX_SUCCESS) status = QKX_NOT_CONVERGED;
    }
    if (st) {
        st->iterations = h.k;
        st->reliable_updates = 0;
        st->converged = (status == QKX_SUCCESS);
        st->iterated_r2 = h.r2;
        st->true_residual = -1.0;
        st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    if (status == QKX_NOT_CONVERGED) qkx::set_error("Warning solver reached number of max iterations");
    if (status == QKX_ERR_BREAKDOWN) qkx::set_error("qkx_cg: breakdown (p.Ap <= 0 or non-finite residual)");
    return status;
}

int qkx_cg(qkx_spinor *x, const qkx_spinor *b, const qkx_gauge *U, const qkx_params *p, qkx_solver_stats *st)
{
    return cg_impl(x, b, U, p, st, 0, nullptr, 0, nullptr);
}

int qkx_cg_record(qkx_spinor *x, const qkx_spinor *b, const qkx_gauge *U, const qkx_params *p, int every,
                  qkx_spinor **store, int nstore, int *nstored, qkx_solver_stats *st)
{
    QKX_CHECK_ARG(every > 0 && nstore >= 0 && (nstore == 0 || store), "qkx_cg_record: bad arguments");
    for (int i = 0; i < nstore; i++)
        QKX_CHECK_ARG(store[i] && store[i]->subset == QKX_SUBSET_FULL && store[i]->nrhs == 1 &&
                          store[i]->L[0] == b->L[0] && store[i]->L[1] == b->L[1] && store[i]->L[2] == b->L[2] &&
                          store[i]->L[3] == b->L[3],
                      "qkx_cg_record: store fields must be single full fields on the lattice of b");
    if (nstored) *nstored = 0;
    return cg_impl(x, b, U, p, st, every, store, nstore, nstored);
}

int qkx_reduce_sum(const double *vals, int n, int mode, double *res)
{
    QKX_CHECK_ARG(vals && res && n > 0 && mode >= 0 && mode <= 2, "qkx_reduce_sum: bad arguments");
    if (int rc = qkx::ensure_init()) return rc;
    constexpr int kProbeGroups = 8;
    double *d = nullptr;
    QKX_HIP(hipMalloc(&d, sizeof(double) * ((size_t)n + kProbeGroups)));
    double *out = d + n;
    hipError_t e = hipMemcpyAsync(d, vals, sizeof(double) * n, hipMemcpyHostToDevice, qkx::g_stream);
    int rc = QKX_SUCCESS;
    if (e == hipSuccess) {
        if (mode == 0) {
            hipLaunchKernelGGL(block_sum_probe_kernel, dim3(kProbeGroups), dim3(kThreads), 0, qkx::g_stream,
                               (const double *)d, n, out);
        } else if (mode == 1) {
            hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(1024), 0, qkx::g_stream, (const double *)d, n, out);
        } else {
            rc = qkx::sum_rows_dev(d, n, 1, out);
        }
        e = hipGetLastError();
    }
    double h[kProbeGroups] = {};
    const int nout = mode == 0 ? kProbeGroups : 1;
    if (e == hipSuccess && rc == QKX_SUCCESS)
        e = hipMemcpyAsync(h, out, sizeof(double) * nout, hipMemcpyDeviceToHost, qkx::g_stream);
    if (e == hipSuccess) e = hipStreamSynchronize(qkx::g_stream);
    (void)hipFree(d);
    QKX_HIP(e);
    if (rc) return rc;
    for (int g = 1; g < nout; g++)
        if (std::memcmp(&h[g], &h[0], sizeof(double)) != 0) {
            qkx::set_error("qkx_reduce_sum: workgroups disagree");
            return QKX_ERR_HIP;
        }
    *res = h[0];
    return QKX_SUCCESS;
}

int qkx_true_residual(double *res, const qkx_spinor *x, const qkx_spinor *b, const qkx_gauge *U, const qkx_params *p)
{
    QKX_CHECK_ARG(res && x && b && U && p, "qkx_true_residual: null argument");
    QKX_CHECK_ARG(x->prec == QKX_PREC_DOUBLE && b->prec == QKX_PREC_DOUBLE && x->subset == QKX_SUBSET_FULL &&
                      b->subset == QKX_SUBSET_FULL,
                  "qkx_true_residual: full FP64 fields required");
    qkx::Temp t(b->L, QKX_PREC_DOUBLE, QKX_SUBSET_FULL);
    if (!t.f) return QKX_ERR_NOMEM;
    double a = 2.0 * p->kappa * p->mu * p->twistSign;
    int rc = qkx::hop(t.f, x, x, U, QKX_BASIS_UKQCD, 0, -1, -p->kappa, 0, 1.0, a, 0);
    if (rc) return rc;
    if ((rc = qkx::axpby_dev(t.f, b, -1.0, 1.0, nullptr, 0))) return rc;
    double r2 = host_norm2(t.f), b2 = host_norm2(b);
    *res = b2 > 0 ? std::sqrt(r2 / b2) : std::sqrt(r2);
    return QKX_SUCCESS;
}

int qkx_solve(qkx_spinor *x, const qkx_spinor *b, const qkx_gauge *U, const qkx_gauge *U_inner_arg,
              const qkx_params *p, const qkx_solve_opts *o, qkx_solver_stats *st)
{
    QKX_CHECK_ARG(x && b && U && p && o, "qkx_solve: null argument");
    QKX_CHECK_ARG(x->prec == QKX_PREC_DOUBLE && b->prec == QKX_PREC_DOUBLE && x->subset == QKX_SUBSET_FULL &&
                      b->subset == QKX_SUBSET_FULL,
                  "qkx_solve: x and b must be full FP64 fields");
    QKX_CHECK_ARG(U->prec == QKX_PREC_DOUBLE && U->recon == QKX_RECON_18, "qkx_solve: U must be FP64 recon 18");
    QKX_CHECK_ARG(o->tol > 0 && o->maxIter > 0, "qkx_solve: bad tol/maxIter");
    if (int rc = ensure_state(b->V / kThreads + 64)) return rc;
    auto t0 = std::chrono::steady_clock::now();
    const int *L = b->L;
    const bool mixed = o->inner_prec != QKX_PREC_DOUBLE;
    const int lp = o->inner_prec;
    const int sub = o->eo ? QKX_SUBSET_EVEN : QKX_SUBSET_FULL;
    int rc;

    // FP64 operator of the checks, the source and the reconstruction (and of
    // a pure FP64 solve): U, or its cached 12-real copy while U is SU(3) to
    // 1e-14; the reported true residual is always formed with U itself
    const qkx_gauge_s *Uop = U;
    if ((rc = qkx::gauge_fp64_operator(U, &Uop))) return rc;
    // inner gauge: caller's, or a cached conversion of U
    const qkx_gauge_s *Ui = Uop;
    if (mixed) {
        if (U_inner_arg) {
            Ui = U_inner_arg;
        } else {
            qkx_gauge_s *Uc = const_cast<qkx_gauge_s *>(U);
            int want_recon = o->inner_recon == QKX_RECON_18 ? QKX_RECON_18 : QKX_RECON_12;
            if (lp == QKX_PREC_HALF || lp == QKX_PREC_BF16) want_recon = QKX_RECON_12;
            if ((rc = qkx::gauge_cached(&Uc->inner, U, lp, want_recon))) return rc;
            Ui = Uc->inner;
        }
    }

    const double kappa = p->kappa, a = 2.0 * p->kappa * p->mu * p->twistSign;
    const double nrm = 1.0 + 4.0 * p->kappa * p->kappa * p->mu * p->mu;

    qkx::Temp hf(L, QKX_PREC_DOUBLE, QKX_SUBSET_FULL), ho(L, QKX_PREC_DOUBLE, QKX_SUBSET_ODD),
        he(L, QKX_PREC_DOUBLE, QKX_SUBSET_EVEN);
    NormalOp Ah{Uop, kappa, a, nrm, o->eo != 0, hf.f, ho.f, he.f};
    qkx::Temp lf(L, lp, QKX_SUBSET_FULL), lo_(L, lp, QKX_SUBSET_ODD), le(L, lp, QKX_SUBSET_EVEN);
    NormalOp Al{Ui, kappa, a, nrm, o->eo != 0, lf.f, lo_.f, le.f};

    // FP32 operator for the parts whose result is stored at FP32 (or BF16)
    // anyway: the normal-equation right-hand side and the residual of each
    // reliable update (r = Op^dag s); s itself, the true residual, stays FP64.
    // An FP16 solve forms both with its FP16 operator and needs none of it.
    const bool need_mid = mixed && lp != QKX_PREC_HALF;
    const qkx_gauge_s *Um = U;
    if (need_mid) {
        qkx_gauge_s *Uc = const_cast<qkx_gauge_s *>(U);
        if ((rc = qkx::gauge_cached(&Uc->mid, U, QKX_PREC_SINGLE, QKX_RECON_12))) return rc;
        Um = Uc->mid;
    }
    std::unique_ptr<qkx::Temp> mo, me, mf, sm, rm;   // FP32 temporaries of a mixed solve
    NormalOp Am{Um, kappa, a, nrm, o->eo != 0, nullptr, nullptr, nullptr};
    if (need_mid) {
        const int S = QKX_PREC_SINGLE;
        if (o->eo) {
            mo.reset(new qkx::Temp(L, S, QKX_SUBSET_ODD));
            me.reset(new qkx::Temp(L, S, QKX_SUBSET_EVEN));
            if (!mo->f || !me->f) return QKX_ERR_NOMEM;
            Am.tmp_odd = mo->f;
            Am.tmp_even = me->f;
        } else {
            mf.reset(new qkx::Temp(L, S, QKX_SUBSET_FULL));
            if (!mf->f) return QKX_ERR_NOMEM;
            Am.tmp_full = mf->f;
        }
        sm.reset(new qkx::Temp(L, S, sub));
        rm.reset(new qkx::Temp(L, S, sub));
        if (!sm->f || !rm->f) return QKX_ERR_NOMEM;
    }

    qkx::Temp bh(L, QKX_PREC_DOUBLE, sub), bn(L, QKX_PREC_DOUBLE, sub), xs(L, QKX_PREC_DOUBLE, sub),
        s(L, QKX_PREC_DOUBLE, sub), bo(L, QKX_PREC_DOUBLE, QKX_SUBSET_ODD),
        to(L, QKX_PREC_DOUBLE, QKX_SUBSET_ODD);
    // p and Ap in the inner storage precision; the solution increment in FP32
    // (FP64 for a pure FP64 solve); the residual in FP32 for BF16 storage
    // (8 mantissa bits cannot carry the recurrence)
    const int xp = lp == QKX_PREC_DOUBLE ? QKX_PREC_DOUBLE : QKX_PREC_SINGLE;
    const int rp = lp == QKX_PREC_BF16 ? QKX_PREC_SINGLE : lp;
    qkx::Temp r(L, rp, sub), pp(L, lp, sub), Ap(L, lp, sub), xl(L, xp, sub);
    if (!hf.f || !ho.f || !he.f || !lf.f || !lo_.f || !le.f || !bh.f || !bn.f || !xs.f || !s.f || !bo.f ||
        !to.f || !r.f || !pp.f || !Ap.f || !xl.f)
        return QKX_ERR_NOMEM;

    const double b2 = host_norm2(b);
    if (!std::isfinite(b2)) { qkx::set_error("qkx_solve: non-finite source"); return QKX_ERR_NONFINITE; }
    if (!(b2 > 0)) {
        qkx_spinor_zero(x);
        if (st) *st = qkx_solver_stats{0, 0, 1, 0.0, 0.0, 0.0};
        return QKX_SUCCESS;
    }

    if (o->eo) {
        // bh_e = b_e + kappa H_eo A^-1 b_o
        if ((rc = qkx_spinor_copy(bh.f, b)) || (rc = qkx_spinor_copy(bo.f, b))) return rc;
        if ((rc = qkx::site(to.f, bo.f, nullptr, QKX_BASIS_UKQCD, -1, 1.0 / nrm, -a / nrm, 0, 0, 0))) return rc;
        if ((rc = qkx::hop(bh.f, to.f, bh.f, Uop, QKX_BASIS_UKQCD, 0, 0, kappa, 0, 1.0, 0, 0))) return rc;
    } else {
        if ((rc = qkx_spinor_copy(bh.f, b))) return rc;
    }
    // normal-equation rhs bn = Op^dag bh (in the inner storage for an FP16
    // solve, FP32 for the other mixed solves: it only seeds r and the first
    // stopping guess); x = 0; r = bn; p = r
    if (r.f->prec == QKX_PREC_HALF) {
        if ((rc = qkx_spinor_copy(Ap.f, bh.f)) || (rc = Al.apply_dag(r.f, Ap.f))) return rc;
    } else {
        if (mixed) {
            if ((rc = qkx_spinor_copy(sm->f, bh.f)) || (rc = Am.apply_dag(rm->f, sm->f)) ||
                (rc = qkx_spinor_copy(bn.f, rm->f)))
                return rc;
        } else if ((rc = Ah.apply_dag(bn.f, bh.f))) {
            return rc;
        }
        if ((rc = qkx_spinor_copy(r.f, bn.f))) return rc;
    }
    if ((rc = qkx_spinor_zero(xs.f)) || (rc = qkx_spinor_zero(xl.f))) return rc;

    const double tol2 = o->tol * o->tol;
    const double bn2 = host_norm2(r.f->prec == QKX_PREC_HALF ? r.f : bn.f);
    Engine E;
    E.delta2 = mixed ? o->delta * o->delta : 0.0;
    E.stop2 = tol2 * bn2;   // first guess, tightened when the true residual disagrees
    if ((rc = write_slot(E, CGSlot{0, 0, 0, 0, 0, 0}))) return rc;
    if ((rc = restart(E, pp.f, r.f, 0))) return rc;
    CGSlot h;
    if ((rc = read_slot(E, h))) return rc;

    int status = QKX_NOT_CONVERGED;
    int nrel = 0;
    qkx_spinor_s *xiter = mixed ? xl.f : xs.f;
    int k_prev = -1;
    double r2_prev = 0;
    IterGraphs graphs;
    while (h.k < o->maxIter) {
        // Enqueue about as many iterations as the residual needs to reach the
        // next event (stop or reliable update) at its recent rate of decrease:
        // launches past an event exit at once but still cost a dispatch each.
        int batch = std::min(qkx::cg_batch_size, o->maxIter - h.k);
        if (k_prev >= 0 && h.k > k_prev && h.r2 > 0 && r2_prev > h.r2 && qkx::cg_batch_size > 2) {
            const double rate = std::log(h.r2 / r2_prev) / (h.k - k_prev);   // < 0
            const double thr = std::max(E.stop2, E.delta2 * h.maxr2);
            if (thr < h.r2) {
                const double need = std::log(thr / h.r2) / rate;
                batch = std::min(batch, std::max(2, (int)std::ceil(need) + 1));
            }
        }
        k_prev = h.k;
        r2_prev = h.r2;
        if ((rc = graphs.run(E, Al, xiter, r.f, pp.f, Ap.f, batch))) return rc;
        if ((rc = read_slot(E, h))) return rc;
        if (h.flag == 0) continue;
        if (h.flag == 3) {
            qkx::set_error("qkx_solve: breakdown (p.Ap <= 0 or non-finite residual)");
            status = QKX_ERR_BREAKDOWN;
            break;
        }
        // flag 1 (reliable update due) or 2 (iterated residual small): check in FP64
        if (mixed) {
            if ((rc = launch_accum(xs.f, xl.f))) return rc;
        }
        // s = bh - Op x ; true relative residual |s|/|b|.  For an FP16 r the
        // same pass also rounds s into Ap, where the reliable update takes it
        // from (Ap is free between iterations).
        if ((rc = Ah.apply_op(s.f, xs.f))) return rc;
        const bool s_to_ap = mixed && r.f->prec == QKX_PREC_HALF && Ap.f->prec == QKX_PREC_HALF &&
                             Ap.f->nsites == s.f->nsites && Ap.f->subset == s.f->subset;
        double s2 = 0.0;
        if ((rc = qkx::resid_norm(s.f, bh.f, s_to_ap ? Ap.f : nullptr, &s2))) return rc;
        if (!std::isfinite(s2)) { status = QKX_ERR_NONFINITE; qkx::set_error("qkx_solve: non-finite residual"); break; }
        if (s2 <= tol2 * b2) { status = QKX_SUCCESS; break; }
        if (h.flag == 2) {
            // the iterated criterion fired before the true residual was met
            E.stop2 = std::min(E.stop2, h.r2) * std::max(1e-4, 0.5 * tol2 * b2 / s2);
        }
        if (mixed) {
            // reliable update: r = Op^dag s, s the FP64 true residual, into the
            // low-precision r; p = r + beta p and the slot restart on the device.
            // An FP16 r is formed by the FP16 operator from s rounded to FP16
            // (Ap is free between iterations): its rounding is that of the FP16
            // storage, and it is cheaper than the FP32 operator.
            nrel++;
            if (r.f->prec == QKX_PREC_HALF) {
                if ((!s_to_ap && (rc = qkx_spinor_copy(Ap.f, s.f))) || (rc = Al.apply_dag(r.f, Ap.f))) return rc;
            } else if ((rc = qkx_spinor_copy(sm->f, s.f))) {
                return rc;
            } else if (r.f->prec == QKX_PREC_SINGLE) {
                if ((rc = Am.apply_dag(r.f, sm->f))) return rc;
            } else if ((rc = Am.apply_dag(rm->f, sm->f)) || (rc = qkx_spinor_copy(r.f, rm->f))) {
                return rc;
            }
            if ((rc = restart(E, pp.f, r.f, 1))) return rc;
        } else {
            // continue the same recurrence: p = r + beta p with the iterated r2
            if ((rc = restart(E, pp.f, r.f, 1))) return rc;
        }
        if ((rc = read_slot(E, h))) return rc;
    }
    if (status == QKX_NOT_CONVERGED) {
        if (mixed) launch_accum(xs.f, xl.f);
        qkx::set_error("qkx_solve: maxIter reached");
    }

    // assemble the full solution
    if (o->eo) {
        // x_o = A^-1 (b_o + kappa H_oe x_e)
        if ((rc = qkx::hop(to.f, xs.f, bo.f, Uop, QKX_BASIS_UKQCD, 0, 1, kappa / nrm, -a * kappa / nrm, 1.0 / nrm,
                           -a / nrm, 0)))
            return rc;
        if ((rc = qkx_spinor_copy(x, xs.f)) || (rc = qkx_spinor_copy(x, to.f))) return rc;
    } else {
        if ((rc = qkx_spinor_copy(x, xs.f))) return rc;
    }
    double tr = -1;
    if ((rc = qkx_true_residual(&tr, x, b, U, p))) return rc;
    if (st) {
        st->iterations = h.k;
        st->reliable_updates = nrel;
        st->converged = status == QKX_SUCCESS;
        st->iterated_r2 = h.r2;
        st->true_residual = tr;
        st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return status;
}

int qkx_solve_multi(int n, qkx_spinor *const *x, const qkx_spinor *const *b, const qkx_gauge *const *U,
                    const qkx_params *p, const qkx_solve_opts *o, qkx_solver_stats *st)
{
    QKX_CHECK_ARG(n >= 1 && n <= kMaxInst, "qkx_solve_multi: 1 to 8 instances");
    QKX_CHECK_ARG(x && b && U && p && o, "qkx_solve_multi: null argument");
    QKX_CHECK_ARG(o->eo, "qkx_solve_multi: the even-odd system only (eo = 1)");
    QKX_CHECK_ARG(o->tol > 0 && o->maxIter > 0, "qkx_solve_multi: bad tol/maxIter");
    const int lp = o->inner_prec;
    QKX_CHECK_ARG(lp >= QKX_PREC_DOUBLE && lp <= QKX_PREC_BF16, "qkx_solve_multi: bad inner precision");
    for (int i = 0; i < n; i++) {
        QKX_CHECK_ARG(x[i] && b[i] && U[i], "qkx_solve_multi: null field");
        QKX_CHECK_ARG(x[i]->prec == QKX_PREC_DOUBLE && b[i]->prec == QKX_PREC_DOUBLE &&
                          x[i]->subset == QKX_SUBSET_FULL && b[i]->subset == QKX_SUBSET_FULL && x[i]->nrhs == 1 &&
                          b[i]->nrhs == 1,
                      "qkx_solve_multi: x and b must be single full FP64 fields");
        QKX_CHECK_ARG(U[i]->prec == QKX_PREC_DOUBLE && U[i]->recon == QKX_RECON_18,
                      "qkx_solve_multi: U must be FP64 recon 18");
        for (int m = 0; m < 4; m++)
            QKX_CHECK_ARG(x[i]->L[m] == U[i]->L[m] && b[i]->L[m] == U[i]->L[m], "qkx_solve_multi: lattices differ");
    }
    auto t0 = std::chrono::steady_clock::now();
    const bool mixed = lp != QKX_PREC_DOUBLE;
    const int xp = mixed ? QKX_PREC_SINGLE : QKX_PREC_DOUBLE;
    const int rp = lp == QKX_PREC_BF16 ? QKX_PREC_SINGLE : lp;
    const double tol2 = o->tol * o->tol;
    int rc;
    std::vector<MInst> I(n);
    size_t part_need = 0;
    for (int i = 0; i < n; i++) part_need += 2 * (size_t)(nblk(b[i]->Vh) + 16);
    if ((rc = ensure_multi(part_need))) return rc;
    size_t off = 0;
    for (int i = 0; i < n; i++) {
        MInst &M = I[i];
        const int *L = b[i]->L;
        M.nsites = b[i]->Vh;
        M.pap = g_ms.part + off;
        off += nblk(M.nsites) + 16;
        M.r2 = g_ms.part + off;
        off += nblk(M.nsites) + 16;
        if ((rc = qkx::gauge_fp64_operator(U[i], &M.Uop))) return rc;
        M.Ui = M.Uop;
        M.Um = U[i];
        qkx_gauge_s *Uc = const_cast<qkx_gauge_s *>(U[i]);
        // FP16 instances form their normal-equation source and reliable-update
        // residuals with the FP16 operator; the others use an FP32 one
        const bool need_mid = mixed && lp != QKX_PREC_HALF;
        if (mixed) {
            if ((rc = qkx::gauge_cached(&Uc->inner, U[i], lp, QKX_RECON_12))) return rc;
            M.Ui = Uc->inner;
        }
        if (need_mid) {
            if ((rc = qkx::gauge_cached(&Uc->mid, U[i], QKX_PREC_SINGLE, QKX_RECON_12))) return rc;
            M.Um = Uc->mid;
        }
        M.kappa = p[i].kappa;
        M.a = 2.0 * p[i].kappa * p[i].mu * p[i].twistSign;
        M.nrm = 1.0 + 4.0 * p[i].kappa * p[i].kappa * p[i].mu * p[i].mu;
        const int D = QKX_PREC_DOUBLE, E = QKX_SUBSET_EVEN, O = QKX_SUBSET_ODD, F32 = QKX_PREC_SINGLE;
        M.ho = M.T(L, D, O);
        M.he = M.T(L, D, E);
        M.lo = mixed ? M.T(L, lp, O) : M.ho;
        M.le = mixed ? M.T(L, lp, E) : M.he;
        M.mo = M.T(L, F32, O, need_mid);
        M.me = M.T(L, F32, E, need_mid);
        M.sm = M.T(L, F32, E, need_mid);
        M.rm = M.T(L, F32, E, need_mid);
        M.bh = M.T(L, D, E);
        M.bn = M.T(L, D, E, rp != QKX_PREC_HALF);
        M.xs = M.T(L, D, E);
        M.s = M.T(L, D, E);
        M.bo = M.T(L, D, O);
        M.to = M.T(L, D, O);
        M.r = M.T(L, rp, E);
        M.pp = M.T(L, lp, E);
        M.Ap = M.T(L, lp, E);
        M.xl = M.T(L, xp, E, mixed);
        for (auto &t : M.temps)
            if (!t->f) return QKX_ERR_NOMEM;
        M.Ah = NormalOp{M.Uop, M.kappa, M.a, M.nrm, true, nullptr, M.ho, M.he};
        M.Al = NormalOp{M.Ui, M.kappa, M.a, M.nrm, true, nullptr, M.lo, M.le};
        M.Am = NormalOp{M.Um, M.kappa, M.a, M.nrm, true, nullptr, M.mo, M.me};
        M.b2 = host_norm2(b[i]);
        if (!std::isfinite(M.b2)) {
            qkx::set_error("qkx_solve_multi: non-finite source");
            return QKX_ERR_NONFINITE;
        }
        // bh_e = b_e + kappa H_eo A^-1 b_o; normal-equation source bn = M^dag bh
        if ((rc = qkx_spinor_copy(M.bh, b[i])) || (rc = qkx_spinor_copy(M.bo, b[i]))) return rc;
        if ((rc = qkx::site(M.to, M.bo, nullptr, QKX_BASIS_UKQCD, -1, 1.0 / M.nrm, -M.a / M.nrm, 0, 0, 0))) return rc;
        if ((rc = qkx::hop(M.bh, M.to, M.bh, M.Uop, QKX_BASIS_UKQCD, 0, 0, M.kappa, 0, 1.0, 0, 0))) return rc;
        if (rp == QKX_PREC_HALF) {
            if ((rc = qkx_spinor_copy(M.Ap, M.bh)) || (rc = M.Al.apply_dag(M.r, M.Ap))) return rc;
        } else {
            if (mixed) {
                if ((rc = qkx_spinor_copy(M.sm, M.bh)) || (rc = M.Am.apply_dag(M.rm, M.sm)) ||
                    (rc = qkx_spinor_copy(M.bn, M.rm)))
                    return rc;
            } else if ((rc = M.Ah.apply_dag(M.bn, M.bh))) {
                return rc;
            }
            if ((rc = qkx_spinor_copy(M.r, M.bn))) return rc;
        }
        if ((rc = qkx_spinor_zero(M.xs)) || (M.xl && (rc = qkx_spinor_zero(M.xl)))) return rc;
        M.stop2 = tol2 * host_norm2(rp == QKX_PREC_HALF ? M.r : M.bn);
        if (!(M.b2 > 0)) M.state = 1;   // zero source: x = 0
    }
    // slots: parity 0 from scratch, every instance restarted into parity 1
    std::vector<CGSlot> h(n, CGSlot{0, 0, 0, 0, 0, 0});
    if ((rc = write_mslots(0, n, h.data()))) return rc;
    for (int i = 0; i < n; i++)
        if ((rc = restart_inst(I[i], i, 0, 0))) return rc;
    int j = 1;
    if ((rc = read_mslots(j, n, h.data()))) return rc;
    for (int i = 0; i < n; i++)
        if (I[i].state) h[i].flag = kDone;
    if ((rc = write_mslots(j, n, h.data()))) return rc;
    const double delta2 = mixed ? o->delta * o->delta : 0.0;
    auto running = [&]() {
        for (int i = 0; i < n; i++)
            if (!I[i].state && h[i].k < o->maxIter) return true;
        return false;
    };
    while (running()) {
        int kmax = 0;
        for (int i = 0; i < n; i++)
            if (!I[i].state) kmax = std::max(kmax, h[i].k);
        const int batch = std::max(1, std::min(qkx::cg_batch_size, o->maxIter - kmax));
        for (int it = 0; it < batch; it++, j++)
            if ((rc = enqueue_multi(lp, n, j, I.data(), delta2))) return rc;
        if ((rc = read_mslots(j, n, h.data()))) return rc;
        std::vector<int> restart_mode(n, -1);
        bool touched = false;
        for (int i = 0; i < n; i++) {
            MInst &M = I[i];
            if (M.state || h[i].flag == 0 || h[i].flag == kDone) continue;
            touched = true;
            if (h[i].flag == 3) {
                M.state = 2;
                h[i].flag = kDone;
                continue;
            }
            // reliable update or iterated stop: the true residual in FP64
            if (mixed) {
                if ((rc = launch_accum(M.xs, M.xl))) return rc;
            }
            if ((rc = M.Ah.apply_op(M.s, M.xs))) return rc;
            // s = bh - Op x and |s|^2 in one pass, s also rounded into an FP16 Ap
            const bool s_to_ap = mixed && M.r->prec == QKX_PREC_HALF && M.Ap->prec == QKX_PREC_HALF &&
                                 M.Ap->nsites == M.s->nsites && M.Ap->subset == M.s->subset;
            double s2 = 0.0;
            if ((rc = qkx::resid_norm(M.s, M.bh, s_to_ap ? M.Ap : nullptr, &s2))) return rc;
            if (!std::isfinite(s2)) {
                M.state = 2;
                h[i].flag = kDone;
                continue;
            }
            if (s2 <= tol2 * M.b2) {
                M.state = 1;
                h[i].flag = kDone;
                continue;
            }
            if (h[i].flag == 2) M.stop2 = std::min(M.stop2, h[i].r2) * std::max(1e-4, 0.5 * tol2 * M.b2 / s2);
            if (mixed) {
                // as in qkx_solve: an FP16 r from the FP16 operator on s rounded to FP16
                M.nrel++;
                if (M.r->prec == QKX_PREC_HALF) {
                    if ((!s_to_ap && (rc = qkx_spinor_copy(M.Ap, M.s))) || (rc = M.Al.apply_dag(M.r, M.Ap))) return rc;
                } else if ((rc = qkx_spinor_copy(M.sm, M.s))) {
                    return rc;
                } else if (M.r->prec == QKX_PREC_SINGLE) {
                    if ((rc = M.Am.apply_dag(M.r, M.sm))) return rc;
                } else if ((rc = M.Am.apply_dag(M.rm, M.sm)) || (rc = qkx_spinor_copy(M.r, M.rm))) {
                    return rc;
                }
            }
            restart_mode[i] = 1;
        }
        if (!touched) continue;
        // next parity: the host's view of every instance (flags cleared for
        // the restarted ones, done for the finished ones), then the restarts
        // overwrite their own slots
        std::vector<CGSlot> nx(h);
        for (int i = 0; i < n; i++)
            if (restart_mode[i] >= 0) nx[i].flag = 0;
        if ((rc = write_mslots(j + 1, n, nx.data()))) return rc;
        for (int i = 0; i < n; i++)
            if (restart_mode[i] >= 0 && (rc = restart_inst(I[i], i, j, restart_mode[i]))) return rc;
        j++;
        if ((rc = read_mslots(j, n, h.data()))) return rc;
    }
    int status = QKX_SUCCESS;
    for (int i = 0; i < n; i++) {
        MInst &M = I[i];
        if (M.state == 0 && mixed) {   // maxIter: keep the increments accumulated so far
            if ((rc = launch_accum(M.xs, M.xl))) return rc;
        }
        // x_o = A^-1 (b_o + kappa H_oe x_e)
        if ((rc = qkx::hop(M.to, M.xs, M.bo, M.Uop, QKX_BASIS_UKQCD, 0, 1, M.kappa / M.nrm, -M.a * M.kappa / M.nrm,
                           1.0 / M.nrm, -M.a / M.nrm, 0)))
            return rc;
        if ((rc = qkx_spinor_copy(x[i], M.xs)) || (rc = qkx_spinor_copy(x[i], M.to))) return rc;
        double tr = -1;
        if ((rc = qkx_true_residual(&tr, x[i], b[i], U[i], &p[i]))) return rc;
        const int code = M.state == 1 ? QKX_SUCCESS : M.state == 2 ? QKX_ERR_BREAKDOWN : QKX_NOT_CONVERGED;
        if (code == QKX_ERR_BREAKDOWN) status = QKX_ERR_BREAKDOWN;
        else if (code == QKX_NOT_CONVERGED && status == QKX_SUCCESS) status = QKX_NOT_CONVERGED;
        if (st) {
            st[i].iterations = h[i].k;
            st[i].reliable_updates = M.nrel;
            st[i].converged = M.state == 1;
            st[i].iterated_r2 = h[i].r2;
            st[i].true_residual = tr;
            st[i].fallback_iterations = 0;
        }
    }
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (st)
        for (int i = 0; i < n; i++) st[i].seconds = secs;
    if (status == QKX_ERR_BREAKDOWN) qkx::set_error("qkx_solve_multi: breakdown (p.Ap <= 0 or non-finite residual)");
    else if (status == QKX_NOT_CONVERGED) qkx::set_error("qkx_solve_multi: maxIter reached");
    return status;
}

} // extern "C"
