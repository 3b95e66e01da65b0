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
// The CG scalars live in device memory.  Each iteration is four or six kernel
// launches (operator, p.Ap, x/r update with |r|^2, p update); the reductions
// finish inside their own launch, so the host enqueues a batch of iterations
// and reads the scalars back once per batch.  When an update kernel raises a
// flag (converged / reliable update due / breakdown) every later kernel of the
// batch returns at entry.
#include "qkx_internal.h"
#include "qkx_core.h"
#include <chrono>
#include <cmath>
#include <vector>

namespace qkx {
void set_hop_skip(const int *flag);
int schur(qkx_spinor_s *out, const qkx_spinor_s *in, const qkx_gauge_s *U, double kappa, double a, double nrm,
          int basis, int pe, int g5out, qkx_spinor_s *tmp_other);
} // namespace qkx

namespace {

constexpr int kRedThreads = 256;
constexpr int kRedBlocks = 1024;

struct CGState {
    double r2;      // current iterated |r|^2
    double r2_old;  // previous
    double pAp;
    double maxr2;   // |r|^2 at the last reliable update
    double delta2;  // reliable-update trigger (relative |r|^2 drop)
    double stop2;   // iterated stopping threshold
    int flag;       // 0 run, 1 reliable update due, 2 iterated convergence, 3 breakdown
    int k;          // iterations done
};

CGState *g_state = nullptr;       // device
CGState *g_state_host = nullptr;  // pinned mirror

int ensure_state()
{
    if (g_state) return QKX_SUCCESS;
    QKX_HIP(hipMalloc(&g_state, sizeof(CGState)));
    QKX_HIP(hipHostMalloc((void **)&g_state_host, sizeof(CGState), hipHostMallocDefault));
    return QKX_SUCCESS;
}

struct DD {
    double hi, lo;
};
__device__ __forceinline__ DD two_sum(double a, double b)
{
    double s = a + b;
    double bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}
__device__ __forceinline__ DD dd_add(DD a, DD b)
{
    DD s = two_sum(a.hi, b.hi);
    return two_sum(s.hi, s.lo + a.lo + b.lo);
}

// Block + grid reduction of one double; returns true in thread 0 of the last
// block with the compensated total in *total.
__device__ bool grid_sum(double v, double *partial, unsigned int *counter, double *total)
{
    __shared__ double lds[kRedThreads / 64];
    __shared__ bool last;
    __shared__ double lh[kRedThreads], ll[kRedThreads];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += lds[w];
        partial[blockIdx.x] = s;
        __threadfence();
        last = (atomicAdd(counter, 1u) == gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return false;
    __threadfence();
    DD acc{0.0, 0.0};
    for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x) acc = dd_add(acc, DD{partial[b], 0.0});
    lh[threadIdx.x] = acc.hi;
    ll[threadIdx.x] = acc.lo;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            DD c = dd_add(DD{lh[threadIdx.x], ll[threadIdx.x]}, DD{lh[threadIdx.x + st], ll[threadIdx.x + st]});
            lh[threadIdx.x] = c.hi;
            ll[threadIdx.x] = c.lo;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        *total = lh[0] + ll[0];
        *counter = 0u;
        return true;
    }
    return false;
}

// st->pAp = Re <p, Ap>
template <class S>
__global__ __launch_bounds__(kRedThreads) void cg_pap_kernel(S p, S Ap, int n, CGState *st, double *partial,
                                                             unsigned int *counter)
{
    if (st->flag) return;
    double acc = 0;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
        Cx<typename S::Real> a[12], b[12];
        p.load(s, a);
        Ap.load(s, b);
#pragma unroll
        for (int k = 0; k < 12; k++) acc += (double)a[k].re * (double)b[k].re + (double)a[k].im * (double)b[k].im;
    }
    double tot;
    if (grid_sum(acc, partial, counter, &tot)) st->pAp = tot;
}

// alpha = r2/pAp; x += alpha p; r -= alpha Ap; r2_new = |r|^2; raise flags.
// x (SX) and r (SR) may be kept in a wider type than p and Ap (S): 16-bit
// storage of p/Ap with FP32 accumulation of the solution and residual.
template <class SX, class SR, class S>
__global__ __launch_bounds__(kRedThreads) void cg_update_kernel(SX x, SR r, S p, S Ap, int n, CGState *st,
                                                                double *partial, unsigned int *counter)
{
    if (st->flag) return;
    using T = typename SX::Real;
    using TR = typename SR::Real;
    using TP = typename S::Real;
    const double alpha = st->r2 / st->pAp;
    const T a = T(alpha);
    const TR ar = TR(alpha);
    double acc = 0;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
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
    double r2n;
    if (grid_sum(acc, partial, counter, &r2n)) {
        double pap = st->pAp;
        st->r2_old = st->r2;
        st->r2 = r2n;
        st->k += 1;
        if (!(pap > 0.0) || !isfinite(r2n)) st->flag = 3;
        else if (r2n <= st->stop2) st->flag = 2;
        else if (r2n < st->delta2 * st->maxr2) st->flag = 1;
    }
}

// p = r + beta p, beta = r2/r2_old
template <class S, class SR>
__global__ __launch_bounds__(256) void cg_p_kernel(S p, SR r, int n, const CGState *st, int force)
{
    if (st->flag && !force) return;
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    using T = typename SR::Real;
    const T beta = T(st->r2 / st->r2_old);
    Cx<typename S::Real> pv[12];
    Cx<T> rv[12];
    p.load(s, pv);
    r.load(s, rv);
    Cx<typename S::Real> out[12];
#pragma unroll
    for (int k = 0; k < 12; k++)
        out[k] = {(typename S::Real)(rv[k].re + beta * T(pv[k].re)), (typename S::Real)(rv[k].im + beta * T(pv[k].im))};
    p.store(s, out);
}

// st->r2 = |r|^2 (also maxr2), no flag check
template <class S>
__global__ __launch_bounds__(kRedThreads) void cg_norm_kernel(S r, int n, CGState *st, double *partial,
                                                              unsigned int *counter, int set_max)
{
    double acc = 0;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
        Cx<typename S::Real> a[12];
        r.load(s, a);
#pragma unroll
        for (int k = 0; k < 12; k++) acc += (double)a[k].re * (double)a[k].re + (double)a[k].im * (double)a[k].im;
    }
    double tot;
    if (grid_sum(acc, partial, counter, &tot)) {
        st->r2 = tot;
        if (set_max) st->maxr2 = tot;
    }
}

// xh += xl (FP64 += low precision)
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
}

int red_blocks(int n) { return std::min(kRedBlocks, (n + kRedThreads - 1) / kRedThreads); }

// ---------------------------------------------------------------------------
// Normal operator A = Op^dag Op in one precision
struct NormalOp {
    const qkx_gauge_s *U;
    double kappa, a, nrm;
    bool eo;
    qkx_spinor_s *tmp_full;  // full-lattice temp (full solve)
    qkx_spinor_s *tmp_odd;   // odd temp (eo)
    qkx_spinor_s *tmp_even;  // even temp (eo)
    // out = A in
    int apply(qkx_spinor_s *out, const qkx_spinor_s *in) const
    {
        if (!eo) {
            int rc = qkx::hop(tmp_full, in, in, U, QKX_BASIS_UKQCD, 0, -1, -kappa, 0, 1.0, a, 0);
            if (rc) return rc;
            return qkx::hop(out, tmp_full, tmp_full, U, QKX_BASIS_UKQCD, 1, -1, -kappa, 0, 1.0, -a, 0);
        }
        int rc = qkx::schur(tmp_even, in, U, kappa, a, nrm, QKX_BASIS_UKQCD, 0, 1, tmp_odd);
        if (rc) return rc;
        return qkx::schur(out, tmp_even, U, kappa, -a, nrm, QKX_BASIS_UKQCD, 0, 1, tmp_odd);
    }
    // out = Op in (the un-squared operator)
    int apply_op(qkx_spinor_s *out, const qkx_spinor_s *in) const
    {
        if (!eo) return qkx::hop(out, in, in, U, QKX_BASIS_UKQCD, 0, -1, -kappa, 0, 1.0, a, 0);
        return qkx::schur(out, in, U, kappa, a, nrm, QKX_BASIS_UKQCD, 0, 0, tmp_odd);
    }
    // out = Op^dag in
    int apply_dag(qkx_spinor_s *out, const qkx_spinor_s *in) const
    {
        if (!eo) return qkx::hop(out, in, in, U, QKX_BASIS_UKQCD, 1, -1, -kappa, 0, 1.0, -a, 0);
        // M(s)^dag = g5 M(-s) g5
        int rc = qkx::site(tmp_even, in, nullptr, QKX_BASIS_UKQCD, -1, 1.0, 0, 0, 0, 1);
        if (rc) return rc;
        return qkx::schur(out, tmp_even, U, kappa, -a, nrm, QKX_BASIS_UKQCD, 0, 1, tmp_odd);
    }
};

template <class F> int with_spinor(const qkx_spinor_s *f, F fn)
{
    return dispatch_prec(f->prec, [&](auto tag) { return fn(make_spinor<decltype(tag)>(f)); });
}

template <class SX, class SR, class S>
int enqueue_blas(qkx_spinor_s *x, qkx_spinor_s *r, qkx_spinor_s *p, qkx_spinor_s *Ap)
{
    int n = p->nsites;
    SX X = make_spinor<SX>(x);
    SR R = make_spinor<SR>(r);
    S P = make_spinor<S>(p), AP = make_spinor<S>(Ap);
    hipLaunchKernelGGL((cg_pap_kernel<S>), dim3(red_blocks(n)), dim3(kRedThreads), 0, qkx::g_stream, P, AP, n,
                       g_state, qkx::reduce_scratch(), qkx::reduce_counter());
    hipLaunchKernelGGL((cg_update_kernel<SX, SR, S>), dim3(red_blocks(n)), dim3(kRedThreads), 0, qkx::g_stream, X, R,
                       P, AP, n, g_state, qkx::reduce_scratch(), qkx::reduce_counter());
    hipLaunchKernelGGL((cg_p_kernel<S, SR>), dim3((n + 255) / 256), dim3(256), 0, qkx::g_stream, P, R, n, g_state, 0);
    QKX_HIP(hipGetLastError());
    return QKX_SUCCESS;
}

int enqueue_iteration(const NormalOp &A, qkx_spinor_s *x, qkx_spinor_s *r, qkx_spinor_s *p, qkx_spinor_s *Ap)
{
    qkx::set_hop_skip(&g_state->flag);
    int rc = A.apply(Ap, p);
    qkx::set_hop_skip(nullptr);
    if (rc) return rc;
    const int px = x->prec, pr = r->prec, pp = p->prec;
    if (px == QKX_PREC_DOUBLE && pr == QKX_PREC_DOUBLE && pp == QKX_PREC_DOUBLE)
        return enqueue_blas<SpinorD, SpinorD, SpinorD>(x, r, p, Ap);
    if (px == QKX_PREC_SINGLE && pr == QKX_PREC_SINGLE && pp == QKX_PREC_SINGLE)
        return enqueue_blas<SpinorF, SpinorF, SpinorF>(x, r, p, Ap);
    if (px == QKX_PREC_SINGLE && pr == QKX_PREC_HALF && pp == QKX_PREC_HALF)
        return enqueue_blas<SpinorF, SpinorH, SpinorH>(x, r, p, Ap);
    if (px == QKX_PREC_SINGLE && pr == QKX_PREC_SINGLE && pp == QKX_PREC_BF16)
        return enqueue_blas<SpinorF, SpinorF, SpinorB>(x, r, p, Ap);
    qkx::set_error("cg: unsupported precision combination");
    return QKX_ERR_ARG;
}

int launch_norm(const qkx_spinor_s *r, int set_max)
{
    int n = r->nsites;
    return with_spinor(r, [&](auto S) {
        hipLaunchKernelGGL((cg_norm_kernel<decltype(S)>), dim3(red_blocks(n)), dim3(kRedThreads), 0, qkx::g_stream, S,
                           n, g_state, qkx::reduce_scratch(), qkx::reduce_counter(), set_max);
        QKX_HIP(hipGetLastError());
        return QKX_SUCCESS;
    });
}

int launch_p_force(qkx_spinor_s *p, const qkx_spinor_s *r)
{
    int n = p->nsites;
    return with_spinor(p, [&](auto S) {
        return with_spinor(r, [&](auto R) {
            hipLaunchKernelGGL((cg_p_kernel<decltype(S), decltype(R)>), dim3((n + 255) / 256), dim3(256), 0,
                               qkx::g_stream, S, R, n, g_state, 1);
            QKX_HIP(hipGetLastError());
            return QKX_SUCCESS;
        });
    });
}

int launch_accum(qkx_spinor_s *xh, const qkx_spinor_s *xl)
{
    int n = xh->nsites;
    return with_spinor(xl, [&](auto S) {
        hipLaunchKernelGGL((accum_kernel<decltype(S)>), dim3((n + 255) / 256), dim3(256), 0, qkx::g_stream,
                           make_spinor<SpinorD>(xh), S, n);
        QKX_HIP(hipGetLastError());
        return QKX_SUCCESS;
    });
}

int read_state(CGState &h)
{
    QKX_HIP(hipMemcpyAsync(g_state_host, g_state, sizeof(CGState), hipMemcpyDeviceToHost, qkx::g_stream));
    QKX_HIP(hipStreamSynchronize(qkx::g_stream));
    h = *g_state_host;
    return QKX_SUCCESS;
}

int write_state(const CGState &h)
{
    *g_state_host = h;
    QKX_HIP(hipMemcpyAsync(g_state, g_state_host, sizeof(CGState), hipMemcpyHostToDevice, qkx::g_stream));
    QKX_HIP(hipStreamSynchronize(qkx::g_stream));
    return QKX_SUCCESS;
}

double host_norm2(const qkx_spinor_s *f)
{
    double v = 0;
    qkx_norm2(&v, f);
    return v;
}

} // namespace

namespace qkx {
int cg_batch_size = 16;
}

extern "C" {

// Reference CG (lib/solvers.cpp:16-81), FP64, full lattice.
int qkx_cg(qkx_spinor *x, const qkx_spinor *b, const qkx_gauge *U, const qkx_params *p, qkx_solver_stats *st)
{
    QKX_CHECK_ARG(x && b && U && p, "qkx_cg: null argument");
    QKX_CHECK_ARG(x->prec == QKX_PREC_DOUBLE && b->prec == QKX_PREC_DOUBLE && U->prec == QKX_PREC_DOUBLE,
                  "qkx_cg: FP64 fields required");
    QKX_CHECK_ARG(x->subset == QKX_SUBSET_FULL && b->subset == QKX_SUBSET_FULL, "qkx_cg: full fields required");
    if (int rc = ensure_state()) return rc;
    auto t0 = std::chrono::steady_clock::now();
    NormalOp A;
    A.U = U;
    A.kappa = p->kappa;
    A.a = 2.0 * p->kappa * p->mu * p->twistSign;
    A.nrm = 1.0 + 4.0 * p->kappa * p->kappa * p->mu * p->mu;
    A.eo = false;
    qkx::Temp tf(b->L, QKX_PREC_DOUBLE, QKX_SUBSET_FULL), r(b->L, QKX_PREC_DOUBLE, QKX_SUBSET_FULL),
        pp(b->L, QKX_PREC_DOUBLE, QKX_SUBSET_FULL), Ap(b->L, QKX_PREC_DOUBLE, QKX_SUBSET_FULL);
    if (!tf.f || !r.f || !pp.f || !Ap.f) return QKX_ERR_NOMEM;
    A.tmp_full = tf.f;
    int rc;
    if ((rc = qkx_spinor_copy(r.f, b)) || (rc = qkx_spinor_copy(pp.f, r.f))) return rc;
    double src2 = host_norm2(b);
    CGState h{};
    h.stop2 = src2 * p->tol * p->tol;
    h.delta2 = 0.0;
    if ((rc = write_state(h))) return rc;
    if ((rc = launch_norm(r.f, 1))) return rc;
    if ((rc = read_state(h))) return rc;
    int maxIter = p->maxIter;
    int status = QKX_SUCCESS;
    if (!(h.r2 > h.stop2)) {
        // the reference loop body never runs
    } else {
        while (true) {
            int batch = std::min(qkx::cg_batch_size, maxIter - h.k);
            if (batch <= 0) break;
            for (int j = 0; j < batch; j++)
                if ((rc = enqueue_iteration(A, x, r.f, pp.f, Ap.f))) return rc;
            if ((rc = read_state(h))) return rc;
            if (h.flag == 3) { status = QKX_ERR_BREAKDOWN; break; }
            if (h.flag == 2) break;
        }
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
    if (int rc = ensure_state()) return rc;
    auto t0 = std::chrono::steady_clock::now();
    const int *L = b->L;
    const bool mixed = o->inner_prec != QKX_PREC_DOUBLE;
    const int lp = o->inner_prec;
    const int sub = o->eo ? QKX_SUBSET_EVEN : QKX_SUBSET_FULL;
    int rc;

    // inner gauge
    const qkx_gauge_s *Ui = U;
    if (mixed) {
        if (U_inner_arg) {
            Ui = U_inner_arg;
        } else {
            qkx_gauge_s *Uc = const_cast<qkx_gauge_s *>(U);
            int want_recon = o->inner_recon == QKX_RECON_18 ? QKX_RECON_18 : QKX_RECON_12;
            int gp = lp;
            if (lp == QKX_PREC_HALF || lp == QKX_PREC_BF16) want_recon = QKX_RECON_12;
            if (!Uc->inner || Uc->inner->prec != gp || Uc->inner->recon != want_recon) {
                if (Uc->inner) qkx_gauge_destroy(Uc->inner);
                Uc->inner = nullptr;
                if ((rc = qkx_gauge_create(&Uc->inner, L, gp, want_recon, U->bc))) return rc;
                if ((rc = qkx_gauge_convert(Uc->inner, U))) return rc;
            }
            Ui = Uc->inner;
        }
    }

    const double kappa = p->kappa, a = 2.0 * p->kappa * p->mu * p->twistSign;
    const double nrm = 1.0 + 4.0 * p->kappa * p->kappa * p->mu * p->mu;

    // FP64 operator
    qkx::Temp hf(L, QKX_PREC_DOUBLE, QKX_SUBSET_FULL), ho(L, QKX_PREC_DOUBLE, QKX_SUBSET_ODD),
        he(L, QKX_PREC_DOUBLE, QKX_SUBSET_EVEN);
    NormalOp Ah{U, kappa, a, nrm, o->eo != 0, hf.f, ho.f, he.f};
    // low-precision operator
    qkx::Temp lf(L, lp, QKX_SUBSET_FULL), lo_(L, lp, QKX_SUBSET_ODD), le(L, lp, QKX_SUBSET_EVEN);
    NormalOp Al{Ui, kappa, a, nrm, o->eo != 0, lf.f, lo_.f, le.f};

    // system right-hand side bh (FP64) on the solve subset, and FP64 work fields
    qkx::Temp bh(L, QKX_PREC_DOUBLE, sub), bn(L, QKX_PREC_DOUBLE, sub), xs(L, QKX_PREC_DOUBLE, sub),
        s(L, QKX_PREC_DOUBLE, sub), rh(L, QKX_PREC_DOUBLE, sub), bo(L, QKX_PREC_DOUBLE, QKX_SUBSET_ODD),
        to(L, QKX_PREC_DOUBLE, QKX_SUBSET_ODD);
    // low-precision CG vectors: p and Ap in the inner storage precision; the
    // solution increment in FP32 (FP64 for a pure FP64 solve); the residual in
    // FP32 for BF16 storage (8 mantissa bits cannot carry the recurrence)
    const int xp = lp == QKX_PREC_DOUBLE ? QKX_PREC_DOUBLE : QKX_PREC_SINGLE;
    const int rp = lp == QKX_PREC_BF16 ? QKX_PREC_SINGLE : lp;
    qkx::Temp r(L, rp, sub), pp(L, lp, sub), Ap(L, lp, sub), xl(L, xp, sub);
    if (!hf.f || !ho.f || !he.f || !lf.f || !lo_.f || !le.f || !bh.f || !bn.f || !xs.f || !s.f || !rh.f || !bo.f ||
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
        if ((rc = qkx::hop(bh.f, to.f, bh.f, U, QKX_BASIS_UKQCD, 0, 0, kappa, 0, 1.0, 0, 0))) return rc;
    } else {
        if ((rc = qkx_spinor_copy(bh.f, b))) return rc;
    }
    // normal-equation rhs bn = Op^dag bh
    if ((rc = Ah.apply_dag(bn.f, bh.f))) return rc;

    // x = 0, r = bn, p = r
    if ((rc = qkx_spinor_zero(xs.f))) return rc;
    if ((rc = qkx_spinor_zero(xl.f))) return rc;
    if ((rc = qkx_spinor_copy(r.f, bn.f)) || (rc = qkx_spinor_copy(pp.f, r.f))) return rc;

    const double tol2 = o->tol * o->tol;
    double bn2 = host_norm2(bn.f);
    CGState h{};
    h.delta2 = mixed ? o->delta * o->delta : 0.0;
    h.stop2 = tol2 * bn2;   // first guess, tightened adaptively
    if ((rc = write_state(h))) return rc;
    if ((rc = launch_norm(r.f, 1))) return rc;
    if ((rc = read_state(h))) return rc;

    int status = QKX_NOT_CONVERGED;
    int nrel = 0;
    double true_res = -1.0;
    qkx_spinor_s *xiter = mixed ? xl.f : xs.f;
    while (h.k < o->maxIter) {
        int batch = std::min(qkx::cg_batch_size, o->maxIter - h.k);
        for (int j = 0; j < batch; j++)
            if ((rc = enqueue_iteration(Al, xiter, r.f, pp.f, Ap.f))) return rc;
        if ((rc = read_state(h))) return rc;
        if (h.flag == 0) continue;
        if (h.flag == 3) {
            qkx::set_error("qkx_solve: breakdown (p.Ap <= 0 or non-finite residual)");
            status = QKX_ERR_BREAKDOWN;
            break;
        }
        // flag 1 (reliable update due) or 2 (iterated residual small): FP64 check
        if (mixed) {
            if ((rc = launch_accum(xs.f, xl.f)) || (rc = qkx_spinor_zero(xl.f))) return rc;
        }
        // s = bh - Op x ; true relative residual |s|/|b|
        if ((rc = Ah.apply_op(s.f, xs.f))) return rc;
        if ((rc = qkx::axpby_dev(s.f, bh.f, -1.0, 1.0, nullptr, 0))) return rc;
        double s2 = host_norm2(s.f);
        true_res = std::sqrt(s2 / b2);
        if (!std::isfinite(s2)) { status = QKX_ERR_NONFINITE; qkx::set_error("qkx_solve: non-finite residual"); break; }
        if (s2 <= tol2 * b2) { status = QKX_SUCCESS; break; }
        if (mixed) {
            // reliable update: r = Op^dag s in FP64, back into the low-precision r
            nrel++;
            if ((rc = Ah.apply_dag(rh.f, s.f))) return rc;
            if ((rc = qkx_spinor_copy(r.f, rh.f))) return rc;
            double r2_old = h.r2_old;
            if ((rc = launch_norm(r.f, 1))) return rc;   // st.r2 = st.maxr2 = |r|^2
            if ((rc = read_state(h))) return rc;
            h.r2_old = r2_old;
        }
        // the iterated criterion fired too early: tighten it
        if (h.flag == 2 || h.r2 <= h.stop2) h.stop2 = std::min(h.stop2, h.r2) * std::max(1e-4, 0.5 * tol2 * b2 / s2);
        h.flag = 0;
        if ((rc = write_state(h))) return rc;
        // p = r + beta p with beta = r2/r2_old
        if ((rc = launch_p_force(pp.f, r.f))) return rc;
    }
    if (status == QKX_NOT_CONVERGED) {
        if (mixed) { launch_accum(xs.f, xl.f); qkx_spinor_zero(xl.f); }
        qkx::set_error("qkx_solve: maxIter reached");
    }

    // assemble the full solution
    if (o->eo) {
        // x_o = A^-1 (b_o + kappa H_oe x_e)
        if ((rc = qkx::hop(to.f, xs.f, bo.f, U, QKX_BASIS_UKQCD, 0, 1, kappa / nrm, -a * kappa / nrm, 1.0 / nrm,
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
    (void)true_res;
    return status;
}

} // extern "C"
