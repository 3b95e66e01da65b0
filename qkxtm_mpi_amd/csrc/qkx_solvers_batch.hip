// qkx_solvers_batch.hip -- CG for a batch of right-hand sides in lockstep.
//
// The reference solves the twelve spin-colour columns of a propagator one
// after another with the same gauge field (bin/test_invert_prop.cpp:83-97).
// Here the columns (or several twisted masses mu on one gauge field) are one
// batched field with the right-hand side innermost, so every stencil launch
// serves all of them and reads each link once per site, and every CG launch
// updates all of them.  Each right-hand side keeps its own CG scalars; one
// that meets its stopping test is frozen (alpha = 0) while the others go on.
// Reliable updates and true-residual checks are done for the whole batch.
//
// Same launch structure and reduction scheme as qkx_solvers.hip: per-workgroup
// partials (one per right-hand side), added by every workgroup of the consuming
// launch in a fixed order, so all workgroups take identical decisions.
#include "qkx_internal.h"
#include "qkx_core.h"
#include <chrono>
#include <cmath>
#include <string>
#include <vector>

namespace {

constexpr int kT = 256;
constexpr int NR = QKX_MAX_RHS;

struct BSlot {
    double r2[NR];      // iterated |r|^2
    double r2_old[NR];
    double pAp[NR];
    double maxr2[NR];   // |r|^2 at the last reliable update
    double stop2[NR];   // iterated stopping threshold per right-hand side
    int active[NR];     // 0: converged (true residual met), frozen
    int flag;           // 0 run, 1 reliable update due, 2 all below stop2, 3 breakdown
    int k;
};

BSlot *g_bslots = nullptr, *g_bslots_host = nullptr;
double *g_bpart = nullptr;
int g_bpart_cap = 0;
double *g_bres = nullptr, *g_bres_host = nullptr;
double *g_bred = nullptr;   // [0, NR) p.Ap, [NR, 2 NR) |r|^2, pre-summed when the partials are many
double *g_bstage = nullptr; // first-stage sums of many partials: kStageRows rows of NR (hi, lo) pairs
constexpr int kStageRows = 256;
constexpr int kDirectPartials = 2048;   // workgroups' partials read directly by each consumer

int ensure_bstate(int nblocks)
{
    if (!g_bslots) {
        QKX_HIP(hipMalloc(&g_bslots, 2 * sizeof(BSlot)));
        QKX_HIP(hipHostMalloc((void **)&g_bslots_host, sizeof(BSlot), hipHostMallocDefault));
        QKX_HIP(hipMalloc(&g_bres, NR * sizeof(double)));
        QKX_HIP(hipHostMalloc((void **)&g_bres_host, NR * sizeof(double), hipHostMallocDefault));
        QKX_HIP(hipMalloc(&g_bred, 2 * NR * sizeof(double)));
        QKX_HIP(hipMalloc(&g_bstage, 2 * kStageRows * NR * sizeof(double)));   // (hi, lo) pairs
    }
    size_t need = (size_t)nblocks * NR;
    if ((int)need > g_bpart_cap) {
        if (g_bpart) (void)hipFree(g_bpart);
        g_bpart = nullptr;
        QKX_HIP(hipMalloc(&g_bpart, sizeof(double) * 2 * need));
        g_bpart_cap = (int)need;
    }
    return QKX_SUCCESS;
}

// sums[r] = sum_b part[b*nr + r] for r < nr, compensated (double-double), in
// a fixed order; every thread of the workgroup sees the same values.
__device__ void block_sums_rhs(const double *part, int nblocks, int nr, double *sums)
{
    __shared__ double lh[kT], ll[kT];
    const int groups = blockDim.x / nr;
    const int r = threadIdx.x % nr, g = threadIdx.x / nr;
    DD s{0.0, 0.0};
    if (g < groups)
        for (int b = g; b < nblocks; b += groups) s = dd_add(s, part[(size_t)b * nr + r]);
    lh[threadIdx.x] = s.hi;
    ll[threadIdx.x] = s.lo;
    __syncthreads();
    if ((int)threadIdx.x < nr) {
        DD t{0.0, 0.0};
        for (int gg = 0; gg < groups; gg++) t = dd_add(t, DD{lh[gg * nr + threadIdx.x], ll[gg * nr + threadIdx.x]});
        sums[threadIdx.x] = dd_value(t);
    }
    __syncthreads();
}

// per-workgroup, per-rhs partial of v, summed in thread order (fixed, whatever
// order the waves run in).  The thread's unit is blockIdx.x * blockDim.x +
// threadIdx.x and its right-hand side that unit mod nr.
__device__ void block_partial_rhs(double v, bool valid, int nr, double *dst)
{
    __shared__ double vals[kT];
    vals[threadIdx.x] = valid ? v : 0.0;
    __syncthreads();
    if ((int)threadIdx.x < nr) {
        const int u0 = (int)(((size_t)blockIdx.x * blockDim.x) % (size_t)nr);
        double s = 0;
        for (int t = ((int)threadIdx.x - u0 + nr) % nr; t < (int)blockDim.x; t += nr) s += vals[t];
        dst[(size_t)blockIdx.x * nr + threadIdx.x] = s;
    }
}

__device__ __forceinline__ bool frozen(const BSlot *c, int r) { return !c->active[r] || c->r2[r] <= c->stop2[r]; }

template <class SX, class SR, class S>
__global__ __launch_bounds__(kT) void cgb_update_kernel(SX x, SR r, S p, S Ap, int n, int nr, const BSlot *cur,
                                                        const double *pap_part, int npap, double *r2_part)
{
    if (cur->flag) return;
    __shared__ double pap[NR];
    block_sums_rhs(pap_part, npap, nr, pap);
    using T = typename SX::Real;
    using TR = typename SR::Real;
    using TP = typename S::Real;
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    const int rr = u % nr;
    const double alpha = frozen(cur, rr) ? 0.0 : cur->r2[rr] / pap[rr];
    double acc = 0;
    const bool valid = u < n;
    if (valid) {
        const T a = T(alpha);
        const TR ar = TR(alpha);
        Cx<T> xv[12];
        Cx<TR> rv[12];
        Cx<TP> pv[12], av[12];
        x.load(u, xv);
        p.load(u, pv);
#pragma unroll
        for (int k = 0; k < 12; k++) xv[k] = {xv[k].re + a * T(pv[k].re), xv[k].im + a * T(pv[k].im)};
        x.store(u, xv);
        r.load(u, rv);
        Ap.load(u, av);
#pragma unroll
        for (int k = 0; k < 12; k++) {
            rv[k] = {rv[k].re - ar * TR(av[k].re), rv[k].im - ar * TR(av[k].im)};
            acc += (double)rv[k].re * (double)rv[k].re + (double)rv[k].im * (double)rv[k].im;
        }
        r.store(u, rv);
    }
    block_partial_rhs(acc, valid, nr, r2_part);
}

// mode 0: iteration; mode 1: restart after a reliable update (beta from r2_old).
// HASX: in mode 0 also x += alpha p with the p of the iteration (before its
// update), alpha as the fused hop forms it (0 for a frozen right-hand side),
// so x gets the same bits as when the hop updates it.
template <class S, class SR, class SX, bool HASX>
__global__ __launch_bounds__(kT) void cgb_dir_kernel(S p, SR r, SX x, int n, int nr, const BSlot *cur, BSlot *next,
                                                     const double *r2_part, int nr2, const double *pap_part, int npap,
                                                     double delta2, int mode)
{
    if (mode == 0 && cur->flag) {
        if (blockIdx.x == 0 && threadIdx.x == 0) *next = *cur;
        return;
    }
    __shared__ double r2n[NR], pap[NR];
    __shared__ int flag_s;
    block_sums_rhs(r2_part, nr2, nr, r2n);
    if (mode == 0) block_sums_rhs(pap_part, npap, nr, pap);
    if (threadIdx.x == 0) {
        int flag = 0;
        if (mode == 0) {
            bool all_frozen = true, any_rel = false, bad = false;
            for (int q = 0; q < nr; q++) {
                if (frozen(cur, q)) continue;
                if (!(pap[q] > 0.0) || !isfinite(r2n[q])) bad = true;
                if (cur->active[q] && r2n[q] > cur->stop2[q]) all_frozen = false;
                if (r2n[q] < delta2 * cur->maxr2[q]) any_rel = true;
            }
            flag = bad ? 3 : all_frozen ? 2 : any_rel ? 1 : 0;
        } else {
            for (int q = 0; q < nr; q++)
                if (!isfinite(r2n[q])) flag = 3;
        }
        flag_s = flag;
    }
    __syncthreads();
    const int flag = flag_s;
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    const int rr = u % nr;
    Cx<typename S::Real> pv[12];
    const bool xupd = HASX && mode == 0 && u < n;
    if (xupd) {
        using TX = typename SX::Real;
        const TX ax = TX(frozen(cur, rr) ? 0.0 : cur->r2[rr] / pap[rr]);
        Cx<TX> xv[12];
        x.load(u, xv);
        p.load(u, pv);
#pragma unroll
        for (int k = 0; k < 12; k++) xv[k] = {xv[k].re + ax * TX(pv[k].re), xv[k].im + ax * TX(pv[k].im)};
        x.store(u, xv);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        BSlot nx = *cur;
        for (int q = 0; q < nr; q++) {
            nx.r2[q] = r2n[q];
            if (mode == 0) {
                nx.r2_old[q] = frozen(cur, q) ? 0.0 : cur->r2[q];   // 0: restart from p = r
                nx.pAp[q] = pap[q];
            } else {
                nx.maxr2[q] = r2n[q];
            }
        }
        nx.flag = flag;
        if (mode == 0) nx.k = cur->k + 1;
        *next = nx;
    }
    if (flag) return;
    if (u >= n) return;
    double beta;
    if (mode == 0) beta = frozen(cur, rr) ? 0.0 : r2n[rr] / cur->r2[rr];
    else beta = cur->r2_old[rr] > 0 ? r2n[rr] / cur->r2_old[rr] : 0.0;
    using T = typename SR::Real;
    const T b = T(beta);
    Cx<T> rv[12];
    if (!xupd) p.load(u, pv);
    r.load(u, rv);
    Cx<typename S::Real> out[12];
#pragma unroll
    for (int k = 0; k < 12; k++)
        out[k] = {(typename S::Real)(rv[k].re + b * T(pv[k].re)), (typename S::Real)(rv[k].im + b * T(pv[k].im))};
    p.store(u, out);
}

template <class S>
__global__ __launch_bounds__(kT) void normb_partial_kernel(S r, int n, int nr, double *part)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    double acc = 0;
    const bool valid = u < n;
    if (valid) {
        Cx<typename S::Real> a[12];
        r.load(u, a);
#pragma unroll
        for (int k = 0; k < 12; k++) acc += (double)a[k].re * (double)a[k].re + (double)a[k].im * (double)a[k].im;
    }
    block_partial_rhs(acc, valid, nr, part);
}

// out[r] = sum_b part[b*nr + r], compensated, fixed order, one workgroup of
// 1024.  dd_in: the rows are (hi, lo) pairs left by sum_rhs_chunks_kernel.
__global__ __launch_bounds__(1024) void sum_rhs_kernel(const double *part, int nblocks, int nr, double *out,
                                                       int dd_in)
{
    __shared__ double lh[1024], ll[1024];
    const int groups = blockDim.x / nr;
    const int r = threadIdx.x % nr, g = threadIdx.x / nr;
    DD v{0.0, 0.0};
    if (g < groups)
        for (int b = g; b < nblocks; b += groups) {
            const size_t i = (size_t)b * nr + r;
            v = dd_in ? dd_add(v, DD{part[2 * i], part[2 * i + 1]}) : dd_add(v, part[i]);
        }
    lh[threadIdx.x] = v.hi;
    ll[threadIdx.x] = v.lo;
    __syncthreads();
    if ((int)threadIdx.x < nr) {
        DD t{0.0, 0.0};
        for (int gg = 0; gg < groups; gg++) t = dd_add(t, DD{lh[gg * nr + threadIdx.x], ll[gg * nr + threadIdx.x]});
        out[threadIdx.x] = dd_value(t);
    }
}

// First stage of a many-row sum: workgroup g sums rows [g chunk, (g+1) chunk)
// of part, compensated, into the (hi, lo) pair out[2 (g*nr + r)], in a fixed
// order (strided per thread, then the thread groups in order).
__global__ __launch_bounds__(kT) void sum_rhs_chunks_kernel(const double *part, int nblocks, int nr, int chunk,
                                                            double *out)
{
    __shared__ double lh[kT], ll[kT];
    const int groups = blockDim.x / nr;
    const int r = threadIdx.x % nr, g = threadIdx.x / nr;
    const int b0 = blockIdx.x * chunk, b1 = min(nblocks, b0 + chunk);
    DD v{0.0, 0.0};
    if (g < groups)
        for (int b = b0 + g; b < b1; b += groups) v = dd_add(v, part[(size_t)b * nr + r]);
    lh[threadIdx.x] = v.hi;
    ll[threadIdx.x] = v.lo;
    __syncthreads();
    if ((int)threadIdx.x < nr) {
        DD t{0.0, 0.0};
        for (int gg = 0; gg < groups; gg++) t = dd_add(t, DD{lh[gg * nr + threadIdx.x], ll[gg * nr + threadIdx.x]});
        const size_t i = (size_t)blockIdx.x * nr + threadIdx.x;
        out[2 * i] = t.hi;
        out[2 * i + 1] = t.lo;
    }
}

template <class SL>
__global__ __launch_bounds__(256) void accumb_kernel(SpinorD xh, SL xl, int n, int zero)
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
    if (zero) {
        // xl restarts from zero (stored as a zero field stores it)
#pragma unroll
        for (int k = 0; k < 12; k++) b[k] = {0, 0};
        xl.store(s, b);
    }
}

// The true-residual step of a batched reliable update in one pass: s = b - s
// (as axpby_kernel forms it with a = -1, b = 1), per-workgroup per-RHS |s|^2
// partials in normb_partial_kernel's grid and order, and, with H, s rounded
// into the 16-bit field h as spinor_convert rounds it.
template <class SH, bool H>
__global__ __launch_bounds__(kT) void residb_kernel(SpinorD s_, SpinorD b_, SH h, int n, int nr, double *part)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    double acc = 0;
    const bool valid = u < n;
    if (valid) {
        Cx<double> a[12], b[12];
        s_.load(u, a);
        b_.load(u, b);
#pragma unroll
        for (int k = 0; k < 12; k++) a[k] = {-1.0 * a[k].re + 1.0 * b[k].re, -1.0 * a[k].im + 1.0 * b[k].im};
        s_.store(u, a);
        if constexpr (H) {
            Cx<typename SH::Real> w[12];
#pragma unroll
            for (int k = 0; k < 12; k++) w[k] = {(typename SH::Real)a[k].re, (typename SH::Real)a[k].im};
            h.store(u, w);
        }
#pragma unroll
        for (int k = 0; k < 12; k++) acc += a[k].re * a[k].re + a[k].im * a[k].im;
    }
    block_partial_rhs(acc, valid, nr, part);
}

int nblk(int n) { return (n + kT - 1) / kT; }

// dst[r] = sum_b part[b*nr + r] in a fixed order: a single workgroup for a
// few hundred rows; beyond that a first stage of up to kStageRows workgroups
// over contiguous chunks (one workgroup reading every row is bound by its
// load latency: 64 us for 16k rows of 12).
void sum_rows(const double *part, int nblocks, int nr, double *dst)
{
    if (nblocks > kStageRows) {
        const int chunk = (nblocks + kStageRows - 1) / kStageRows;
        const int g = (nblocks + chunk - 1) / chunk;
        hipLaunchKernelGGL(sum_rhs_chunks_kernel, dim3(g), dim3(kT), 0, qkx::g_stream, part, nblocks, nr, chunk,
                           g_bstage);
        hipLaunchKernelGGL(sum_rhs_kernel, dim3(1), dim3(1024), 0, qkx::g_stream, (const double *)g_bstage, g, nr,
                           dst, 1);
        return;
    }
    hipLaunchKernelGGL(sum_rhs_kernel, dim3(1), dim3(1024), 0, qkx::g_stream, part, nblocks, nr, dst, 0);
}

// per-rhs partials (part, nblocks) as consumers should read them: pre-summed
// into dst when there are many
void reduce_if_many(const double *&part, int &nblocks, int nr, double *dst)
{
    if (nblocks <= kDirectPartials) return;
    sum_rows(part, nblocks, nr, dst);
    part = dst;
    nblocks = 1;
}

template <class F> int with_spinor(const qkx_spinor_s *f, F fn)
{
    return dispatch_prec(f->prec, [&](auto tag) { return fn(make_spinor<decltype(tag)>(f)); });
}

// per-rhs squared norms of a batched field, to the host
// s = b - s and out[r] = |s_r|^2 (the bits norms() gives) in one pass over
// FP64 fields; with h (FP16, same sites and RHS) also h = s rounded
int resid_norms(qkx_spinor_s *s, const qkx_spinor_s *b, qkx_spinor_s *h, std::vector<double> &out)
{
    const int n = entries(s), nr = s->nrhs, nb = nblk(n);
    if (h)
        hipLaunchKernelGGL((residb_kernel<SpinorH, true>), dim3(nb), dim3(kT), 0, qkx::g_stream,
                           make_spinor<SpinorD>(s), make_spinor<SpinorD>(b), make_spinor<SpinorH>(h), n, nr,
                           g_bpart + g_bpart_cap);
    else
        hipLaunchKernelGGL((residb_kernel<SpinorD, false>), dim3(nb), dim3(kT), 0, qkx::g_stream,
                           make_spinor<SpinorD>(s), make_spinor<SpinorD>(b), make_spinor<SpinorD>(s), n, nr,
                           g_bpart + g_bpart_cap);
    QKX_HIP(hipGetLastError());
    sum_rows(g_bpart + g_bpart_cap, nb, nr, g_bres);
    QKX_HIP(hipGetLastError());
    QKX_HIP(hipMemcpyAsync(g_bres_host, g_bres, nr * sizeof(double), hipMemcpyDeviceToHost, qkx::g_stream));
    QKX_HIP(hipStreamSynchronize(qkx::g_stream));
    out.assign(g_bres_host, g_bres_host + nr);
    return QKX_SUCCESS;
}

int norms(const qkx_spinor_s *f, std::vector<double> &out)
{
    const int n = entries(f), nr = f->nrhs, nb = nblk(n);
    int rc = with_spinor(f, [&](auto S) {
        hipLaunchKernelGGL((normb_partial_kernel<decltype(S)>), dim3(nb), dim3(kT), 0, qkx::g_stream, S, n, nr,
                           g_bpart + g_bpart_cap);
        QKX_HIP(hipGetLastError());
        return QKX_SUCCESS;
    });
    if (rc) return rc;
    sum_rows(g_bpart + g_bpart_cap, nb, nr, g_bres);
    QKX_HIP(hipGetLastError());
    QKX_HIP(hipMemcpyAsync(g_bres_host, g_bres, nr * sizeof(double), hipMemcpyDeviceToHost, qkx::g_stream));
    QKX_HIP(hipStreamSynchronize(qkx::g_stream));
    out.assign(g_bres_host, g_bres_host + nr);
    return QKX_SUCCESS;
}

// Batched operator with per-rhs twist a_r and A-normalisation n_r
struct BatchOp {
    const qkx_gauge_s *U;
    double kappa;
    double a[NR], nrm[NR];
    int nr;
    bool eo;
    qkx_spinor_s *tmp_full, *tmp_odd, *tmp_even;

    // Coefficients of the two hops of the Schur operator (tm_xx,
    // lib/lattice_util.cpp:2001-2022) with twist sign sg: the odd
    // intermediate A^-1 H_oe in, then A in - kappa^2 H_eo (that)
    void schur_coefs(double sg, double c1[NR][4], double c2[NR][4]) const
    {
        for (int r = 0; r < NR; r++) {
            const int q = r < nr ? r : 0;
            const double ar = sg * a[q];
            c1[r][0] = 1.0 / nrm[q];
            c1[r][1] = -ar / nrm[q];
            c1[r][2] = 0;
            c1[r][3] = 0;
            c2[r][0] = -kappa * kappa;
            c2[r][1] = 0;
            c2[r][2] = 1.0;
            c2[r][3] = ar;
        }
    }
    int schur(qkx_spinor_s *out, const qkx_spinor_s *in, double sg, int g5out, double *part = nullptr,
              int *np = nullptr) const
    {
        double c1[NR][4], c2[NR][4];
        schur_coefs(sg, c1, c2);
        int rc = qkx::hop_batch(tmp_odd, in, nullptr, U, QKX_BASIS_UKQCD, 0, 1, c1, 0);
        if (rc) return rc;
        return qkx::hop_batch(out, tmp_odd, in, U, QKX_BASIS_UKQCD, 0, 0, c2, g5out, part, np);
    }
    // The same with the CG update fused into the last hop (Ap is not stored)
    int schur_cg(const qkx_spinor_s *in, double sg, int g5out, const qkx::CgBatchFuse &f, int *np) const
    {
        double c1[NR][4], c2[NR][4];
        schur_coefs(sg, c1, c2);
        int rc = qkx::hop_batch(tmp_odd, in, nullptr, U, QKX_BASIS_UKQCD, 0, 1, c1, 0);
        if (rc) return rc;
        return qkx::hop_cg_batch(tmp_odd, in, U, 0, 0, c2, g5out, f, np);
    }
    // full-lattice D (dag = 0) or D^dag (dag = 1), aux = in
    int dirac(qkx_spinor_s *out, const qkx_spinor_s *in, int dag, double *part = nullptr, int *np = nullptr) const
    {
        double c[NR][4];
        for (int r = 0; r < NR; r++) {
            const int q = r < nr ? r : 0;
            c[r][0] = -kappa;
            c[r][1] = 0;
            c[r][2] = 1.0;
            c[r][3] = dag ? -a[q] : a[q];
        }
        return qkx::hop_batch(out, in, in, U, QKX_BASIS_UKQCD, dag, -1, c, 0, part, np);
    }
    int apply(qkx_spinor_s *out, const qkx_spinor_s *in, double *part, int *np) const
    {
        if (!eo) {
            int rc = dirac(tmp_full, in, 0, part, np);
            if (rc) return rc;
            return dirac(out, tmp_full, 1);
        }
        int rc = schur(tmp_even, in, +1, 1, part, np);
        if (rc) return rc;
        return schur(out, tmp_even, -1, 1);
    }
    int apply_op(qkx_spinor_s *out, const qkx_spinor_s *in) const
    {
        return eo ? schur(out, in, +1, 0) : dirac(out, in, 0);
    }
    int apply_dag(qkx_spinor_s *out, const qkx_spinor_s *in) const
    {
        if (!eo) return dirac(out, in, 1);
        int rc = qkx::site(tmp_even, in, nullptr, QKX_BASIS_UKQCD, -1, 1.0, 0, 0, 0, 1);
        if (rc) return rc;
        return schur(out, tmp_even, -1, 1);
    }
};

struct BEngine {
    int j = 0;
    double delta2 = 0;
    BSlot *cur() const { return g_bslots + (j & 1); }
    BSlot *next() const { return g_bslots + ((j + 1) & 1); }
};

int read_bslot(const BEngine &E, BSlot &h)
{
    QKX_HIP(hipMemcpyAsync(g_bslots_host, E.cur(), sizeof(BSlot), hipMemcpyDeviceToHost, qkx::g_stream));
    QKX_HIP(hipStreamSynchronize(qkx::g_stream));
    h = *g_bslots_host;
    return QKX_SUCCESS;
}

int write_bslot(const BEngine &E, const BSlot &h)
{
    *g_bslots_host = h;
    QKX_HIP(hipMemcpyAsync(E.cur(), g_bslots_host, sizeof(BSlot), hipMemcpyHostToDevice, qkx::g_stream));
    QKX_HIP(hipStreamSynchronize(qkx::g_stream));
    return QKX_SUCCESS;
}

template <class SX, class SR, class S>
int enqueue_bblas(BEngine &E, qkx_spinor_s *x, qkx_spinor_s *r, qkx_spinor_s *p, qkx_spinor_s *Ap, int npap)
{
    const int n = entries(p), nr = p->nrhs;
    const double *pap_part = g_bpart;
    double *r2_part = g_bpart + g_bpart_cap;
    SX X = make_spinor<SX>(x);
    SR R = make_spinor<SR>(r);
    S P = make_spinor<S>(p), AP = make_spinor<S>(Ap);
    reduce_if_many(pap_part, npap, nr, g_bred);
    hipLaunchKernelGGL((cgb_update_kernel<SX, SR, S>), dim3(nblk(n)), dim3(kT), 0, qkx::g_stream, X, R, P, AP, n, nr,
                       E.cur(), pap_part, npap, r2_part);
    const double *r2_src = r2_part;
    int nr2 = nblk(n);
    reduce_if_many(r2_src, nr2, nr, g_bred + NR);
    hipLaunchKernelGGL((cgb_dir_kernel<S, SR, S, false>), dim3(nblk(n)), dim3(kT), 0, qkx::g_stream, P, R, P, n, nr,
                       E.cur(), E.next(), r2_src, nr2, pap_part, npap, E.delta2, 0);
    QKX_HIP(hipGetLastError());
    return QKX_SUCCESS;
}

// One iteration with the update fused into the last hop (even-odd FP16
// batches of 3-16 RHS that share links): p.Ap from the partials of the first Schur
// application, summed to one row before the fused hop reads it.
int enqueue_biter_fused(BEngine &E, const BatchOp &A, qkx_spinor_s *x, qkx_spinor_s *r, qkx_spinor_s *p)
{
    const int n = entries(p), nr = p->nrhs;
    qkx::set_hop_skip(&E.cur()->flag);
    int npap = 0;
    int rc = A.schur(A.tmp_even, p, +1, 1, g_bpart, &npap);
    if (rc) {
        qkx::set_hop_skip(nullptr);
        return rc;
    }
    sum_rows(g_bpart, npap, nr, g_bred);
    double *r2_part = g_bpart + g_bpart_cap;
    BSlot *cur = E.cur();
    qkx::CgBatchFuse f{x, r, p, cur->r2, cur->stop2, cur->active, g_bred, r2_part};
    // qkx_set_cg_fuse(2): x += alpha p in the direction kernel, which reads p anyway
    f.x_in_dir = qkx::cg_fuse_on() == 2;
    int nr2 = 0;
    rc = A.schur_cg(A.tmp_even, -1, 1, f, &nr2);
    qkx::set_hop_skip(nullptr);
    if (rc) return rc;
    const double *r2_src = r2_part;
    reduce_if_many(r2_src, nr2, nr, g_bred + NR);
    const SpinorH P = make_spinor<SpinorH>(p), R = make_spinor<SpinorH>(r);
    if (f.x_in_dir)
        hipLaunchKernelGGL((cgb_dir_kernel<SpinorH, SpinorH, SpinorF, true>), dim3(nblk(n)), dim3(kT), 0,
                           qkx::g_stream, P, R, make_spinor<SpinorF>(x), n, nr, E.cur(), E.next(), r2_src, nr2,
                           (const double *)g_bred, 1, E.delta2, 0);
    else
        hipLaunchKernelGGL((cgb_dir_kernel<SpinorH, SpinorH, SpinorH, false>), dim3(nblk(n)), dim3(kT), 0,
                           qkx::g_stream, P, R, P, n, nr, E.cur(), E.next(), r2_src, nr2, (const double *)g_bred, 1,
                           E.delta2, 0);
    QKX_HIP(hipGetLastError());
    E.j++;
    return QKX_SUCCESS;
}

// the update and direction kernels of one iteration, for the inner precisions
int enqueue_bblas_any(BEngine &E, qkx_spinor_s *x, qkx_spinor_s *r, qkx_spinor_s *p, qkx_spinor_s *Ap, int npap)
{
    const int px = x->prec, pr = r->prec, pp = p->prec;
    if (px == QKX_PREC_DOUBLE && pr == QKX_PREC_DOUBLE && pp == QKX_PREC_DOUBLE)
        return enqueue_bblas<SpinorD, SpinorD, SpinorD>(E, x, r, p, Ap, npap);
    if (px == QKX_PREC_SINGLE && pr == QKX_PREC_SINGLE && pp == QKX_PREC_SINGLE)
        return enqueue_bblas<SpinorF, SpinorF, SpinorF>(E, x, r, p, Ap, npap);
    if (px == QKX_PREC_SINGLE && pr == QKX_PREC_HALF && pp == QKX_PREC_HALF)
        return enqueue_bblas<SpinorF, SpinorH, SpinorH>(E, x, r, p, Ap, npap);
    if (px == QKX_PREC_SINGLE && pr == QKX_PREC_SINGLE && pp == QKX_PREC_BF16)
        return enqueue_bblas<SpinorF, SpinorF, SpinorB>(E, x, r, p, Ap, npap);
    qkx::set_error("batched cg: unsupported precision combination");
    return QKX_ERR_ARG;
}

int enqueue_biter(BEngine &E, const BatchOp &A, qkx_spinor_s *x, qkx_spinor_s *r, qkx_spinor_s *p,
                  qkx_spinor_s *Ap)
{
    if (A.eo && qkx::cg_fuse_on() && p->prec == QKX_PREC_HALF && qkx::hop_cg_batch_available(p, A.U, x, r))
        return enqueue_biter_fused(E, A, x, r, p);
    qkx::set_hop_skip(&E.cur()->flag);
    int npap = 0;
    int rc = A.apply(Ap, p, g_bpart, &npap);
    qkx::set_hop_skip(nullptr);
    if (rc) return rc;
    rc = enqueue_bblas_any(E, x, r, p, Ap, npap);
    E.j++;
    return rc;
}

// The learned right preconditioner of the even-odd system: the inner CG
// runs on (Mh N)^dag (Mh N) y = (Mh N)^dag r with N = N_e the network on the
// even sites, and x = N y.  While `on`, every inner operator, the iterate's
// accumulation and the residual of each reliable update go through N; once
// switched off (the fallback) the solver is the plain CG.
struct NetPre {
    qkx_net_s *net = nullptr;
    const qkx_gauge_s *Un = nullptr;   // the network's FP32 links
    bool on = false;
    qkx_spinor_s *t = nullptr, *w = nullptr;   // inner-precision even temporaries
    qkx_spinor_s *t32 = nullptr;               // FP32 even temporary: N applied to the iterate
};

// One iteration on N^dag Mh^dag Mh N: p.Ap = |Mh N p|^2 from the partials of
// the first Schur application.  The network launches follow the hops' skip
// flag, so iterations enqueued past an event cost only their dispatches.
int enqueue_biter_net(BEngine &E, const BatchOp &A, NetPre &np, qkx_spinor_s *x, qkx_spinor_s *r, qkx_spinor_s *p,
                      qkx_spinor_s *Ap)
{
    const int *skip = &E.cur()->flag;
    qkx::set_hop_skip(skip);
    int npap = 0, rc;
    if ((rc = qkx::net_apply_eo(np.net, np.t, p, np.Un, 0, skip)) ||
        (rc = A.schur(A.tmp_even, np.t, +1, 1, g_bpart, &npap)) || (rc = A.schur(np.w, A.tmp_even, -1, 1)) ||
        (rc = qkx::net_apply_eo(np.net, Ap, np.w, np.Un, 1, skip))) {
        qkx::set_hop_skip(nullptr);
        return rc;
    }
    qkx::set_hop_skip(nullptr);
    rc = enqueue_bblas_any(E, x, r, p, Ap, npap);
    E.j++;
    return rc;
}

// One batched iteration: the preconditioned form while the network is on.
int enqueue_any(BEngine &E, const BatchOp &A, NetPre *np, qkx_spinor_s *x, qkx_spinor_s *r, qkx_spinor_s *p,
                qkx_spinor_s *Ap)
{
    return (np && np->on) ? enqueue_biter_net(E, A, *np, x, r, p, Ap) : enqueue_biter(E, A, x, r, p, Ap);
}

// Pairs of batched iterations replayed from captured HIP graphs.  After two
// iterations the slot parity is back where it started, so one executable per
// starting parity serves the whole solve: every per-RHS scalar the kernels
// use (alpha, beta, the stopping thresholds and flags) lives in the device
// slots, not in kernel arguments, so nothing has to be captured again -- only
// when the iteration changes form (the network's fallback).  The captured
// work is kernels only (no memset or copy nodes).  The first iterations of a
// solve are launched one by one, so that the hop autotuner has timed every
// launch kind before a capture records its shapes.  Captures that fail (or
// instrumentation, whose events cannot be captured) fall back to launching
// the iterations one by one.
struct BIterGraphs {
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    bool broken = false;
    bool net_on = false;   // the iteration form the executables hold
    int eager = 2;         // iterations still to launch one by one
    BIterGraphs() = default;
    BIterGraphs(const BIterGraphs &) = delete;
    BIterGraphs &operator=(const BIterGraphs &) = delete;
    ~BIterGraphs() { drop(); }
    void drop()
    {
        for (auto &e : exec)
            if (e) (void)hipGraphExecDestroy(e);
        exec[0] = exec[1] = nullptr;
    }
    bool capture(BEngine &E, const BatchOp &A, NetPre *np, qkx_spinor_s *x, qkx_spinor_s *r, qkx_spinor_s *p,
                 qkx_spinor_s *Ap, int par)
    {
        if (hipStreamBeginCapture(qkx::g_stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        const int j0 = E.j;
        int rc = enqueue_any(E, A, np, x, r, p, Ap);
        if (!rc) rc = enqueue_any(E, A, np, x, r, p, Ap);
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
        return ok;
    }
    int run(BEngine &E, const BatchOp &A, NetPre *np, qkx_spinor_s *x, qkx_spinor_s *r, qkx_spinor_s *p,
            qkx_spinor_s *Ap, int n)
    {
        const bool on = np && np->on;
        if (on != net_on) {
            drop();
            net_on = on;
        }
        const bool use = qkx::cg_graph_batched() && !qkx::prof_active();
        while (n > 0) {
            const int par = E.j & 1;
            if (use && !broken && eager <= 0 && n >= 2) {
                if (!exec[par] && !capture(E, A, np, x, r, p, Ap, par)) broken = true;
                if (!broken) {
                    QKX_HIP(hipGraphLaunch(exec[par], qkx::g_stream));
                    E.j += 2;
                    n -= 2;
                    continue;
                }
            }
            if (int rc = enqueue_any(E, A, np, x, r, p, Ap)) return rc;
            eager--;
            n--;
        }
        return QKX_SUCCESS;
    }
};

int brestart(BEngine &E, qkx_spinor_s *p, qkx_spinor_s *r, int initial)
{
    const int n = entries(r), nr = r->nrhs;
    int rc = with_spinor(r, [&](auto R) {
        hipLaunchKernelGGL((normb_partial_kernel<decltype(R)>), dim3(nblk(n)), dim3(kT), 0, qkx::g_stream, R, n, nr,
                           g_bpart + g_bpart_cap);
        QKX_HIP(hipGetLastError());
        return QKX_SUCCESS;
    });
    if (rc) return rc;
    if (initial && (rc = qkx_spinor_copy(p, r))) return rc;
    const double *r2_src = g_bpart + g_bpart_cap;
    int nr2 = nblk(n);
    reduce_if_many(r2_src, nr2, nr, g_bred + NR);
    rc = with_spinor(p, [&](auto P) {
        return with_spinor(r, [&](auto R) {
            hipLaunchKernelGGL((cgb_dir_kernel<decltype(P), decltype(R), decltype(P), false>), dim3(nblk(n)),
                               dim3(kT), 0, qkx::g_stream, P, R, P, initial ? 0 : n, nr, E.cur(), E.next(), r2_src,
                               nr2, g_bpart, 0, 0.0, 1);
            QKX_HIP(hipGetLastError());
            return QKX_SUCCESS;
        });
    });
    E.j++;
    return rc;
}

} // namespace

namespace qkx {
extern int cg_batch_size;
// the batched solver's per-right-hand-side row sums (two stages beyond
// kStageRows rows), for qkx_reduce_sum
int sum_rows_dev(const double *part, int nblocks, int nr, double *dst)
{
    if (int rc = ensure_bstate(64)) return rc;
    sum_rows(part, nblocks, nr, dst);
    QKX_HIP(hipGetLastError());
    return QKX_SUCCESS;
}
}

extern "C" {

// qkx_solve_batch, and with np the learned-preconditioned solver
// (qkx_solve_learned): np->on selects the preconditioned operator; stall is
// the true-residual reduction a reliable-update cycle must reach for the
// network to be kept (otherwise the solve goes on as the plain CG).
static int solve_batch_impl(qkx_spinor *x, const qkx_spinor *b, const qkx_gauge *U, const qkx_gauge *U_inner_arg,
                            const qkx_params *p, const double *mu, const qkx_solve_opts *o, qkx_solver_stats *st,
                            NetPre *np, double stall)
{
    QKX_CHECK_ARG(x && b && U && p && o, "qkx_solve_batch: null argument");
    QKX_CHECK_ARG(x->prec == QKX_PREC_DOUBLE && b->prec == QKX_PREC_DOUBLE && x->subset == QKX_SUBSET_FULL &&
                      b->subset == QKX_SUBSET_FULL && x->nrhs == b->nrhs,
                  "qkx_solve_batch: x and b must be full FP64 batches of the same size");
    QKX_CHECK_ARG(U->prec == QKX_PREC_DOUBLE && U->recon == QKX_RECON_18, "qkx_solve_batch: U must be FP64 recon 18");
    QKX_CHECK_ARG(o->tol > 0 && o->maxIter > 0, "qkx_solve_batch: bad tol/maxIter");
    const int nr = b->nrhs;
    const int *L = b->L;
    // hop workgroups hold at least 128 (site, RHS) units (whole sites when
    // links are shared): partial capacity for that many workgroups
    if (int rc = ensure_bstate(b->V * nr / 128 + 64)) return rc;
    auto t0 = std::chrono::steady_clock::now();
    const bool mixed = o->inner_prec != QKX_PREC_DOUBLE;
    const int lp = o->inner_prec;
    const int sub = o->eo ? QKX_SUBSET_EVEN : QKX_SUBSET_FULL;
    int rc;

    // FP64 operator of the checks, the source and the reconstruction (and of
    // a pure FP64 solve): U, or its cached 12-real copy while U is SU(3) to
    // 1e-14; the reported true residuals are always formed with U itself
    const qkx_gauge_s *Uop = U;
    if ((rc = qkx::gauge_fp64_operator(U, &Uop))) return rc;
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

    // FP32 links for the parts whose result is stored at FP32 (or BF16)
    // anyway: the normal-equation right-hand side and the residual r = Op^dag s
    // of each reliable update (s itself, the true residual, stays FP64).  An
    // FP16 solve forms both with its FP16 operator and needs none of it.
    const bool need_mid = mixed && lp != QKX_PREC_HALF;
    const qkx_gauge_s *Um = U;
    if (need_mid) {
        qkx_gauge_s *Uc = const_cast<qkx_gauge_s *>(U);
        if ((rc = qkx::gauge_cached(&Uc->mid, U, QKX_PREC_SINGLE, QKX_RECON_12))) return rc;
        Um = Uc->mid;
    }

    BatchOp Ah{}, Al{}, Am{};
    Ah.U = Uop;
    Al.U = Ui;
    Am.U = Um;
    Ah.kappa = Al.kappa = Am.kappa = p->kappa;
    Ah.nr = Al.nr = Am.nr = nr;
    Ah.eo = Al.eo = Am.eo = o->eo != 0;
    for (int r = 0; r < NR; r++) {
        const double m = (mu && r < nr) ? mu[r] : p->mu;
        Ah.a[r] = Al.a[r] = Am.a[r] = 2.0 * p->kappa * m * p->twistSign;
        Ah.nrm[r] = Al.nrm[r] = Am.nrm[r] = 1.0 + 4.0 * p->kappa * p->kappa * m * m;
    }
    // temporaries; those a variant does not use are not allocated (a 12-rhs
    // FP64 full field at 48^3x96 is 24.5 GB)
    const bool eo = o->eo != 0;
    std::vector<qkx_spinor_s *> temps;
    bool nomem = false;
    auto T = [&](int prec, int subset, bool need = true) -> qkx_spinor_s * {
        if (!need) return nullptr;
        qkx_spinor_s *f = qkx::pool_get_batch(L, prec, subset, nr);
        if (!f) nomem = true;
        temps.push_back(f);
        return f;
    };
    struct Release {
        std::vector<qkx_spinor_s *> &t;
        ~Release()
        {
            for (auto f : t)
                if (f) qkx::pool_put(f);
        }
    } release{temps};
    const int xp = lp == QKX_PREC_DOUBLE ? QKX_PREC_DOUBLE : QKX_PREC_SINGLE;
    const int rp = lp == QKX_PREC_BF16 ? QKX_PREC_SINGLE : lp;
    qkx_spinor_s *hf = T(QKX_PREC_DOUBLE, QKX_SUBSET_FULL, !eo), *ho = T(QKX_PREC_DOUBLE, QKX_SUBSET_ODD, eo),
                 *he = T(QKX_PREC_DOUBLE, QKX_SUBSET_EVEN, eo), *lf = T(lp, QKX_SUBSET_FULL, !eo && lp != QKX_PREC_DOUBLE),
                 *lo = T(lp, QKX_SUBSET_ODD, eo && lp != QKX_PREC_DOUBLE),
                 *le = T(lp, QKX_SUBSET_EVEN, eo && lp != QKX_PREC_DOUBLE);
    if (lp == QKX_PREC_DOUBLE) {
        lf = hf;
        lo = ho;
        le = he;
    }
    qkx_spinor_s *bh = T(QKX_PREC_DOUBLE, sub), *bn = T(QKX_PREC_DOUBLE, sub, rp != QKX_PREC_HALF),
                 *xs = T(QKX_PREC_DOUBLE, sub),
                 *s = T(QKX_PREC_DOUBLE, sub),
                 *bo = T(QKX_PREC_DOUBLE, QKX_SUBSET_ODD, eo), *to = T(QKX_PREC_DOUBLE, QKX_SUBSET_ODD, eo),
                 *r = T(rp, sub), *pp = T(lp, sub), *Ap = T(lp, sub), *xl = T(xp, sub, mixed),
                 *tf = eo ? T(QKX_PREC_DOUBLE, QKX_SUBSET_FULL) : hf;
    const int S32 = QKX_PREC_SINGLE;
    qkx_spinor_s *mo = T(S32, QKX_SUBSET_ODD, need_mid && eo), *me = T(S32, QKX_SUBSET_EVEN, need_mid && eo),
                 *sm = T(S32, sub, need_mid), *rm = T(S32, sub, need_mid);
    if (np) {
        np->t = T(lp, QKX_SUBSET_EVEN);
        np->w = T(lp, QKX_SUBSET_EVEN);
        np->t32 = T(S32, QKX_SUBSET_EVEN);
    }
    if (nomem) return QKX_ERR_NOMEM;
    Ah.tmp_full = hf; Ah.tmp_odd = ho; Ah.tmp_even = he;
    Al.tmp_full = lf; Al.tmp_odd = lo; Al.tmp_even = le;
    Am.tmp_full = nullptr; Am.tmp_odd = mo; Am.tmp_even = me;   // Op^dag needs no full temporary
    // r_out = Op^dag s in FP32 (mixed) or FP64, s an FP64 field; r_out any
    // precision.  An FP16 r_out is formed by the FP16 operator from s rounded
    // to FP16 (Ap is free between iterations): the rounding of its storage.
    // ap_has_s: Ap already holds s rounded to FP16 (left by the residual check)
    bool ap_has_s = false;
    auto dag_residual0 = [&](qkx_spinor_s *r_out, qkx_spinor_s *src64) -> int {
        if (!mixed) return Ah.apply_dag(r_out, src64);
        int e;
        if (r_out->prec == QKX_PREC_HALF) {
            const bool have = ap_has_s && src64 == s;
            ap_has_s = false;
            if (!have && (e = qkx_spinor_copy(Ap, src64))) return e;
            return Al.apply_dag(r_out, Ap);
        }
        e = qkx_spinor_copy(sm, src64);
        if (e) return e;
        if (r_out->prec == QKX_PREC_SINGLE) return Am.apply_dag(r_out, sm);
        if ((e = Am.apply_dag(rm, sm))) return e;
        return qkx_spinor_copy(r_out, rm);
    };
    // with the network: r_out = N^dag Op^dag s
    auto dag_residual = [&](qkx_spinor_s *r_out, qkx_spinor_s *src64) -> int {
        int e = dag_residual0(r_out, src64);
        if (e || !np || !np->on) return e;
        return qkx::net_apply_eo(np->net, r_out, r_out, np->Un, 1);
    };

    std::vector<double> b2, bn2, s2;
    if ((rc = norms(b, b2))) return rc;
    for (int q = 0; q < nr; q++)
        if (!std::isfinite(b2[q])) { qkx::set_error("qkx_solve_batch: non-finite source"); return QKX_ERR_NONFINITE; }

    if (o->eo) {
        double c[NR][4];
        for (int q = 0; q < NR; q++) {
            const int z = q < nr ? q : 0;
            c[q][0] = 1.0 / Ah.nrm[z];
            c[q][1] = -Ah.a[z] / Ah.nrm[z];
            c[q][2] = c[q][3] = 0;
        }
        if ((rc = qkx_spinor_copy(bh, b)) || (rc = qkx_spinor_copy(bo, b))) return rc;
        if ((rc = qkx::site_batch(to, bo, nullptr, QKX_BASIS_UKQCD, -1, c, 0))) return rc;
        if ((rc = qkx::hop(bh, to, bh, Uop, QKX_BASIS_UKQCD, 0, 0, p->kappa, 0, 1.0, 0, 0))) return rc;
    } else {
        if ((rc = qkx_spinor_copy(bh, b))) return rc;
    }
    // normal-equation source Op^dag bh; with the network N^dag Op^dag bh.  An
    // FP16 solve forms it in r itself, with its FP16 operator.
    std::vector<double> bn2_plain;
    qkx_spinor_s *src = r->prec == QKX_PREC_HALF ? r : bn;
    if ((rc = dag_residual0(src, bh)) || (rc = norms(src, bn2_plain))) return rc;
    if (np && np->on && (rc = qkx::net_apply_eo(np->net, src, src, np->Un, 1))) return rc;
    if ((rc = qkx_spinor_zero(xs)) || (xl && (rc = qkx_spinor_zero(xl)))) return rc;
    if (src != r && (rc = qkx_spinor_copy(r, bn))) return rc;
    if ((rc = norms(src, bn2))) return rc;

    const double tol2 = o->tol * o->tol;
    BEngine E;
    E.delta2 = mixed ? o->delta * o->delta : 0.0;
    BSlot h{};
    for (int q = 0; q < nr; q++) {
        h.active[q] = b2[q] > 0 ? 1 : 0;
        h.stop2[q] = tol2 * bn2[q];
    }
    if ((rc = write_bslot(E, h))) return rc;
    if ((rc = brestart(E, pp, r, 1))) return rc;
    if ((rc = read_bslot(E, h))) return rc;

    int status = QKX_NOT_CONVERGED;
    int nrel = 0, k_fallback = -1, weak_cycles = 0;
    std::vector<double> true_res(nr, -1.0), prev_s2(nr, 0.0);
    if ((rc = norms(bh, prev_s2))) return rc;   // the first cycle's reduction is against |bh|
    qkx_spinor_s *xiter = mixed ? xl : xs;
    auto all_done = [&]() {
        for (int q = 0; q < nr; q++)
            if (h.active[q]) return false;
        return true;
    };
    if (all_done()) status = QKX_SUCCESS;
    BIterGraphs graphs;
    while (status != QKX_SUCCESS && h.k < o->maxIter) {
        int batch = std::min(qkx::cg_batch_size, o->maxIter - h.k);
        if ((rc = graphs.run(E, Al, np, xiter, r, pp, Ap, batch))) return rc;
        if ((rc = read_bslot(E, h))) return rc;
        if (h.flag == 0) continue;
        if (h.flag == 3) {
            qkx::set_error("qkx_solve_batch: breakdown (p.Ap <= 0 or non-finite residual)");
            status = QKX_ERR_BREAKDOWN;
            break;
        }
        if (mixed) {
            const int n = entries(xs);
            qkx_spinor_s *xacc = xl;
            if (np && np->on) {   // x = N y
                if ((rc = qkx::net_apply_eo(np->net, np->t32, xl, np->Un, 0))) return rc;
                xacc = np->t32;
            }
            // xl is reset in the same pass when it is what was accumulated
            rc = with_spinor(xacc, [&](auto S) {
                hipLaunchKernelGGL((accumb_kernel<decltype(S)>), dim3((n + 255) / 256), dim3(256), 0, qkx::g_stream,
                                   make_spinor<SpinorD>(xs), S, n, xacc == xl ? 1 : 0);
                QKX_HIP(hipGetLastError());
                return QKX_SUCCESS;
            });
            if (rc || (xacc != xl && (rc = qkx_spinor_zero(xl)))) return rc;
        }
        if ((rc = Ah.apply_op(s, xs))) return rc;
        // s = bh - Op x and its per-RHS norms in one pass; for an FP16 r, s is
        // also rounded into Ap, where the reliable update takes it from
        ap_has_s = mixed && r->prec == QKX_PREC_HALF && Ap->prec == QKX_PREC_HALF && Ap->nsites == s->nsites &&
                   Ap->nrhs == s->nrhs && Ap->subset == s->subset;
        if ((rc = resid_norms(s, bh, ap_has_s ? Ap : nullptr, s2))) return rc;
        bool nonfinite = false;
        for (int q = 0; q < nr; q++) {
            if (!std::isfinite(s2[q])) nonfinite = true;
            true_res[q] = std::sqrt(s2[q] / b2[q]);
            if (h.active[q] && s2[q] <= tol2 * b2[q]) h.active[q] = 0;
            else if (h.active[q] && h.r2[q] <= h.stop2[q])
                h.stop2[q] = h.stop2[q] * std::max(1e-4, 0.5 * tol2 * b2[q] / s2[q]);
        }
        if (nonfinite) { status = QKX_ERR_NONFINITE; qkx::set_error("qkx_solve_batch: non-finite residual"); break; }
        if (all_done()) { status = QKX_SUCCESS; break; }
        // The network stays while the preconditioned cycles make progress:
        // the plain CG takes over from the current solution (fresh
        // directions, the plain normal-equation stopping threshold) when a
        // cycle that ended on its reliable-update trigger cut the worst active
        // true residual by less than `stall`, when a cycle left it unchanged,
        // or after three weak cycles in a row.  (A cycle that ended on the
        // iterated stopping test near convergence may cut little and is
        // not judged alone: its threshold is tightened instead.)
        bool switched = false;
        if (np && np->on) {
            double worst = 0.0;
            for (int q = 0; q < nr; q++)
                if (h.active[q] && prev_s2[q] > 0) worst = std::max(worst, std::sqrt(s2[q] / prev_s2[q]));
            weak_cycles = worst > stall ? weak_cycles + 1 : 0;
            if ((h.flag == 1 && worst > stall) || worst > 0.99 || weak_cycles >= 3) {
                np->on = false;
                switched = true;
                k_fallback = h.k;
                for (int q = 0; q < nr; q++) h.stop2[q] = tol2 * bn2_plain[q];
            }
        }
        prev_s2 = s2;
        h.flag = 0;
        if ((rc = write_bslot(E, h))) return rc;
        if (mixed) {
            nrel++;
            if ((rc = dag_residual(r, s))) return rc;
        }
        if ((rc = brestart(E, pp, r, switched ? 1 : 0))) return rc;
        if ((rc = read_bslot(E, h))) return rc;
    }
    if (status == QKX_NOT_CONVERGED) qkx::set_error("qkx_solve_batch: maxIter reached");

    // assemble: x_o = A^-1 (b_o + kappa H_oe x_e), per right-hand side
    if (o->eo) {
        double c[NR][4];
        for (int q = 0; q < NR; q++) {
            const int z = q < nr ? q : 0;
            const double ar = Ah.a[z], nz = Ah.nrm[z];
            c[q][0] = p->kappa / nz;
            c[q][1] = -ar * p->kappa / nz;
            c[q][2] = 1.0 / nz;
            c[q][3] = -ar / nz;
        }
        if ((rc = qkx::hop_batch(to, xs, bo, Uop, QKX_BASIS_UKQCD, 0, 1, c, 0))) return rc;
        if ((rc = qkx_spinor_copy(x, xs)) || (rc = qkx_spinor_copy(x, to))) return rc;
    } else {
        if ((rc = qkx_spinor_copy(x, xs))) return rc;
    }
    // true residuals of the full system, FP64
    {
        BatchOp F = Ah;
        F.U = U;
        F.eo = false;
        if ((rc = F.dirac(tf, x, 0))) return rc;
        if ((rc = qkx::axpby_dev(tf, b, -1.0, 1.0, nullptr, 0))) return rc;
        std::vector<double> t2;
        if ((rc = norms(tf, t2))) return rc;
        for (int q = 0; q < nr; q++) true_res[q] = b2[q] > 0 ? std::sqrt(t2[q] / b2[q]) : 0.0;
    }
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (st)
        for (int q = 0; q < nr; q++) {
            st[q].iterations = h.k;
            st[q].reliable_updates = nrel;
            st[q].converged = status == QKX_SUCCESS;
            st[q].iterated_r2 = h.r2[q];
            st[q].true_residual = true_res[q];
            st[q].seconds = secs;
            st[q].fallback_iterations = k_fallback >= 0 ? h.k - k_fallback : 0;
        }
    return status;
}

int qkx_solve_batch(qkx_spinor *x, const qkx_spinor *b, const qkx_gauge *U, const qkx_gauge *U_inner_arg,
                    const qkx_params *p, const double *mu, const qkx_solve_opts *o, qkx_solver_stats *st)
{
    return solve_batch_impl(x, b, U, U_inner_arg, p, mu, o, st, nullptr, 0.0);
}

int qkx_solve_learned(qkx_spinor *x, const qkx_spinor *b, const qkx_gauge *U, const qkx_gauge *U_net, qkx_net *net,
                      const qkx_params *p, const qkx_solve_opts *o, double stall, qkx_solver_stats *st)
{
    QKX_CHECK_ARG(x && b && U && p && o, "qkx_solve_learned: null argument");
    QKX_CHECK_ARG(o->eo == 1 && o->inner_prec != QKX_PREC_DOUBLE,
                  "qkx_solve_learned: even-odd with a low-precision inner solver required");
    QKX_CHECK_ARG(stall > 0 && stall <= 1, "qkx_solve_learned: stall must be in (0, 1]");
    if (!net) return solve_batch_impl(x, b, U, nullptr, p, nullptr, o, st, nullptr, 0.0);
    QKX_CHECK_ARG(U_net && U_net->prec == QKX_PREC_SINGLE && U_net->recon == QKX_RECON_18 &&
                      U_net->L[0] == b->L[0] && U_net->L[1] == b->L[1] && U_net->L[2] == b->L[2] &&
                      U_net->L[3] == b->L[3],
                  "qkx_solve_learned: U_net must be FP32 recon 18 on the lattice of b");
    if (int rc = qkx::net_prep_eo(net)) return rc;
    NetPre np;
    np.net = net;
    np.Un = U_net;
    np.on = true;
    return solve_batch_impl(x, b, U, nullptr, p, nullptr, o, st, &np, stall);
}

} // extern "C"
