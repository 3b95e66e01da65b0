// qkx_gcr.hip -- flexible GCR on D x = b (full lattice, FP64 vectors) with an
// optional learned right preconditioner z = M r, and a plain-CG fallback.
//
// No reference counterpart (the reference has CG only, lib/solvers.cpp).
// Restarted GCR(m): per iteration z_i = M r, w_i = D z_i, w_i and z_i made
// orthogonal (in w) to the previous directions of the cycle, then
// x += alpha z_i, r -= alpha w_i with alpha = <w_i, r>/<w_i, w_i>.  Because z
// is formed anew each iteration the method tolerates any (even changing)
// preconditioner.  All coefficients stay on the device; the host reads |r|^2
// once per iteration.  A cycle that reduces |r| by less than the stall factor
// hands the residual to the even-odd mixed-precision CG (qkx_solve), which
// judges the final residual as always in FP64.
#include "qkx_internal.h"
#include "qkx_core.h"
#include <chrono>
#include <cmath>
#include <vector>

namespace qkx {
int net_apply(qkx_net_s *n, qkx_spinor_s *out, const qkx_spinor_s *in, const qkx_gauge_s *U);
}

namespace {

constexpr int kT = 256;
constexpr int MAXM = 16;

struct VecSet {
    const double2 *p[MAXM];
};

__device__ __forceinline__ double2 cj_mul(double2 a, double2 b)   // conj(a) b
{
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 c_mul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// sum the NV doubles of every thread over the workgroup; thread j < NV gets total j
template <int NV> __device__ void block_sum(double v[NV], double *out)
{
    __shared__ double red[kT / 64][NV];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < NV; j++) {
        double a = v[j];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
        if (lane == 0) red[wv][j] = a;
    }
    __syncthreads();
    if ((int)threadIdx.x < NV) {
        double s = 0;
        for (int q = 0; q < (int)(blockDim.x >> 6); q++) s += red[q][threadIdx.x];
        out[threadIdx.x] = s;
    }
}

// c_j = <W_j, v> = sum conj(W_j) v for j < nw; per-workgroup partials [block][MAXM][2]
__global__ __launch_bounds__(kT) void mdot_kernel(VecSet W, int nw, const double2 *__restrict__ v, int V,
                                                  double *part)
{
    double acc[2 * MAXM];
#pragma unroll
    for (int j = 0; j < 2 * MAXM; j++) acc[j] = 0;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < V; s += gridDim.x * blockDim.x) {
#pragma unroll
        for (int k = 0; k < 12; k++) {
            const double2 b = v[(size_t)k * V + s];
#pragma unroll
            for (int j = 0; j < MAXM; j++)
                if (j < nw) {
                    const double2 z = cj_mul(W.p[j][(size_t)k * V + s], b);
                    acc[2 * j] += z.x;
                    acc[2 * j + 1] += z.y;
                }
        }
    }
    __shared__ double tot[2 * MAXM];
    block_sum<2 * MAXM>(acc, tot);
    if ((int)threadIdx.x < 2 * MAXM) part[(size_t)blockIdx.x * 2 * MAXM + threadIdx.x] = tot[threadIdx.x];
}

// beta_j = c_j / |W_j|^2 (fixed-order sum over the workgroup partials)
__global__ void beta_kernel(const double *part, int nb, int nw, const double *wn2, double2 *beta)
{
    const int j = threadIdx.x;
    if (j >= nw) return;
    DD re{0.0, 0.0}, im{0.0, 0.0};
    for (int b = 0; b < nb; b++) {
        re = dd_add(re, part[(size_t)b * 2 * MAXM + 2 * j]);
        im = dd_add(im, part[(size_t)b * 2 * MAXM + 2 * j + 1]);
    }
    beta[j] = make_double2(dd_value(re) / wn2[j], dd_value(im) / wn2[j]);
}

// w -= sum beta_j W_j, z -= sum beta_j Z_j; partials of |w|^2 and <w, r>
__global__ __launch_bounds__(kT) void orth_kernel(double2 *w, double2 *z, VecSet W, VecSet Z, int nw,
                                                  const double2 *beta, const double2 *__restrict__ r, int V,
                                                  double *part)
{
    double acc[3] = {0, 0, 0};
    __shared__ double2 bs[MAXM];
    if ((int)threadIdx.x < nw) bs[threadIdx.x] = beta[threadIdx.x];
    __syncthreads();
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < V; s += gridDim.x * blockDim.x) {
#pragma unroll
        for (int k = 0; k < 12; k++) {
            const size_t i = (size_t)k * V + s;
            double2 wv = w[i], zv = z[i];
#pragma unroll
            for (int j = 0; j < MAXM; j++)
                if (j < nw) {
                    const double2 bw = c_mul(bs[j], W.p[j][i]);
                    const double2 bz = c_mul(bs[j], Z.p[j][i]);
                    wv.x -= bw.x;
                    wv.y -= bw.y;
                    zv.x -= bz.x;
                    zv.y -= bz.y;
                }
            w[i] = wv;
            z[i] = zv;
            const double2 c = cj_mul(wv, r[i]);
            acc[0] += wv.x * wv.x + wv.y * wv.y;
            acc[1] += c.x;
            acc[2] += c.y;
        }
    }
    __shared__ double tot[3];
    block_sum<3>(acc, tot);
    if (threadIdx.x < 3) part[(size_t)blockIdx.x * 3 + threadIdx.x] = tot[threadIdx.x];
}

__global__ void alpha_kernel(const double *part, int nb, double *wn2_i, double2 *alpha)
{
    if (threadIdx.x != 0) return;
    DD dn{0.0, 0.0}, dre{0.0, 0.0}, dim{0.0, 0.0};
    for (int b = 0; b < nb; b++) {
        dn = dd_add(dn, part[(size_t)b * 3]);
        dre = dd_add(dre, part[(size_t)b * 3 + 1]);
        dim = dd_add(dim, part[(size_t)b * 3 + 2]);
    }
    const double n = dd_value(dn), re = dd_value(dre), im = dd_value(dim);
    *wn2_i = n;
    *alpha = n > 0 ? make_double2(re / n, im / n) : make_double2(0.0, 0.0);
}

// x += alpha z, r -= alpha w; partials of |r|^2
__global__ __launch_bounds__(kT) void xr_kernel(double2 *x, double2 *r, const double2 *__restrict__ z,
                                                const double2 *__restrict__ w, const double2 *alpha, int V,
                                                double *part)
{
    const double2 a = *alpha;
    double acc[1] = {0};
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < V; s += gridDim.x * blockDim.x) {
#pragma unroll
        for (int k = 0; k < 12; k++) {
            const size_t i = (size_t)k * V + s;
            const double2 az = c_mul(a, z[i]), aw = c_mul(a, w[i]);
            double2 xv = x[i], rv = r[i];
            xv.x += az.x;
            xv.y += az.y;
            rv.x -= aw.x;
            rv.y -= aw.y;
            x[i] = xv;
            r[i] = rv;
            acc[0] += rv.x * rv.x + rv.y * rv.y;
        }
    }
    __shared__ double tot[1];
    block_sum<1>(acc, tot);
    if (threadIdx.x == 0) part[blockIdx.x] = tot[0];
}

__global__ void sum_kernel(const double *part, int nb, double *out)
{
    if (threadIdx.x != 0) return;
    DD s{0.0, 0.0};
    for (int b = 0; b < nb; b++) s = dd_add(s, part[b]);
    *out = dd_value(s);
}

struct State {
    double *part = nullptr;
    double *wn2 = nullptr;       // [MAXM]
    double2 *beta = nullptr;     // [MAXM]
    double2 *alpha = nullptr;
    double *r2 = nullptr;
    double *r2_host = nullptr;
    int cap = 0;
};
State g_st;

int ensure_state(int nb)
{
    if (!g_st.wn2) {
        QKX_HIP(hipMalloc(&g_st.wn2, sizeof(double) * MAXM));
        QKX_HIP(hipMalloc(&g_st.beta, sizeof(double2) * MAXM));
        QKX_HIP(hipMalloc(&g_st.alpha, sizeof(double2)));
        QKX_HIP(hipMalloc(&g_st.r2, sizeof(double)));
        QKX_HIP(hipHostMalloc((void **)&g_st.r2_host, sizeof(double), hipHostMallocDefault));
    }
    const int need = nb * 2 * MAXM;
    if (need > g_st.cap) {
        if (g_st.part) (void)hipFree(g_st.part);
        g_st.part = nullptr;
        QKX_HIP(hipMalloc(&g_st.part, sizeof(double) * need));
        g_st.cap = need;
    }
    return QKX_SUCCESS;
}

const double2 *dp(const qkx_spinor_s *f) { return (const double2 *)f->data; }
double2 *dp(qkx_spinor_s *f) { return (double2 *)f->data; }

int read_r2(double &r2)
{
    QKX_HIP(hipMemcpyAsync(g_st.r2_host, g_st.r2, sizeof(double), hipMemcpyDeviceToHost, qkx::g_stream));
    QKX_HIP(hipStreamSynchronize(qkx::g_stream));
    r2 = *g_st.r2_host;
    return QKX_SUCCESS;
}

} // namespace

extern "C" {

int qkx_solve_gcr(qkx_spinor *x, const qkx_spinor *b, const qkx_gauge *U, const qkx_gauge *U_net, qkx_net *net,
                  const qkx_params *p, const qkx_gcr_opts *o, qkx_solver_stats *st)
{
    QKX_CHECK_ARG(x && b && U && p && o, "qkx_solve_gcr: null argument");
    QKX_CHECK_ARG(x->prec == QKX_PREC_DOUBLE && b->prec == QKX_PREC_DOUBLE && x->subset == QKX_SUBSET_FULL &&
                      b->subset == QKX_SUBSET_FULL && x->nrhs == 1 && b->nrhs == 1,
                  "qkx_solve_gcr: x and b must be single full FP64 fields");
    QKX_CHECK_ARG(U->prec == QKX_PREC_DOUBLE && U->recon == QKX_RECON_18, "qkx_solve_gcr: U must be FP64 recon 18");
    QKX_CHECK_ARG(!net || U_net, "qkx_solve_gcr: the network needs its FP32 gauge field");
    QKX_CHECK_ARG(o->restart >= 1 && o->restart <= MAXM && o->maxIter > 0 && o->tol > 0,
                  "qkx_solve_gcr: restart must be 1..16, maxIter > 0, tol > 0");
    const int V = b->V;
    const int nb = std::min((V + kT - 1) / kT, 1024);
    if (int rc = ensure_state(nb)) return rc;
    auto t0 = std::chrono::steady_clock::now();
    const double kappa = p->kappa, a = 2.0 * p->kappa * p->mu * p->twistSign;
    const int m = o->restart;
    int rc;

    std::vector<qkx::Temp *> temps;
    struct Release {
        std::vector<qkx::Temp *> &t;
        ~Release()
        {
            for (auto f : t) delete f;
        }
    } release{temps};
    auto T = [&]() {
        temps.push_back(new qkx::Temp(b->L, QKX_PREC_DOUBLE, QKX_SUBSET_FULL));
        return temps.back()->f;
    };
    qkx_spinor_s *r = T(), *d = T();
    std::vector<qkx_spinor_s *> Z(m), W(m);
    for (int i = 0; i < m; i++) {
        Z[i] = T();
        W[i] = T();
    }
    for (auto t : temps)
        if (!t->f) return QKX_ERR_NOMEM;

    double b2 = 0;
    if ((rc = qkx_norm2(&b2, b))) return rc;
    if (!std::isfinite(b2)) { qkx::set_error("qkx_solve_gcr: non-finite source"); return QKX_ERR_NONFINITE; }
    if ((rc = qkx_spinor_zero(x))) return rc;
    const double target = o->tol * o->tol * b2;
    int status = QKX_NOT_CONVERGED, it = 0, cycles = 0, fallbacks = 0, cg_iters = 0;
    double r2 = b2;
    if (b2 == 0) status = QKX_SUCCESS;
    if ((rc = qkx_spinor_copy(r, b))) return rc;

    while (status == QKX_NOT_CONVERGED && it < o->maxIter) {
        const double r2_cycle = r2;
        cycles++;
        for (int i = 0; i < m && it < o->maxIter; i++) {
            if (net) {
                if ((rc = qkx::net_apply(net, Z[i], r, U_net))) return rc;
            } else if ((rc = qkx_spinor_copy(Z[i], r))) {
                return rc;
            }
            if ((rc = qkx::hop(W[i], Z[i], Z[i], U, QKX_BASIS_UKQCD, 0, -1, -kappa, 0, 1.0, a, 0))) return rc;
            VecSet Wp{}, Zp{};
            for (int j = 0; j < i; j++) {
                Wp.p[j] = dp(W[j]);
                Zp.p[j] = dp(Z[j]);
            }
            if (i > 0) {
                hipLaunchKernelGGL(mdot_kernel, dim3(nb), dim3(kT), 0, qkx::g_stream, Wp, i, dp(W[i]), V, g_st.part);
                hipLaunchKernelGGL(beta_kernel, dim3(1), dim3(64), 0, qkx::g_stream, g_st.part, nb, i, g_st.wn2,
                                   g_st.beta);
            }
            hipLaunchKernelGGL(orth_kernel, dim3(nb), dim3(kT), 0, qkx::g_stream, dp(W[i]), dp(Z[i]), Wp, Zp, i,
                               g_st.beta, dp(r), V, g_st.part);
            hipLaunchKernelGGL(alpha_kernel, dim3(1), dim3(64), 0, qkx::g_stream, g_st.part, nb, g_st.wn2 + i,
                               g_st.alpha);
            hipLaunchKernelGGL(xr_kernel, dim3(nb), dim3(kT), 0, qkx::g_stream, dp(x), dp(r), dp(Z[i]), dp(W[i]),
                               g_st.alpha, V, g_st.part);
            hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(64), 0, qkx::g_stream, g_st.part, nb, g_st.r2);
            QKX_HIP(hipGetLastError());
            it++;
            if ((rc = read_r2(r2))) return rc;
            if (!std::isfinite(r2)) {
                qkx::set_error("qkx_solve_gcr: non-finite residual");
                return QKX_ERR_NONFINITE;
            }
            if (r2 <= target) break;
        }
        // true residual at the end of each cycle
        if ((rc = qkx::hop(r, x, x, U, QKX_BASIS_UKQCD, 0, -1, -kappa, 0, 1.0, a, 0))) return rc;
        if ((rc = qkx::axpby_dev(r, b, -1.0, 1.0, nullptr, 0))) return rc;
        if ((rc = qkx_norm2(&r2, r))) return rc;
        if (r2 <= target) {
            status = QKX_SUCCESS;
            break;
        }
        if (o->stall > 0 && r2 > o->stall * o->stall * r2_cycle) {
            // stalled: the mixed-precision even-odd CG finishes the job from here
            fallbacks++;
            qkx_solve_opts so{1, QKX_PREC_SINGLE, QKX_RECON_12, 0.003, 50, std::max(1, o->maxIter - it),
                              std::min(0.5, std::sqrt(target / r2))};
            qkx_solver_stats cs{};
            rc = qkx_solve(d, r, U, nullptr, p, &so, &cs);
            if (rc != QKX_SUCCESS && rc != QKX_NOT_CONVERGED) return rc;
            cg_iters += cs.iterations;
            it += cs.iterations;
            if ((rc = qkx_axpy(x, 1.0, d))) return rc;
            if ((rc = qkx::hop(r, x, x, U, QKX_BASIS_UKQCD, 0, -1, -kappa, 0, 1.0, a, 0))) return rc;
            if ((rc = qkx::axpby_dev(r, b, -1.0, 1.0, nullptr, 0))) return rc;
            if ((rc = qkx_norm2(&r2, r))) return rc;
            if (r2 <= target) status = QKX_SUCCESS;
        }
    }
    if (status == QKX_NOT_CONVERGED) qkx::set_error("qkx_solve_gcr: maxIter reached");
    double tr = -1;
    if ((rc = qkx_true_residual(&tr, x, b, U, p))) return rc;
    if (st) {
        st->iterations = it;
        st->reliable_updates = cycles;
        st->converged = status == QKX_SUCCESS;
        st->iterated_r2 = r2;
        st->true_residual = tr;
        st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        st->fallback_iterations = cg_iters;
    }
    (void)fallbacks;
    return status;
}

} // extern "C"
