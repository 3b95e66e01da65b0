// qkx_blas.hip -- solver BLAS: real dot, norm, axpy/xpay and the fused CG
// update, with single-launch reductions.
//
// Reference: reCdotXdagY lib/blas_qkxTM.cpp:4-28 (Re sum conj(x) y, then an
// MPI_Allreduce), X_eq_X_p_aY :31-42, X_eq_aX_p_Y :45-56.
//
// Reductions: every thread accumulates in FP64, a wavefront reduces with
// cross-lane shuffles, the workgroup through LDS and writes one partial; a
// second one-workgroup launch adds the partials in double-double (compensated)
// arithmetic in a fixed order.  No atomics and no device-wide fences (which
// write back the L2 of every XCD), and the result is reproducible.
#include "qkx_internal.h"
#include "qkx_core.h"

namespace {

constexpr int kRedBlocks = 1024;
constexpr int kRedThreads = 256;

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Per-workgroup partial of v (wavefront shuffles, then LDS).
__device__ void block_partial(double v, double *dst)
{
    __shared__ double lds[kRedThreads / 64];
    double w = wave_sum(v);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = w;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0;
        for (int i = 0; i < (int)(blockDim.x >> 6); i++) s += lds[i];
        *dst = s;
    }
}

template <class S>
__global__ __launch_bounds__(kRedThreads) void redot_kernel(S x, S y, int n, int mode, double *partial)
{
    double acc = 0.0;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
        Cx<typename S::Real> a[12];
        x.load(s, a);
        if (mode == 0) {
            Cx<typename S::Real> b[12];
            y.load(s, b);
#pragma unroll
            for (int k = 0; k < 12; k++)
                acc += (double)a[k].re * (double)b[k].re + (double)a[k].im * (double)b[k].im;
        } else {
#pragma unroll
            for (int k = 0; k < 12; k++)
                acc += (double)a[k].re * (double)a[k].re + (double)a[k].im * (double)a[k].im;
        }
    }
    block_partial(acc, partial + blockIdx.x);
}

// The true-residual step of a reliable update in one pass: s = b - s (as
// axpby_kernel forms it with a = -1, b = 1), per-workgroup |s|^2 partials in
// redot_kernel's grid and order (so finish_kernel gives the bits qkx_norm2
// gives), and, with H, s rounded into the 16-bit field h as spinor_convert
// rounds it.
template <class S, class SH, bool H>
__global__ __launch_bounds__(kRedThreads) void resid_norm_kernel(S s_, S b_, SH h, int n, double *partial)
{
    using T = typename S::Real;
    double acc = 0.0;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
        Cx<T> a[12], b[12];
        s_.load(s, a);
        b_.load(s, b);
#pragma unroll
        for (int k = 0; k < 12; k++) a[k] = {T(-1.0) * a[k].re + T(1.0) * b[k].re, T(-1.0) * a[k].im + T(1.0) * b[k].im};
        s_.store(s, a);
        if constexpr (H) {
            Cx<typename SH::Real> w[12];
#pragma unroll
            for (int k = 0; k < 12; k++) w[k] = {(typename SH::Real)a[k].re, (typename SH::Real)a[k].im};
            h.store(s, w);
        }
#pragma unroll
        for (int k = 0; k < 12; k++) acc += (double)a[k].re * (double)a[k].re + (double)a[k].im * (double)a[k].im;
    }
    block_partial(acc, partial + blockIdx.x);
}

// One workgroup: result = compensated (double-double) sum of partial[0..n).
__global__ __launch_bounds__(kRedThreads) void finish_kernel(const double *partial, int n, double *result)
{
    __shared__ double lh[kRedThreads], ll[kRedThreads];
    DD acc{0.0, 0.0};
    for (int b = threadIdx.x; b < n; b += blockDim.x) acc = dd_add(acc, partial[b]);
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
    if (threadIdx.x == 0) *result = dd_value(DD{lh[0], ll[0]});
}

// x = ax*x + ay*y   (scalars from host or from device memory *sc)
template <class S>
__global__ __launch_bounds__(256) void axpby_kernel(S x, S y, int n, double ax, double ay, const double *sc,
                                                    int sc_mode)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    using T = typename S::Real;
    double a = ax, b = ay;
    if (sc_mode == 1) b = sc[0];          // x = x + sc*y
    else if (sc_mode == 2) a = sc[0];     // x = sc*x + y
    else if (sc_mode == 3) b = -sc[0];    // x = x - sc*y
    Cx<T> xv[12], yv[12];
    x.load(s, xv);
    y.load(s, yv);
#pragma unroll
    for (int k = 0; k < 12; k++) xv[k] = {T(a) * xv[k].re + T(b) * yv[k].re, T(a) * xv[k].im + T(b) * yv[k].im};
    x.store(s, xv);
}

int check_pair(const qkx_spinor_s *x, const qkx_spinor_s *y, const char *who)
{
    if (!x || !y) { qkx::set_error(std::string(who) + ": null field"); return QKX_ERR_ARG; }
    if (x->prec != y->prec || x->nsites != y->nsites || x->nrhs != y->nrhs || x->subset != y->subset) {
        qkx::set_error(std::string(who) + ": fields differ in precision, subset or size");
        return QKX_ERR_ARG;
    }
    return QKX_SUCCESS;
}

} // namespace

namespace qkx {

// device reduction to result slot 'slot' of the reduce buffer (no host sync)
int redot_dev(const qkx_spinor_s *x, const qkx_spinor_s *y, double *result)
{
    int rc = check_pair(x, y ? y : x, "redot");
    if (rc) return rc;
    int n = entries(x);
    int nb = std::min(kRedBlocks, (n + kRedThreads - 1) / kRedThreads);
    int mode = y ? 0 : 1;
    return dispatch_prec(x->prec, [&](auto tag) {
        using S = decltype(tag);
        S X = make_spinor<S>(x), Y = make_spinor<S>(y ? y : x);
        hipLaunchKernelGGL((redot_kernel<S>), dim3(nb), dim3(kRedThreads), 0, g_stream, X, Y, n, mode,
                           reduce_scratch());
        hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(kRedThreads), 0, g_stream, reduce_scratch(), nb, result);
        QKX_HIP(hipGetLastError());
        return QKX_SUCCESS;
    });
}

int finish_sum_dev(const double *partial, int n, double *result)
{
    hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(kRedThreads), 0, g_stream, partial, n, result);
    QKX_HIP(hipGetLastError());
    return QKX_SUCCESS;
}

int resid_norm(qkx_spinor_s *s, const qkx_spinor_s *b, qkx_spinor_s *h, double *s2)
{
    int rc = check_pair(s, b, "resid_norm");
    if (rc) return rc;
    if (s->prec != QKX_PREC_DOUBLE || (h && (h->prec != QKX_PREC_HALF || h->nsites != s->nsites || h->nrhs != s->nrhs ||
                                             h->subset != s->subset))) {
        qkx::set_error("resid_norm: FP64 s and b, and an FP16 h of the same sites");
        return QKX_ERR_ARG;
    }
    const int n = entries(s);
    const int nb = std::min(kRedBlocks, (n + kRedThreads - 1) / kRedThreads);
    const SpinorD S = make_spinor<SpinorD>(s), Bv = make_spinor<SpinorD>(b);
    if (h)
        hipLaunchKernelGGL((resid_norm_kernel<SpinorD, SpinorH, true>), dim3(nb), dim3(kRedThreads), 0, g_stream, S,
                           Bv, make_spinor<SpinorH>(h), n, reduce_scratch());
    else
        hipLaunchKernelGGL((resid_norm_kernel<SpinorD, SpinorD, false>), dim3(nb), dim3(kRedThreads), 0, g_stream, S,
                           Bv, S, n, reduce_scratch());
    hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(kRedThreads), 0, g_stream, reduce_scratch(), nb, reduce_result());
    QKX_HIP(hipGetLastError());
    QKX_HIP(hipMemcpyAsync(s2, reduce_result(), sizeof(double), hipMemcpyDeviceToHost, g_stream));
    QKX_HIP(hipStreamSynchronize(g_stream));
    return QKX_SUCCESS;
}

int axpby_dev(qkx_spinor_s *x, const qkx_spinor_s *y, double a, double b, const double *sc, int sc_mode)
{
    int rc = check_pair(x, y, "axpby");
    if (rc) return rc;
    int n = entries(x);
    return dispatch_prec(x->prec, [&](auto tag) {
        using S = decltype(tag);
        S X = make_spinor<S>(x), Y = make_spinor<S>(y);
        hipLaunchKernelGGL((axpby_kernel<S>), dim3((n + 255) / 256), dim3(256), 0, g_stream, X, Y, n, a, b,
                           sc, sc_mode);
        QKX_HIP(hipGetLastError());
        return QKX_SUCCESS;
    });
}

} // namespace qkx

extern "C" {

int qkx_redot(double *res, const qkx_spinor *x, const qkx_spinor *y)
{
    QKX_CHECK_ARG(res, "qkx_redot: null result");
    int rc = qkx::redot_dev(x, y, qkx::reduce_result());
    if (rc) return rc;
    QKX_HIP(hipMemcpyAsync(res, qkx::reduce_result(), sizeof(double), hipMemcpyDeviceToHost, qkx::g_stream));
    QKX_HIP(hipStreamSynchronize(qkx::g_stream));
    return QKX_SUCCESS;
}

int qkx_norm2(double *res, const qkx_spinor *x)
{
    QKX_CHECK_ARG(res && x, "qkx_norm2: null argument");
    int rc = qkx::redot_dev(x, nullptr, qkx::reduce_result());
    if (rc) return rc;
    QKX_HIP(hipMemcpyAsync(res, qkx::reduce_result(), sizeof(double), hipMemcpyDeviceToHost, qkx::g_stream));
    QKX_HIP(hipStreamSynchronize(qkx::g_stream));
    return QKX_SUCCESS;
}

int qkx_axpy(qkx_spinor *x, double a, const qkx_spinor *y) { return qkx::axpby_dev(x, y, 1.0, a, nullptr, 0); }
int qkx_xpay(qkx_spinor *x, double a, const qkx_spinor *y) { return qkx::axpby_dev(x, y, a, 1.0, nullptr, 0); }

} // extern "C"
