// qkx_prop.hip -- propagators on the GPU: the 12 columns of a point-source
// propagator as one batch field, their assembly, rescaling and rotation to
// the physical (twisted) basis, FP64.
//
// Reference: Propagator (include/lattice_util.h:158-185): absorbVector
// lib/lattice_util.cpp:2073-2082, copyVector :574-581, zero/copy/norm2
// :561-598, rotateToPhysicalBasePlus/Minus :2323-2453; Vector::rescale
// :2314-2321.
//
// A propagator is a batch field of 12 members (qkx_spinor_create_batch,
// FP64, full lattice): member r = 3 nu + c2 holds the column of source spin
// nu and source colour c2, so Propagator::M[iv][mu][nu][c1][c2] is component
// (mu, c1) of member 3 nu + c2 at site iv.  The batched solver's solution of
// the 12 point sources is therefore the propagator with no assembly copy.
#include "qkx_internal.h"
#include "qkx_core.h"

namespace {

constexpr int kThreads = 256;

// member r of dst <- src (a single field)
__global__ __launch_bounds__(kThreads) void absorb_kernel(SpinorD dst, SpinorD src, int V, int nr, int r)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= V) return;
    Cx<double> v[12];
    src.load(s, v);
    dst.store(s * nr + r, v);
}

// S -> (1 + i s g5) S (1 + i s g5) / 2 for the 4x4 spin matrix S[mu][nu] of
// one site and colour pair (c1, c2), g5 of the UKQCD basis (g5 swaps spins
// 0 <-> 2 and 1 <-> 3).  R = 1 + i s g5 has R[a][a] = 1 and R[a][a^2] = i s,
// so every product is exact and each entry is the sum of four terms, added in
// the reference's order (mu', nu' ascending) so that the result matches it
// bit for bit.  One thread per (site, c1, c2).
__global__ __launch_bounds__(kThreads) void rotate_kernel(SpinorD p, int V, int sign)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= V * 9) return;
    const int s = u / 9, cc = u - s * 9, c1 = cc / 3, c2 = cc - c1 * 3;
    Cx<double> S[4][4];
#pragma unroll
    for (int mu = 0; mu < 4; mu++)
#pragma unroll
        for (int nu = 0; nu < 4; nu++) {
            const double2 d = ld(p.p + (size_t)(mu * 3 + c1) * p.stride, (uint32_t)(s * 12 + nu * 3 + c2));
            S[mu][nu] = {d.x, d.y};
        }
    const double sg = (double)sign;
    // R[a][b] times z: 1 -> z, i s -> i s z
    auto rmul = [&](int a, int b, Cx<double> z) -> Cx<double> {
        return a == b ? z : Cx<double>{-sg * z.im, sg * z.re};
    };
#pragma unroll
    for (int mu = 0; mu < 4; mu++)
#pragma unroll
        for (int nu = 0; nu < 4; nu++) {
            const int m0 = mu < (mu ^ 2) ? mu : (mu ^ 2), m1 = m0 ^ 2;
            const int n0 = nu < (nu ^ 2) ? nu : (nu ^ 2), n1 = n0 ^ 2;
            Cx<double> acc = rmul(n0, nu, rmul(mu, m0, S[m0][n0]));
            acc = acc + rmul(n1, nu, rmul(mu, m0, S[m0][n1]));
            acc = acc + rmul(n0, nu, rmul(mu, m1, S[m1][n0]));
            acc = acc + rmul(n1, nu, rmul(mu, m1, S[m1][n1]));
            st(p.p + (size_t)(mu * 3 + c1) * p.stride, (uint32_t)(s * 12 + nu * 3 + c2),
               make_double2(0.5 * acc.re, 0.5 * acc.im));
        }
}

// reference layout Complex[V][4][4][3][3] <-> batch layout; one thread per
// (site, member)
__global__ __launch_bounds__(kThreads) void prop_from_aos(SpinorD dst, const double2 *src, Geom g)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= g.V * 12) return;
    const int s = u / 12, r = u - s * 12, nu = r / 3, c2 = r - nu * 3;
    const size_t lex = (size_t)blocked_to_lex(s, g);
    Cx<double> v[12];
#pragma unroll
    for (int mu = 0; mu < 4; mu++)
#pragma unroll
        for (int c1 = 0; c1 < 3; c1++) {
            const double2 d = src[(((lex * 4 + mu) * 4 + nu) * 3 + c1) * 3 + c2];
            v[mu * 3 + c1] = {d.x, d.y};
        }
    dst.store(u, v);
}

__global__ __launch_bounds__(kThreads) void prop_to_aos(double2 *dst, SpinorD src, Geom g)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= g.V * 12) return;
    const int s = u / 12, r = u - s * 12, nu = r / 3, c2 = r - nu * 3;
    const size_t lex = (size_t)blocked_to_lex(s, g);
    Cx<double> v[12];
    src.load(u, v);
#pragma unroll
    for (int mu = 0; mu < 4; mu++)
#pragma unroll
        for (int c1 = 0; c1 < 3; c1++)
            dst[(((lex * 4 + mu) * 4 + nu) * 3 + c1) * 3 + c2] = make_double2(v[mu * 3 + c1].re, v[mu * 3 + c1].im);
}

bool is_prop(const qkx_spinor_s *p)
{
    return p && p->prec == QKX_PREC_DOUBLE && p->subset == QKX_SUBSET_FULL && p->nrhs == 12;
}

} // namespace

extern "C" {

int qkx_prop_absorb(qkx_spinor *prop, const qkx_spinor *phi, int nu, int c2)
{
    QKX_CHECK_ARG(is_prop(prop), "qkx_prop_absorb: the propagator must be a full FP64 batch of 12");
    QKX_CHECK_ARG(phi && phi->prec == QKX_PREC_DOUBLE && phi->subset == QKX_SUBSET_FULL && phi->nrhs == 1 &&
                      phi->V == prop->V,
                  "qkx_prop_absorb: phi must be a single full FP64 field on the propagator's lattice");
    QKX_CHECK_ARG(nu >= 0 && nu < 4 && c2 >= 0 && c2 < 3, "qkx_prop_absorb: nu must be 0..3 and c2 0..2");
    const int V = prop->V;
    hipLaunchKernelGGL(absorb_kernel, dim3((V + kThreads - 1) / kThreads), dim3(kThreads), 0, qkx::g_stream,
                       make_spinor<SpinorD>(prop), make_spinor<SpinorD>(phi), V, 12, nu * 3 + c2);
    QKX_HIP(hipGetLastError());
    return QKX_SUCCESS;
}

int qkx_prop_rotate_physical(qkx_spinor *prop, int sign)
{
    QKX_CHECK_ARG(is_prop(prop), "qkx_prop_rotate_physical: the propagator must be a full FP64 batch of 12");
    QKX_CHECK_ARG(sign == 1 || sign == -1, "qkx_prop_rotate_physical: sign must be +1 or -1");
    QKX_CHECK_ARG((long long)prop->V * 9 < (1ll << 31), "qkx_prop_rotate_physical: lattice too large");
    const int V = prop->V;
    hipLaunchKernelGGL(rotate_kernel, dim3((V * 9 + kThreads - 1) / kThreads), dim3(kThreads), 0, qkx::g_stream,
                       make_spinor<SpinorD>(prop), V, sign);
    QKX_HIP(hipGetLastError());
    return QKX_SUCCESS;
}

int qkx_spinor_rescale(qkx_spinor *x, double a)
{
    QKX_CHECK_ARG(x, "qkx_spinor_rescale: null field");
    return qkx::axpby_dev(x, x, a, 0.0, nullptr, 0);
}

int qkx_prop_upload(qkx_spinor *prop, const double *M)
{
    QKX_CHECK_ARG(is_prop(prop) && M, "qkx_prop_upload: the propagator must be a full FP64 batch of 12");
    const size_t bytes = sizeof(double) * 288 * (size_t)prop->V;
    void *st = qkx::stage_buffer(bytes);
    if (!st) { qkx::set_error("qkx_prop_upload: staging allocation failed"); return QKX_ERR_NOMEM; }
    QKX_HIP(hipMemcpyAsync(st, M, bytes, hipMemcpyHostToDevice, qkx::g_stream));
    const Geom g = qkx::make_geom(prop->L, 1);
    hipLaunchKernelGGL(prop_from_aos, dim3((prop->V * 12 + kThreads - 1) / kThreads), dim3(kThreads), 0,
                       qkx::g_stream, make_spinor<SpinorD>(prop), (const double2 *)st, g);
    QKX_HIP(hipGetLastError());
    QKX_HIP(hipStreamSynchronize(qkx::g_stream));
    return QKX_SUCCESS;
}

int qkx_prop_download(const qkx_spinor *prop, double *M)
{
    QKX_CHECK_ARG(is_prop(prop) && M, "qkx_prop_download: the propagator must be a full FP64 batch of 12");
    const size_t bytes = sizeof(double) * 288 * (size_t)prop->V;
    void *st = qkx::stage_buffer(bytes);
    if (!st) { qkx::set_error("qkx_prop_download: staging allocation failed"); return QKX_ERR_NOMEM; }
    const Geom g = qkx::make_geom(prop->L, 1);
    hipLaunchKernelGGL(prop_to_aos, dim3((prop->V * 12 + kThreads - 1) / kThreads), dim3(kThreads), 0,
                       qkx::g_stream, (double2 *)st, make_spinor<SpinorD>(prop), g);
    QKX_HIP(hipGetLastError());
    QKX_HIP(hipMemcpyAsync(M, st, bytes, hipMemcpyDeviceToHost, qkx::g_stream));
    QKX_HIP(hipStreamSynchronize(qkx::g_stream));
    return QKX_SUCCESS;
}

} // extern "C"
