// qkx_fields.hip -- device fields: allocation, host (reference AoS layout)
// <-> device (parity-blocked SoA) conversion, precision conversion, copies.
//
// Host layout of the reference: Vector::M Complex[V][4][3]
// (include/lattice_util.h:90), Gauge::M Complex[V][4][3][3] (:196).
#include "qkx_internal.h"
#include "qkx_core.h"
#include <string>
#include <vector>

namespace {

template <class S>
__global__ void spinor_from_aos(S dst, const double2 *src, Geom g, int par0, int n, int nrhs, int rhs)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int s_full = par0 * g.Vh + i;   // par0 = 0 for full fields
    int lex = blocked_to_lex(s_full, g);
    Cx<typename S::Real> v[12];
#pragma unroll
    for (int k = 0; k < 12; k++) {
        double2 d = src[(size_t)lex * 12 + k];
        v[k] = {(typename S::Real)d.x, (typename S::Real)d.y};
    }
    dst.store(i * nrhs + rhs, v);
}

template <class S>
__global__ void spinor_to_aos(double2 *dst, S src, Geom g, int par0, int n, int zero_other, int nrhs, int rhs)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int s_full = par0 * g.Vh + i;
    int lex = blocked_to_lex(s_full, g);
    Cx<typename S::Real> v[12];
    src.load(i * nrhs + rhs, v);
#pragma unroll
    for (int k = 0; k < 12; k++) dst[(size_t)lex * 12 + k] = make_double2((double)v[k].re, (double)v[k].im);
    if (zero_other) {
        // the other parity's site with the same cb
        int lex2 = blocked_to_lex((1 - par0) * g.Vh + i, g);
#pragma unroll
        for (int k = 0; k < 12; k++) dst[(size_t)lex2 * 12 + k] = make_double2(0.0, 0.0);
    }
}

// dst site i (in dst storage) <- src site (i + src_shift)
template <class SD, class SS>
__global__ void spinor_convert(SD dst, SS src, int n, int dst_shift, int src_shift)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Cx<typename SS::Real> v[12];
    src.load(i + src_shift, v);
    Cx<typename SD::Real> w[12];
#pragma unroll
    for (int k = 0; k < 12; k++) w[k] = {(typename SD::Real)v[k].re, (typename SD::Real)v[k].im};
    dst.store(i + dst_shift, w);
}

template <class S>
__global__ void spinor_zero(S dst, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Cx<typename S::Real> v[12];
#pragma unroll
    for (int k = 0; k < 12; k++) v[k] = {0, 0};
    dst.store(i, v);
}

template <class S>
__global__ void spinor_zero_range(S dst, int n, int shift)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Cx<typename S::Real> v[12];
#pragma unroll
    for (int k = 0; k < 12; k++) v[k] = {0, 0};
    dst.store(i + shift, v);
}

// Gaussian noise (variance 1 per real component) from a counter-based hash of
// (seed, entry, component): splitmix64 finalisers (Steele, Lea and Flood,
// OOPSLA 2014) give two uniforms, the Box-Muller transform the pair of
// normals.  The same seed gives the same field on any grid.
__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

template <class S>
__global__ void spinor_gaussian(S dst, int n, unsigned long long seed)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    using R = typename S::Real;
    Cx<R> v[12];
    const uint64_t base = mix64(seed);
#pragma unroll
    for (int k = 0; k < 12; k++) {
        const uint64_t h1 = mix64(base ^ (((uint64_t)i << 4) | (uint64_t)k));
        const uint64_t h2 = mix64(h1);
        const double u1 = ((double)(h1 >> 11) + 0.5) * (1.0 / 9007199254740992.0);   // (0, 1)
        const double u2 = (double)(h2 >> 11) * (1.0 / 9007199254740992.0);          // [0, 1)
        const double rad = sqrt(-2.0 * log(u1)), ph = 6.283185307179586 * u2;
        v[k] = {(R)(rad * cos(ph)), (R)(rad * sin(ph))};
    }
    dst.store(i, v);
}

// ---- gauge ------------------------------------------------------------------
// Store one link U[9] (complex) at (mu, s) in the requested format.
template <int PREC, int RECON>
__device__ __forceinline__ void gauge_store(void *base, int V, int mu, int s, const Cx<double> U[9])
{
    if constexpr (PREC == QKX_PREC_DOUBLE) {
        double2 *p = (double2 *)base;
        constexpr int NC = RECON == QKX_RECON_18 ? 9 : 6;
#pragma unroll
        for (int k = 0; k < NC; k++) p[(size_t)(mu * NC + k) * V + s] = make_double2(U[k].re, U[k].im);
    } else if constexpr (PREC == QKX_PREC_SINGLE && RECON == QKX_RECON_18) {
        float2 *p = (float2 *)base;
#pragma unroll
        for (int k = 0; k < 9; k++) p[(size_t)(mu * 9 + k) * V + s] = make_float2((float)U[k].re, (float)U[k].im);
    } else if constexpr (PREC == QKX_PREC_SINGLE) {
        float4 *p = (float4 *)base;
#pragma unroll
        for (int k = 0; k < 3; k++)
            p[(size_t)(mu * 3 + k) * V + s] = make_float4((float)U[2 * k].re, (float)U[2 * k].im,
                                                          (float)U[2 * k + 1].re, (float)U[2 * k + 1].im);
    } else {
        // 16-bit: U0..U3 in a uint4 array [4][V], U4, U5 in a uint2 array [4][V]
        uint4 *p4 = (uint4 *)base;
        uint2 *p2 = (uint2 *)((char *)base + (size_t)16 * 4 * V);
        float f[8];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            f[2 * k] = (float)U[k].re;
            f[2 * k + 1] = (float)U[k].im;
        }
        if constexpr (PREC == QKX_PREC_HALF) {
            p4[(size_t)mu * V + s] = pack_h8(f);
            p2[(size_t)mu * V + s] = pack_h4((float)U[4].re, (float)U[4].im, (float)U[5].re, (float)U[5].im);
        } else {
            p4[(size_t)mu * V + s] = pack_b8(f);
            p2[(size_t)mu * V + s] = pack_b4((float)U[4].re, (float)U[4].im, (float)U[5].re, (float)U[5].im);
        }
    }
}

// flip_t: the t-links of the last time slice are stored negated (links that
// carry the antiperiodic sign, stored in 12 reals with boundaryT = -1 instead)
template <int PREC, int RECON>
__global__ void gauge_from_aos(void *dst, const double2 *src, Geom g, int flip_t)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.V) return;
    int lex = blocked_to_lex(s, g);
    const bool last_t = flip_t && lex / (g.L[0] * g.L[1] * g.L[2]) == g.L[3] - 1;
    for (int mu = 0; mu < 4; mu++) {
        Cx<double> U[9];
        const double sg = (mu == 3 && last_t) ? -1.0 : 1.0;
#pragma unroll
        for (int k = 0; k < 9; k++) {
            double2 d = src[((size_t)lex * 4 + mu) * 9 + k];
            U[k] = {sg * d.x, sg * d.y};
        }
        gauge_store<PREC, RECON>(dst, g.V, mu, s, U);
    }
}

// Per-workgroup max over the links of a host-layout field (U[lex][mu][9]) of
// |row2 - conj(row0 x row1)| as stored (part[2b]) and with the last time
// slice's t-links negated (part[2b + 1]); a NaN gives NaN.
__device__ __forceinline__ double max_nan_aos(double a, double b) { return (b > a || b != b) ? b : a; }
__global__ __launch_bounds__(256) void recon_dev_aos_kernel(const double2 *src, Geom g, double *part)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;   // (lex, mu), mu fastest
    double m = 0, mf = 0;
    if (t < 4 * g.V) {
        const int lex = t >> 2, mu = t & 3;
        const bool last_t = mu == 3 && lex / (g.L[0] * g.L[1] * g.L[2]) == g.L[3] - 1;
        Cx<double> u[9], r[9];
#pragma unroll
        for (int k = 0; k < 9; k++) {
            const double2 d = src[(size_t)t * 9 + k];
            u[k] = {d.x, d.y};
        }
#pragma unroll
        for (int k = 0; k < 6; k++) r[k] = u[k];
        recon_row2(r);   // unchanged when rows 0 and 1 are both negated
#pragma unroll
        for (int k = 6; k < 9; k++) {
            const double dre = fabs(u[k].re - r[k].re), dim = fabs(u[k].im - r[k].im);
            const double fre = last_t ? fabs(u[k].re + r[k].re) : dre, fim = last_t ? fabs(u[k].im + r[k].im) : dim;
            m = max_nan_aos(max_nan_aos(m, dre), dim);
            mf = max_nan_aos(max_nan_aos(mf, fre), fim);
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        m = max_nan_aos(m, __shfl_xor(m, off, 64));
        mf = max_nan_aos(mf, __shfl_xor(mf, off, 64));
    }
    __shared__ double lds[2][4];
    if ((threadIdx.x & 63) == 0) {
        lds[0][threadIdx.x >> 6] = m;
        lds[1][threadIdx.x >> 6] = mf;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = lds[0][0], vf = lds[1][0];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
            v = max_nan_aos(v, lds[0][w]);
            vf = max_nan_aos(vf, lds[1][w]);
        }
        part[2 * blockIdx.x] = v;
        part[2 * blockIdx.x + 1] = vf;
    }
}

template <int PREC, int RECON>
// flip_t: the t-links of the last time slice are stored negated (a field
// whose links carry the antiperiodic sign, copied with boundaryT = -1 instead)
__global__ void gauge_from_d18(void *dst, const double2 *src, Geom g, int flip_t)
{
    const int V = g.V;
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= V) return;
    const bool last_t = flip_t && blocked_to_lex(s, g) / (g.L[0] * g.L[1] * g.L[2]) == g.L[3] - 1;
    for (int mu = 0; mu < 4; mu++) {
        Cx<double> U[9];
        const double sg = (mu == 3 && last_t) ? -1.0 : 1.0;
#pragma unroll
        for (int k = 0; k < 9; k++) {
            double2 d = src[(size_t)(mu * 9 + k) * V + s];
            U[k] = {sg * d.x, sg * d.y};
        }
        gauge_store<PREC, RECON>(dst, V, mu, s, U);
    }
}

// FP64 recon-18 device links -> reference layout U[lex][mu][row][col]
__global__ void gauge_to_aos_d18(double2 *dst, const double2 *src, Geom g)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.V) return;
    int lex = blocked_to_lex(s, g);
    for (int mu = 0; mu < 4; mu++)
#pragma unroll
        for (int k = 0; k < 9; k++) dst[((size_t)lex * 4 + mu) * 9 + k] = src[(size_t)(mu * 9 + k) * g.V + s];
}

// ---- staging buffer ---------------------------------------------------------
void *g_stage = nullptr;
size_t g_stage_bytes = 0;

void *stage(size_t bytes)
{
    if (bytes > g_stage_bytes) {
        if (g_stage) (void)hipFree(g_stage);
        g_stage = nullptr;
        g_stage_bytes = 0;
        if (hipMalloc(&g_stage, bytes) != hipSuccess) return nullptr;
        g_stage_bytes = bytes;
    }
    return g_stage;
}

template <class F>
int dispatch_spinor(const qkx_spinor_s *f, F fn)
{
    return dispatch_prec(f->prec, [&](auto tag) { return fn(make_spinor<decltype(tag)>(f)); });
}

bool valid_dims(const int L[4])
{
    for (int m = 0; m < 4; m++)
        if (L[m] < 2 || (L[m] & 1)) return false;
    long V = (long)L[0] * L[1] * L[2] * L[3];
    return V < (1L << 28);
}

} // namespace

namespace qkx {

// every storage chunk is 16 bytes: DOUBLE 12 per site, SINGLE 6, 16-bit 3
size_t spinor_elem_bytes(int) { return 16; }
int spinor_nchunks(int prec) { return prec == QKX_PREC_DOUBLE ? 12 : prec == QKX_PREC_SINGLE ? 6 : 3; }

// storage offset (in entries) of each parity's first site
void parity_offsets(const qkx_spinor_s *f, int off[2])
{
    off[0] = 0;
    off[1] = f->subset == QKX_SUBSET_FULL ? f->Vh * f->nrhs : 0;
}

Geom make_geom(const int L[4], int bc)
{
    Geom g;
    for (int m = 0; m < 4; m++) g.L[m] = L[m];
    g.Lh0 = L[0] / 2;
    g.V = L[0] * L[1] * L[2] * L[3];
    g.Vh = g.V / 2;
    g.bc = bc;
    g.dLh0 = make_fastdiv(g.Lh0);
    g.dL1 = make_fastdiv(L[1]);
    g.dL2 = make_fastdiv(L[2]);
    return g;
}

void gauge_modified(qkx_gauge_s *g) { g->gen++; }

void *stage_buffer(size_t bytes) { return stage(bytes); }

// Per-workgroup max over (site, mu) of |row2 - conj(row0 x row1)|, the
// reconstruction the 12-real fields use (part[2b]), and the same with the
// t-links of the last time slice negated (part[2b + 1]); a NaN gives NaN.
__device__ __forceinline__ double max_nan(double a, double b) { return (b > a || b != b) ? b : a; }
__global__ __launch_bounds__(256) void recon_dev_kernel(GaugeD18 G, Geom g, double *part)
{
    const int V = g.V;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double m = 0, mf = 0;
    if (t < 4 * V) {
        const int mu = t / V, s = t - mu * V;
        const bool last_t = mu == 3 && blocked_to_lex(s, g) / (g.L[0] * g.L[1] * g.L[2]) == g.L[3] - 1;
        Cx<double> u[9], r[9];
        G.load(mu, s, u);
#pragma unroll
        for (int k = 0; k < 6; k++) r[k] = u[k];
        recon_row2(r);   // unchanged when rows 0 and 1 are both negated
#pragma unroll
        for (int k = 6; k < 9; k++) {
            const double dre = fabs(u[k].re - r[k].re), dim = fabs(u[k].im - r[k].im);
            const double fre = last_t ? fabs(u[k].re + r[k].re) : dre, fim = last_t ? fabs(u[k].im + r[k].im) : dim;
            m = max_nan(max_nan(m, dre), dim);
            mf = max_nan(max_nan(mf, fre), fim);
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        m = max_nan(m, __shfl_xor(m, off, 64));
        mf = max_nan(mf, __shfl_xor(mf, off, 64));
    }
    __shared__ double lds[2][4];
    if ((threadIdx.x & 63) == 0) {
        lds[0][threadIdx.x >> 6] = m;
        lds[1][threadIdx.x >> 6] = mf;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = lds[0][0], vf = lds[1][0];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
            v = max_nan(v, lds[0][w]);
            vf = max_nan(vf, lds[1][w]);
        }
        part[2 * blockIdx.x] = v;
        part[2 * blockIdx.x + 1] = vf;
    }
}

// dst <- src (FP64 18-real), with the last time slice's t-links negated when
// flip_t (dst then applies boundaryT = -1 itself)
int gauge_convert_flip(qkx_gauge_s *dst, const qkx_gauge_s *src, int flip_t)
{
    dst->bc = flip_t ? -1 : src->bc;
    dim3 grid((dst->V + 255) / 256), blk(256);
    const double2 *s2 = (const double2 *)src->data;
    const Geom g = make_geom(src->L, src->bc);
    if (dst->prec == QKX_PREC_DOUBLE && dst->recon == 18) hipLaunchKernelGGL((gauge_from_d18<QKX_PREC_DOUBLE, 18>), grid, blk, 0, g_stream, dst->data, s2, g, flip_t);
    else if (dst->prec == QKX_PREC_DOUBLE) hipLaunchKernelGGL((gauge_from_d18<QKX_PREC_DOUBLE, 12>), grid, blk, 0, g_stream, dst->data, s2, g, flip_t);
    else if (dst->prec == QKX_PREC_SINGLE && dst->recon == 18) hipLaunchKernelGGL((gauge_from_d18<QKX_PREC_SINGLE, 18>), grid, blk, 0, g_stream, dst->data, s2, g, flip_t);
    else if (dst->prec == QKX_PREC_SINGLE) hipLaunchKernelGGL((gauge_from_d18<QKX_PREC_SINGLE, 12>), grid, blk, 0, g_stream, dst->data, s2, g, flip_t);
    else if (dst->prec == QKX_PREC_HALF) hipLaunchKernelGGL((gauge_from_d18<QKX_PREC_HALF, 12>), grid, blk, 0, g_stream, dst->data, s2, g, flip_t);
    else hipLaunchKernelGGL((gauge_from_d18<QKX_PREC_BF16, 12>), grid, blk, 0, g_stream, dst->data, s2, g, flip_t);
    QKX_HIP(hipGetLastError());
    gauge_modified(dst);
    dst->src_gen = src->gen;
    return QKX_SUCCESS;
}

// How a 12-real copy at precision prec can carry U's links: 0 as stored (SU(3)
// to 1e-14 for FP64, 1e-6 below), 1 with the last time slice's t-links
// negated and boundaryT = -1 (links that carry the antiperiodic sign, with
// boundaryT = +1), -1 not at all.
int recon12_mode(const qkx_gauge_s *U, int prec, int *mode)
{
    double dev = 0;
    int rc = gauge_recon_deviation(U, &dev);
    if (rc) return rc;
    const double tol = prec == QKX_PREC_DOUBLE ? 1e-14 : 1e-6;
    if (dev <= tol) *mode = 0;
    else if (U->bc == 1 && U->recon_dev_flip <= tol) *mode = 1;
    else *mode = -1;
    return QKX_SUCCESS;
}

int gauge_cached(qkx_gauge_s **cache, const qkx_gauge_s *U, int prec, int recon)
{
    int rc, mode = 0;
    if (recon == QKX_RECON_12 && U->prec == QKX_PREC_DOUBLE && U->recon == QKX_RECON_18) {
        if ((rc = recon12_mode(U, prec, &mode))) return rc;
        if (mode < 0) {
            set_error("12-real copy: the links are not SU(3) (row 2 != conj(row0 x row1) by " +
                      std::to_string(U->recon_dev) + ", also with the boundary sign removed); upload them "
                      "without the boundary sign with boundaryT = -1, or use 18-real storage / FP64");
            return QKX_ERR_ARG;
        }
    }
    const int bc = mode == 1 ? -1 : U->bc;
    if (*cache && ((*cache)->prec != prec || (*cache)->recon != recon || (*cache)->bc != bc)) {
        qkx_gauge_destroy(*cache);
        *cache = nullptr;
    }
    if (!*cache) {
        if ((rc = qkx_gauge_create(cache, U->L, prec, recon, bc))) return rc;
    } else if ((*cache)->src_gen == U->gen) {
        return QKX_SUCCESS;
    }
    return gauge_convert_flip(*cache, U, mode);   // records U->gen
}

int gauge_recon_deviation(const qkx_gauge_s *U, double *dev)
{
    if (U->prec != QKX_PREC_DOUBLE || U->recon != QKX_RECON_18) {
        set_error("gauge_recon_deviation: FP64 18-real fields only");
        return QKX_ERR_ARG;
    }
    qkx_gauge_s *Um = const_cast<qkx_gauge_s *>(U);
    if (Um->recon_gen != U->gen) {
        const int n = 4 * U->V, nb = (n + 255) / 256;
        double *part = nullptr;
        QKX_HIP(hipMalloc(&part, 2 * nb * sizeof(double)));
        hipLaunchKernelGGL(recon_dev_kernel, dim3(nb), dim3(256), 0, g_stream, make_gauge<GaugeD18>(U),
                           make_geom(U->L, U->bc), part);
        std::vector<double> h(2 * nb);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(h.data(), part, 2 * nb * sizeof(double), hipMemcpyDeviceToHost, g_stream);
        if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
        (void)hipFree(part);
        QKX_HIP(e);
        double m = 0, mf = 0;
        for (int b = 0; b < nb; b++) {   // a NaN link propagates
            m = h[2 * b] > m || h[2 * b] != h[2 * b] ? h[2 * b] : m;
            mf = h[2 * b + 1] > mf || h[2 * b + 1] != h[2 * b + 1] ? h[2 * b + 1] : mf;
        }
        Um->recon_dev = m;
        Um->recon_dev_flip = mf;
        Um->recon_gen = U->gen;
    }
    *dev = U->recon_dev;
    return QKX_SUCCESS;
}

int gauge_fp64_operator(const qkx_gauge_s *U, const qkx_gauge_s **op)
{
    *op = U;
    if (U->prec != QKX_PREC_DOUBLE || U->recon != QKX_RECON_18) return QKX_SUCCESS;
    int mode = 0;
    int rc = recon12_mode(U, QKX_PREC_DOUBLE, &mode);
    if (rc) return rc;
    if (mode < 0) return QKX_SUCCESS;
    qkx_gauge_s *Uc = const_cast<qkx_gauge_s *>(U);
    if ((rc = gauge_cached(&Uc->outer, U, QKX_PREC_DOUBLE, QKX_RECON_12))) return rc;
    *op = Uc->outer;
    return QKX_SUCCESS;
}

// zero the sites of one parity of a full field
int zero_parity(qkx_spinor_s *f, int parity)
{
    if (f->subset != QKX_SUBSET_FULL) return QKX_SUCCESS;
    int n = f->Vh * f->nrhs, shift = parity * f->Vh * f->nrhs;
    return dispatch_spinor(f, [&](auto S) {
        hipLaunchKernelGGL((spinor_zero_range<decltype(S)>), dim3((n + 255) / 256), dim3(256), 0, g_stream, S, n,
                           shift);
        QKX_HIP(hipGetLastError());
        return QKX_SUCCESS;
    });
}

} // namespace qkx

extern "C" {

int qkx_spinor_create(qkx_spinor **out, const int L[4], int prec, int subset)
{
    return qkx_spinor_create_batch(out, L, prec, subset, 1);
}

int qkx_spinor_create_batch(qkx_spinor **out, const int L[4], int prec, int subset, int nrhs)
{
    int rc = qkx::ensure_init();
    if (rc) return rc;
    QKX_CHECK_ARG(out && L, "qkx_spinor_create: null argument");
    QKX_CHECK_ARG(valid_dims(L), "qkx_spinor_create: every extent must be even and >= 2");
    QKX_CHECK_ARG(prec >= QKX_PREC_DOUBLE && prec <= QKX_PREC_BF16, "qkx_spinor_create: bad precision");
    QKX_CHECK_ARG(subset >= QKX_SUBSET_FULL && subset <= QKX_SUBSET_ODD, "qkx_spinor_create: bad subset");
    QKX_CHECK_ARG(nrhs >= 1 && nrhs <= QKX_MAX_RHS, "qkx_spinor_create: nrhs out of range");
    auto *f = new qkx_spinor_s();
    for (int m = 0; m < 4; m++) f->L[m] = L[m];
    f->V = L[0] * L[1] * L[2] * L[3];
    f->Vh = f->V / 2;
    f->prec = prec;
    f->subset = subset;
    f->nsites = subset == QKX_SUBSET_FULL ? f->V : f->Vh;
    f->nrhs = nrhs;
    f->bytes = qkx::spinor_elem_bytes(prec) * qkx::spinor_nchunks(prec) * (size_t)f->nsites * nrhs;
    if (hipMalloc(&f->data, f->bytes) != hipSuccess) {
        delete f;
        qkx::set_error("qkx_spinor_create: out of device memory");
        return QKX_ERR_NOMEM;
    }
    f->nrm = nullptr;
    if (prec == QKX_PREC_HALF) {
        if (hipMalloc(&f->nrm, sizeof(float) * (size_t)f->nsites * nrhs) != hipSuccess) {
            (void)hipFree(f->data);
            delete f;
            qkx::set_error("qkx_spinor_create: out of device memory");
            return QKX_ERR_NOMEM;
        }
    }
    *out = f;
    return qkx_spinor_zero(f);
}

int qkx_spinor_destroy(qkx_spinor *f)
{
    if (!f) return QKX_SUCCESS;
    (void)hipStreamSynchronize(qkx::g_stream);
    if (f->data) (void)hipFree(f->data);
    if (f->nrm) (void)hipFree(f->nrm);
    delete f;
    return QKX_SUCCESS;
}

int qkx_spinor_zero(qkx_spinor *f)
{
    QKX_CHECK_ARG(f, "qkx_spinor_zero: null field");
    int n = entries(f);
    return dispatch_spinor(f, [&](auto S) {
        hipLaunchKernelGGL((spinor_zero<decltype(S)>), dim3((n + 255) / 256), dim3(256), 0, qkx::g_stream, S, n);
        QKX_HIP(hipGetLastError());
        return QKX_SUCCESS;
    });
}

int qkx_spinor_gaussian(qkx_spinor *f, unsigned long long seed)
{
    QKX_CHECK_ARG(f, "qkx_spinor_gaussian: null field");
    int n = entries(f);
    return dispatch_spinor(f, [&](auto S) {
        hipLaunchKernelGGL((spinor_gaussian<decltype(S)>), dim3((n + 255) / 256), dim3(256), 0, qkx::g_stream, S, n,
                           seed);
        QKX_HIP(hipGetLastError());
        return QKX_SUCCESS;
    });
}

int qkx_spinor_upload(qkx_spinor *f, const double *psi) { return qkx_spinor_upload_rhs(f, 0, psi); }
int qkx_spinor_download(const qkx_spinor *f, double *psi) { return qkx_spinor_download_rhs(f, 0, psi); }

int qkx_spinor_upload_rhs(qkx_spinor *f, int rhs, const double *psi)
{
    QKX_CHECK_ARG(f && psi, "qkx_spinor_upload: null argument");
    QKX_CHECK_ARG(rhs >= 0 && rhs < f->nrhs, "qkx_spinor_upload: rhs out of range");
    size_t bytes = sizeof(double) * 24 * (size_t)f->V;
    void *st = stage(bytes);
    if (!st) { qkx::set_error("qkx_spinor_upload: staging allocation failed"); return QKX_ERR_NOMEM; }
    QKX_HIP(hipMemcpyAsync(st, psi, bytes, hipMemcpyHostToDevice, qkx::g_stream));
    Geom g = qkx::make_geom(f->L, 1);
    int par0 = f->subset == QKX_SUBSET_ODD ? 1 : 0;
    int n = f->nsites;
    int rc = dispatch_spinor(f, [&](auto S) {
        hipLaunchKernelGGL((spinor_from_aos<decltype(S)>), dim3((n + 255) / 256), dim3(256), 0, qkx::g_stream,
                           S, (const double2 *)st, g, par0, n, f->nrhs, rhs);
        QKX_HIP(hipGetLastError());
        return QKX_SUCCESS;
    });
    if (rc) return rc;
    QKX_HIP(hipStreamSynchronize(qkx::g_stream));
    return QKX_SUCCESS;
}

int qkx_spinor_download_rhs(const qkx_spinor *f, int rhs, double *psi)
{
    QKX_CHECK_ARG(f && psi, "qkx_spinor_download: null argument");
    QKX_CHECK_ARG(rhs >= 0 && rhs < f->nrhs, "qkx_spinor_download: rhs out of range");
    size_t bytes = sizeof(double) * 24 * (size_t)f->V;
    void *st = stage(bytes);
    if (!st) { qkx::set_error("qkx_spinor_download: staging allocation failed"); return QKX_ERR_NOMEM; }
    Geom g = qkx::make_geom(f->L, 1);
    int par0 = f->subset == QKX_SUBSET_ODD ? 1 : 0;
    int n = f->nsites;
    int zero_other = f->subset != QKX_SUBSET_FULL;
    int rc = dispatch_spinor(f, [&](auto S) {
        hipLaunchKernelGGL((spinor_to_aos<decltype(S)>), dim3((n + 255) / 256), dim3(256), 0, qkx::g_stream,
                           (double2 *)st, S, g, par0, n, zero_other, f->nrhs, rhs);
        QKX_HIP(hipGetLastError());
        return QKX_SUCCESS;
    });
    if (rc) return rc;
    QKX_HIP(hipMemcpyAsync(psi, st, bytes, hipMemcpyDeviceToHost, qkx::g_stream));
    QKX_HIP(hipStreamSynchronize(qkx::g_stream));
    return QKX_SUCCESS;
}

// Copy with precision conversion.  Full<->full and subset<->subset (same parity)
// copy everything; full -> subset takes that parity; subset -> full writes only
// that parity of the destination.
int qkx_spinor_copy(qkx_spinor *dst, const qkx_spinor *src)
{
    QKX_CHECK_ARG(dst && src, "qkx_spinor_copy: null argument");
    for (int m = 0; m < 4; m++) QKX_CHECK_ARG(dst->L[m] == src->L[m], "qkx_spinor_copy: lattices differ");
    QKX_CHECK_ARG(dst->nrhs == src->nrhs, "qkx_spinor_copy: numbers of right-hand sides differ");
    const int nr = dst->nrhs;
    int n, dshift = 0, sshift = 0;
    if (dst->subset == src->subset) {
        n = entries(dst);
    } else if (dst->subset != QKX_SUBSET_FULL && src->subset == QKX_SUBSET_FULL) {
        n = entries(dst);
        sshift = dst->subset == QKX_SUBSET_ODD ? src->Vh * nr : 0;
    } else if (dst->subset == QKX_SUBSET_FULL && src->subset != QKX_SUBSET_FULL) {
        n = entries(src);
        dshift = src->subset == QKX_SUBSET_ODD ? dst->Vh * nr : 0;
    } else {
        qkx::set_error("qkx_spinor_copy: even and odd subsets cannot be copied onto each other");
        return QKX_ERR_ARG;
    }
    return dispatch_spinor(dst, [&](auto SD) {
        return dispatch_spinor(src, [&](auto SS) {
            hipLaunchKernelGGL((spinor_convert<decltype(SD), decltype(SS)>), dim3((n + 255) / 256), dim3(256), 0,
                               qkx::g_stream, SD, SS, n, dshift, sshift);
            QKX_HIP(hipGetLastError());
            return QKX_SUCCESS;
        });
    });
}

int qkx_gauge_create(qkx_gauge **out, const int L[4], int prec, int recon, int boundaryT)
{
    int rc = qkx::ensure_init();
    if (rc) return rc;
    QKX_CHECK_ARG(out && L, "qkx_gauge_create: null argument");
    QKX_CHECK_ARG(valid_dims(L), "qkx_gauge_create: every extent must be even and >= 2");
    QKX_CHECK_ARG(recon == QKX_RECON_18 || recon == QKX_RECON_12, "qkx_gauge_create: recon must be 18 or 12");
    QKX_CHECK_ARG(prec >= QKX_PREC_DOUBLE && prec <= QKX_PREC_BF16, "qkx_gauge_create: bad precision");
    QKX_CHECK_ARG(!((prec == QKX_PREC_HALF || prec == QKX_PREC_BF16) && recon == QKX_RECON_18),
                  "qkx_gauge_create: 16-bit gauge storage needs recon 12");
    QKX_CHECK_ARG(boundaryT == 1 || boundaryT == -1, "qkx_gauge_create: boundaryT must be +1 or -1");
    auto *g = new qkx_gauge_s();
    for (int m = 0; m < 4; m++) g->L[m] = L[m];
    g->V = L[0] * L[1] * L[2] * L[3];
    g->Vh = g->V / 2;
    g->prec = prec;
    g->recon = recon;
    g->bc = boundaryT;
    g->inner = nullptr;
    g->mid = nullptr;
    g->outer = nullptr;
    g->gen = 0;
    g->src_gen = 0;
    g->recon_dev = -1.0;
    g->recon_dev_flip = -1.0;
    g->recon_gen = ~0u;   // not measured
    size_t per = (prec == QKX_PREC_DOUBLE ? 16 : prec == QKX_PREC_SINGLE ? 8 : 4) * (recon == 18 ? 9 : 6);
    g->bytes = per * 4 * (size_t)g->V;
    if (hipMalloc(&g->data, g->bytes) != hipSuccess) {
        delete g;
        qkx::set_error("qkx_gauge_create: out of device memory");
        return QKX_ERR_NOMEM;
    }
    *out = g;
    return QKX_SUCCESS;
}

int qkx_gauge_destroy(qkx_gauge *g)
{
    if (!g) return QKX_SUCCESS;
    (void)hipStreamSynchronize(qkx::g_stream);
    if (g->inner) qkx_gauge_destroy(g->inner);
    if (g->mid) qkx_gauge_destroy(g->mid);
    if (g->outer) qkx_gauge_destroy(g->outer);
    if (g->data) (void)hipFree(g->data);
    delete g;
    return QKX_SUCCESS;
}

int qkx_gauge_upload(qkx_gauge *g, const double *U)
{
    QKX_CHECK_ARG(g && U, "qkx_gauge_upload: null argument");
    size_t bytes = sizeof(double) * 72 * (size_t)g->V;
    void *st = stage(bytes);
    if (!st) { qkx::set_error("qkx_gauge_upload: staging allocation failed"); return QKX_ERR_NOMEM; }
    QKX_HIP(hipMemcpyAsync(st, U, bytes, hipMemcpyHostToDevice, qkx::g_stream));
    const double2 *s2 = (const double2 *)st;
    // 12-real storage rebuilds row 2: as for qkx_gauge_convert, links that
    // carry the antiperiodic sign (boundaryT = +1) are stored with the last
    // time slice's t-links negated and boundaryT = -1; links that are not
    // SU(3) are refused
    int flip = 0;
    if (g->recon == QKX_RECON_12) {
        const int n = 4 * g->V, nb = (n + 255) / 256;
        double *part = nullptr;
        QKX_HIP(hipMalloc(&part, 2 * nb * sizeof(double)));
        hipLaunchKernelGGL(recon_dev_aos_kernel, dim3(nb), dim3(256), 0, qkx::g_stream, s2,
                           qkx::make_geom(g->L, g->bc), part);
        std::vector<double> h(2 * nb);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(h.data(), part, 2 * nb * sizeof(double), hipMemcpyDeviceToHost, qkx::g_stream);
        if (e == hipSuccess) e = hipStreamSynchronize(qkx::g_stream);
        (void)hipFree(part);
        QKX_HIP(e);
        double m = 0, mf = 0;
        for (int b = 0; b < nb; b++) {   // a NaN link propagates
            m = h[2 * b] > m || h[2 * b] != h[2 * b] ? h[2 * b] : m;
            mf = h[2 * b + 1] > mf || h[2 * b + 1] != h[2 * b + 1] ? h[2 * b + 1] : mf;
        }
        const double tol = g->prec == QKX_PREC_DOUBLE ? 1e-14 : 1e-6;
        if (m <= tol) {
            flip = 0;
        } else if (g->bc == 1 && mf <= tol) {
            flip = 1;
            g->bc = -1;
        } else {
            qkx::set_error("qkx_gauge_upload: 12-real field, but the links are not SU(3) (row 2 != conj(row0 x "
                           "row1) by " + std::to_string(m) + ", also with the boundary sign removed)");
            return QKX_ERR_ARG;
        }
    }
    Geom geo = qkx::make_geom(g->L, g->bc);
    dim3 grid((g->V + 255) / 256), blk(256);
    if (g->prec == QKX_PREC_DOUBLE && g->recon == 18) hipLaunchKernelGGL((gauge_from_aos<QKX_PREC_DOUBLE, 18>), grid, blk, 0, qkx::g_stream, g->data, s2, geo, flip);
    else if (g->prec == QKX_PREC_DOUBLE) hipLaunchKernelGGL((gauge_from_aos<QKX_PREC_DOUBLE, 12>), grid, blk, 0, qkx::g_stream, g->data, s2, geo, flip);
    else if (g->prec == QKX_PREC_SINGLE && g->recon == 18) hipLaunchKernelGGL((gauge_from_aos<QKX_PREC_SINGLE, 18>), grid, blk, 0, qkx::g_stream, g->data, s2, geo, flip);
    else if (g->prec == QKX_PREC_SINGLE) hipLaunchKernelGGL((gauge_from_aos<QKX_PREC_SINGLE, 12>), grid, blk, 0, qkx::g_stream, g->data, s2, geo, flip);
    else if (g->prec == QKX_PREC_HALF) hipLaunchKernelGGL((gauge_from_aos<QKX_PREC_HALF, 12>), grid, blk, 0, qkx::g_stream, g->data, s2, geo, flip);
    else hipLaunchKernelGGL((gauge_from_aos<QKX_PREC_BF16, 12>), grid, blk, 0, qkx::g_stream, g->data, s2, geo, flip);
    QKX_HIP(hipGetLastError());
    QKX_HIP(hipStreamSynchronize(qkx::g_stream));
    qkx::gauge_modified(g);
    return QKX_SUCCESS;
}

int qkx_gauge_convert(qkx_gauge *dst, const qkx_gauge *src)
{
    QKX_CHECK_ARG(dst && src, "qkx_gauge_convert: null argument");
    QKX_CHECK_ARG(src->prec == QKX_PREC_DOUBLE && src->recon == 18, "qkx_gauge_convert: source must be FP64 recon 18");
    QKX_CHECK_ARG(dst->V == src->V, "qkx_gauge_convert: lattices differ");
    // a 12-real destination rebuilds row 2 as conj(row0 x row1): links that
    // carry the antiperiodic sign are stored with the last time slice's
    // t-links negated and boundaryT = -1 (the same operator); links that are
    // not SU(3) are refused
    int mode = 0;
    if (dst->recon == QKX_RECON_12) {
        int rc = qkx::recon12_mode(src, dst->prec, &mode);
        if (rc) return rc;
        if (mode < 0) {
            qkx::set_error("qkx_gauge_convert: 12-real destination, but the links are not SU(3) (row 2 != "
                           "conj(row0 x row1) by " + std::to_string(src->recon_dev) +
                           ", also with the boundary sign removed)");
            return QKX_ERR_ARG;
        }
    }
    return qkx::gauge_convert_flip(dst, src, mode);
}

int qkx_gauge_su3_deviation(const qkx_gauge *g, double *dev)
{
    QKX_CHECK_ARG(g && dev, "qkx_gauge_su3_deviation: null argument");
    return qkx::gauge_recon_deviation(g, dev);
}

int qkx_gauge_download(const qkx_gauge *g, double *U)
{
    QKX_CHECK_ARG(g && U, "qkx_gauge_download: null argument");
    QKX_CHECK_ARG(g->prec == QKX_PREC_DOUBLE && g->recon == QKX_RECON_18,
                  "qkx_gauge_download: only FP64 recon-18 fields are downloaded");
    size_t bytes = sizeof(double) * 72 * (size_t)g->V;
    void *st = stage(bytes);
    if (!st) { qkx::set_error("qkx_gauge_download: staging allocation failed"); return QKX_ERR_NOMEM; }
    Geom geo = qkx::make_geom(g->L, g->bc);
    hipLaunchKernelGGL(gauge_to_aos_d18, dim3((g->V + 255) / 256), dim3(256), 0, qkx::g_stream, (double2 *)st,
                       (const double2 *)g->data, geo);
    QKX_HIP(hipGetLastError());
    QKX_HIP(hipMemcpyAsync(U, st, bytes, hipMemcpyDeviceToHost, qkx::g_stream));
    QKX_HIP(hipStreamSynchronize(qkx::g_stream));
    return QKX_SUCCESS;
}

} // extern "C"
