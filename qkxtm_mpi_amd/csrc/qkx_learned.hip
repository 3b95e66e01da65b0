// qkx_learned.hip -- learned preconditioner: a gauge-equivariant
// parallel-transport convolution (PTC) network, its training kernels, and the
// flexible GCR solver that uses it as a right preconditioner.
//
// No reference counterpart.  The layer is the published PTC layer (C. Lehner
// and T. Wettig, Phys. Rev. D 108, 034503 (2023), arXiv:2302.05419) with
// paths of length 0 and 1:
//   out_a(x) = sum_p sum_b W[p][a][b] (T_p in_b)(x),   p = 0..8
//   T_0 = 1,  T_{1+2mu} f(x) = U_mu(x) f(x+mu),  T_{2+2mu} f(x) = U_mu(x-mu)^dag f(x-mu)
// The network is M = L2 L1 with 4 -> ch -> 4 colour-vector channels (a spinor
// is 4 channels, one per spin).  It is linear in its input, so it is a fixed
// linear operator once trained.
//
// Feature fields: float2 (one complex) per (channel, colour, site), component
// k = channel*3 + colour at base[k*V + s], s the parity-blocked site index of
// the spinor/gauge layout.  Weights are W[9][cout][cin] complex; the master
// copy is FP32, inference reads a BF16 copy (FP32 accumulation).
#include "qkx_internal.h"
#include "qkx_core.h"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <type_traits>
#include <vector>

struct qkx_net_s {
    int ch;
    int n1, n2;            // complex weights of layer 1 ([9][ch][4]) and layer 2 ([9][4][ch])
    float2 *w;             // FP32 master weights, W1 then W2
    uint32_t *wb;          // BF16 copies (re in the low half, im in the high half)
    double2 *grad;         // accumulated gradient (dL/dRe + i dL/dIm)
    float2 *m, *v;         // Adam moments
    double *loss;          // accumulated loss
    int step;              // Adam steps taken
    int bf16;              // inference/training forward reads the BF16 copy
    int L[4];              // lattice of the work buffers
    float2 *fin, *fh, *fout, *gout, *gh;   // work buffers: 4ch, ch, 4ch, 4ch, ch channels
    double *part;
    int part_cap;
    float2 *weo;           // even-odd weight tables (forward, then adjoint), see eo_weights_kernel
    float2 *ehid;          // even-odd hidden channels on the odd sites, [3 ch][Vh][nrhs]
    size_t ehid_cap;       // float2 entries of ehid
    size_t ehid_limit;     // bound on ehid (float2 entries); 0: kEoHiddenCap
};

namespace {

constexpr int kT = 256;

__device__ __forceinline__ uint32_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);   // round to nearest even
    return u >> 16;
}
__device__ __forceinline__ float bf2f(uint32_t h) { return __uint_as_float(h << 16); }

__device__ __forceinline__ int rev_path(int p) { return p == 0 ? 0 : (((p - 1) & 1) == 0 ? p + 1 : p - 1); }

// coordinates of parity-blocked site s
__device__ __forceinline__ void coords_of(const Geom &g, int s, int x[4])
{
    const int par = s >= g.Vh ? 1 : 0;
    const int cb = s - par * g.Vh;
    const int xh = cb % g.Lh0;
    const int rest = cb / g.Lh0;
    x[1] = rest % g.L[1];
    const int zt = rest / g.L[1];
    x[2] = zt % g.L[2];
    x[3] = zt / g.L[2];
    x[0] = 2 * xh + ((par + x[1] + x[2] + x[3]) & 1);
}
__device__ __forceinline__ int site_of(const Geom &g, const int x[4])
{
    const int par = (x[0] + x[1] + x[2] + x[3]) & 1;
    return par * g.Vh + (x[0] >> 1) + g.Lh0 * (x[1] + g.L[1] * (x[2] + g.L[2] * x[3]));
}

// path p at site x: neighbour site, link site, dagger flag and boundary sign
struct Hop {
    int sn, sl;
    bool dag;
    float sign;
};
__device__ __forceinline__ Hop path_hop(const Geom &g, const int x[4], int s, int p)
{
    Hop h{s, s, false, 1.0f};
    if (p == 0) return h;
    const int mu = (p - 1) >> 1;
    int y[4] = {x[0], x[1], x[2], x[3]};
    if (((p - 1) & 1) == 0) {   // forward
        y[mu] = (x[mu] + 1 == g.L[mu]) ? 0 : x[mu] + 1;
        h.sn = site_of(g, y);
        h.sl = s;
        h.dag = false;
        if (mu == 3 && x[3] + 1 == g.L[3]) h.sign = (float)g.bc;
    } else {                    // backward
        y[mu] = (x[mu] == 0) ? g.L[mu] - 1 : x[mu] - 1;
        h.sn = site_of(g, y);
        h.sl = h.sn;
        h.dag = true;
        if (mu == 3 && x[3] == 0) h.sign = (float)g.bc;
    }
    return h;
}

__device__ __forceinline__ void load_link(const GaugeF18 &U, const Hop &h, int p, Cx<float> u[9])
{
    const int mu = (p - 1) >> 1;
    Cx<float> l[9];
    U.load(mu, h.sl, l);
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const Cx<float> e = h.dag ? Cx<float>{l[c * 3 + r].re, -l[c * 3 + r].im} : l[r * 3 + c];
            u[r * 3 + c] = {h.sign * e.re, h.sign * e.im};
        }
}

__device__ __forceinline__ void matvec(const Cx<float> u[9], const Cx<float> v[3], Cx<float> o[3])
{
#pragma unroll
    for (int r = 0; r < 3; r++) {
        Cx<float> a = cmul(u[r * 3], v[0]);
        a = a + cmul(u[r * 3 + 1], v[1]);
        a = a + cmul(u[r * 3 + 2], v[2]);
        o[r] = a;
    }
}

__device__ __forceinline__ Cx<float> ldf(const float2 *f, int k, int V, int s)
{
    const float2 v = f[(size_t)k * V + s];
    return {v.x, v.y};
}

// One PTC layer (CIN -> COUT channels), or the adjoint of the layer whose
// weights are given (ADJ: A[p][a][b] = conj(W[rev p][b][a]), W of shape
// [9][CIN][COUT]).  One thread per site; the input channels are streamed so
// that only the COUT x 3 accumulators stay live.
template <int CIN, int COUT, bool ADJ, bool BF>
__global__ __launch_bounds__(kT) void ptc_kernel(const float2 *__restrict__ in, float2 *__restrict__ out, GaugeF18 U,
                                                 Geom g, const void *__restrict__ wsrc)
{
    __shared__ float2 w[9 * COUT * CIN];
    for (int i = threadIdx.x; i < 9 * COUT * CIN; i += blockDim.x) {
        const int p = i / (COUT * CIN), a = (i / CIN) % COUT, b = i % CIN;
        const int src = ADJ ? (rev_path(p) * CIN + b) * COUT + a : i;
        float2 v;
        if (BF) {
            const uint32_t h = ((const uint32_t *)wsrc)[src];
            v = make_float2(bf2f(h & 0xffffu), bf2f(h >> 16));
        } else {
            v = ((const float2 *)wsrc)[src];
        }
        if (ADJ) v.y = -v.y;
        w[i] = v;
    }
    __syncthreads();
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.V) return;
    const int V = g.V;
    int x[4];
    coords_of(g, s, x);
    Cx<float> acc[COUT][3];
#pragma unroll
    for (int a = 0; a < COUT; a++)
#pragma unroll
        for (int c = 0; c < 3; c++) acc[a][c] = {0.f, 0.f};

    for (int p = 0; p < 9; p++) {
        const Hop h = path_hop(g, x, s, p);
        Cx<float> u[9];
        if (p > 0) load_link(U, h, p, u);
        if (CIN <= COUT) {
            // transport each input channel, then mix into all outputs
#pragma unroll
            for (int b = 0; b < CIN; b++) {
                Cx<float> v[3], t[3];
#pragma unroll
                for (int c = 0; c < 3; c++) v[c] = ldf(in, b * 3 + c, V, h.sn);
                if (p > 0) matvec(u, v, t);
                else {
#pragma unroll
                    for (int c = 0; c < 3; c++) t[c] = v[c];
                }
#pragma unroll
                for (int a = 0; a < COUT; a++) {
                    const float2 wv = w[(p * COUT + a) * CIN + b];
                    const Cx<float> wc{wv.x, wv.y};
#pragma unroll
                    for (int c = 0; c < 3; c++) acc[a][c] = acc[a][c] + cmul(wc, t[c]);
                }
            }
        } else {
            // mix the input channels at the neighbour, then transport COUT vectors
            Cx<float> m[COUT][3];
#pragma unroll
            for (int a = 0; a < COUT; a++)
#pragma unroll
                for (int c = 0; c < 3; c++) m[a][c] = {0.f, 0.f};
#pragma unroll
            for (int b = 0; b < CIN; b++) {
                Cx<float> v[3];
#pragma unroll
                for (int c = 0; c < 3; c++) v[c] = ldf(in, b * 3 + c, V, h.sn);
#pragma unroll
                for (int a = 0; a < COUT; a++) {
                    const float2 wv = w[(p * COUT + a) * CIN + b];
                    const Cx<float> wc{wv.x, wv.y};
#pragma unroll
                    for (int c = 0; c < 3; c++) m[a][c] = m[a][c] + cmul(wc, v[c]);
                }
            }
#pragma unroll
            for (int a = 0; a < COUT; a++) {
                Cx<float> t[3];
                if (p > 0) matvec(u, m[a], t);
                else {
#pragma unroll
                    for (int c = 0; c < 3; c++) t[c] = m[a][c];
                }
#pragma unroll
                for (int c = 0; c < 3; c++) acc[a][c] = acc[a][c] + t[c];
            }
        }
    }
#pragma unroll
    for (int a = 0; a < COUT; a++)
#pragma unroll
        for (int c = 0; c < 3; c++) out[(size_t)(a * 3 + c) * V + s] = make_float2(acc[a][c].re, acc[a][c].im);
}

// Weight gradient of a layer CIN -> COUT: G[p][a][b] = sum_x sum_c g_a(x)_c
// conj((T_p in_b)(x)_c).  grid = (blocks over sites, 9 paths, 4x4 tiles of
// (a, b)); per-workgroup partials, summed in a fixed order by wgrad_finish.
template <int CIN, int COUT>
__global__ __launch_bounds__(kT) void ptc_wgrad_kernel(const float2 *__restrict__ gin, const float2 *__restrict__ in,
                                                       GaugeF18 U, Geom geo, double *__restrict__ part)
{
    const int p = blockIdx.y;
    const int a0 = (blockIdx.z / (CIN / 4)) * 4, b0 = (blockIdx.z % (CIN / 4)) * 4;
    const int V = geo.V;
    float acc[4][4][2];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j][0] = acc[i][j][1] = 0.f;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < V; s += gridDim.x * blockDim.x) {
        int x[4];
        coords_of(geo, s, x);
        const Hop h = path_hop(geo, x, s, p);
        Cx<float> u[9];
        if (p > 0) load_link(U, h, p, u);
        Cx<float> t[4][3];
#pragma unroll
        for (int bb = 0; bb < 4; bb++) {
            Cx<float> v[3];
#pragma unroll
            for (int c = 0; c < 3; c++) v[c] = ldf(in, (b0 + bb) * 3 + c, V, h.sn);
            if (p > 0) matvec(u, v, t[bb]);
            else {
#pragma unroll
                for (int c = 0; c < 3; c++) t[bb][c] = v[c];
            }
        }
#pragma unroll
        for (int aa = 0; aa < 4; aa++) {
            Cx<float> gv[3];
#pragma unroll
            for (int c = 0; c < 3; c++) gv[c] = ldf(gin, (a0 + aa) * 3 + c, V, s);
#pragma unroll
            for (int bb = 0; bb < 4; bb++) {
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const Cx<float> z = cjmul(t[bb][c], gv[c]);   // g * conj(t)
                    acc[aa][bb][0] += z.re;
                    acc[aa][bb][1] += z.im;
                }
            }
        }
    }
    // workgroup sum of the 32 values: wavefront butterflies, then LDS
    __shared__ float red[kT / 64][32];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int k = 0; k < 2; k++) {
                float v = acc[i][j][k];
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
                if (lane == 0) red[wv][(i * 4 + j) * 2 + k] = v;
            }
    __syncthreads();
    if (threadIdx.x < 32) {
        double s = 0;
        for (int q = 0; q < (int)(blockDim.x >> 6); q++) s += red[q][threadIdx.x];
        const int i = threadIdx.x >> 3, j = (threadIdx.x >> 1) & 3, k = threadIdx.x & 1;
        const int widx = (p * COUT + a0 + i) * CIN + b0 + j;
        part[((size_t)blockIdx.x * 9 * COUT * CIN + widx) * 2 + k] = s;
    }
}

__global__ void wgrad_finish_kernel(const double *__restrict__ part, int nblocks, int nw, double2 *grad)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nw) return;
    double re = 0, im = 0;
    for (int b = 0; b < nblocks; b++) {
        re += part[((size_t)b * nw + i) * 2];
        im += part[((size_t)b * nw + i) * 2 + 1];
    }
    grad[i].x += re;
    grad[i].y += im;
}

// loss of one pair: d = o - e, gout = 2 d, per-workgroup |d|^2
__global__ __launch_bounds__(kT) void loss_kernel(const float2 *__restrict__ o, const float2 *__restrict__ e,
                                                  float2 *__restrict__ gout, int n, double *part)
{
    double acc = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float2 a = o[i], b = e[i];
        const float dx = a.x - b.x, dy = a.y - b.y;
        gout[i] = make_float2(2.f * dx, 2.f * dy);
        acc += (double)dx * dx + (double)dy * dy;
    }
    __shared__ double red[kT / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0;
        for (int q = 0; q < (int)(blockDim.x >> 6); q++) s += red[q];
        part[blockIdx.x] = s;
    }
}

__global__ void loss_finish_kernel(const double *part, int nblocks, double *loss)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0;
    for (int b = 0; b < nblocks; b++) s += part[b];
    *loss += s;
}

__global__ void adam_kernel(float2 *w, uint32_t *wb, double2 *grad, float2 *m, float2 *v, int n, float lr,
                            float b1, float b2, float eps, float c1, float c2, float scale)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gx = (float)grad[i].x * scale, gy = (float)grad[i].y * scale;
    float2 mm = m[i], vv = v[i], ww = w[i];
    mm.x = b1 * mm.x + (1.f - b1) * gx;
    mm.y = b1 * mm.y + (1.f - b1) * gy;
    vv.x = b2 * vv.x + (1.f - b2) * gx * gx;
    vv.y = b2 * vv.y + (1.f - b2) * gy * gy;
    ww.x -= lr * (mm.x / c1) / (sqrtf(vv.x / c2) + eps);
    ww.y -= lr * (mm.y / c1) / (sqrtf(vv.y / c2) + eps);
    m[i] = mm;
    v[i] = vv;
    w[i] = ww;
    wb[i] = f2bf(ww.x) | (f2bf(ww.y) << 16);
    grad[i] = make_double2(0.0, 0.0);
}

__global__ void to_bf16_kernel(const float2 *w, uint32_t *wb, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) wb[i] = f2bf(w[i].x) | (f2bf(w[i].y) << 16);
}

// spinor field (accessor S, full lattice) <-> 4-channel feature field
template <class S> __global__ __launch_bounds__(kT) void spinor_to_feat(S f, float2 *out, int V)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= V) return;
    Cx<typename S::Real> v[12];
    f.load(s, v);
#pragma unroll
    for (int k = 0; k < 12; k++) out[(size_t)k * V + s] = make_float2((float)v[k].re, (float)v[k].im);
}
template <class S> __global__ __launch_bounds__(kT) void feat_to_spinor(const float2 *in, S f, int V)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= V) return;
    using R = typename S::Real;
    Cx<R> v[12];
#pragma unroll
    for (int k = 0; k < 12; k++) {
        const float2 a = in[(size_t)k * V + s];
        v[k] = {(R)a.x, (R)a.y};
    }
    f.store(s, v);
}

// ---------------------------------------------------------------------------
// The network on even-parity fields, N_e = P_e L2 L1 P_e (P_e: restriction to
// the even sites), batched over right-hand sides (RHS innermost).  With the
// input zero on the odd sites, L1's hidden channels are W1[0] in(x) at an even
// site x and sum_{p>0} W1[p] (T_p in)(y) at an odd site y (all neighbours of y
// are even), and L2's output at x is
//   C0 in(x) + sum_{p>0} W2[p] (T_p h)(x),   C0 = W2[0] W1[0]  (4 x 4),
// whose hops read only odd-site hidden channels.  So N_e is two launches:
// eo_hidden_kernel (odd sites) and eo_out_kernel (even sites).  The adjoint
// N_e^dag = P_e L1^dag L2^dag P_e has the same form with the adjoint layers'
// weights, A[p][a][b] = conj(W[rev p][b][a]).
//
// Weight table of one direction (forward or adjoint), complex:
//   [8][CH][4] first-layer hops (p = 1..8), [8][4][CH] second-layer hops,
//   [4][4] C0.
constexpr int eo_table(int ch) { return 8 * ch * 4 * 2 + 16; }

__device__ __forceinline__ float2 net_w(const float2 *w, const uint32_t *wb, int bf, int i)
{
    if (!bf) return w[i];
    const uint32_t h = wb[i];
    return make_float2(bf2f(h & 0xffffu), bf2f(h >> 16));
}

// Both tables from the FP32 master weights or their BF16 copies: thread i
// writes entry i of [forward table][adjoint table].
__global__ void eo_weights_kernel(const float2 *w, const uint32_t *wb, int bf, int ch, float2 *out)
{
    const int T = eo_table(ch), i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 2 * T) return;
    const int adj = i / T, j = i % T, n1 = 9 * ch * 4;
    auto W1 = [&](int p, int a, int b) { return net_w(w, wb, bf, (p * ch + a) * 4 + b); };    // [9][ch][4]
    auto W2 = [&](int p, int a, int b) { return net_w(w, wb, bf, n1 + (p * 4 + a) * ch + b); };   // [9][4][ch]
    auto cj = [](float2 v) { return make_float2(v.x, -v.y); };
    const int h1 = 8 * ch * 4;
    float2 v;
    if (j < h1) {                       // first layer, p = 1..8: [a < ch][b < 4]
        const int p = j / (ch * 4) + 1, a = (j / 4) % ch, b = j % 4;
        v = adj ? cj(W2(rev_path(p), b, a)) : W1(p, a, b);
    } else if (j < 2 * h1) {            // second layer: [a < 4][b < ch]
        const int k = j - h1, p = k / (4 * ch) + 1, a = (k / ch) % 4, b = k % ch;
        v = adj ? cj(W1(rev_path(p), b, a)) : W2(p, a, b);
    } else {                            // C0 = W2[0] W1[0], or its adjoint
        const int k = j - 2 * h1, a = adj ? k % 4 : k / 4, b = adj ? k / 4 : k % 4;
        float re = 0.f, im = 0.f;
        for (int c = 0; c < ch; c++) {
            const float2 x = W2(0, a, c), y = W1(0, c, b);
            re += x.x * y.x - x.y * y.y;
            im += x.x * y.y + x.y * y.x;
        }
        v = adj ? make_float2(re, -im) : make_float2(re, im);
    }
    out[i] = v;
}

// Hidden channels at the odd sites: h_a(y) = sum_{p>0} sum_b w[p][a][b] (T_p in_b)(y).
// One thread per (odd site, RHS); in is an even-parity field.
template <int CH, class S>
__global__ __launch_bounds__(kT) void eo_hidden_kernel(S in, GaugeF18 U, Geom g, const float2 *__restrict__ wsrc,
                                                       float2 *__restrict__ hid, int nr, int r0, int nc,
                                                       const int *skip)
{
    if (skip && *skip) return;   // uniform over the grid
    __shared__ float2 w[8 * CH * 4];
    for (int i = threadIdx.x; i < 8 * CH * 4; i += blockDim.x) w[i] = wsrc[i];
    __syncthreads();
    // units of this launch: (odd site, rhs r0 .. r0 + nc - 1) of a batch of nr
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= g.Vh * nc) return;
    const int cb = u / nc, r = r0 + (u - cb * nc);
    const int s = g.Vh + cb;
    int x[4];
    coords_of(g, s, x);
    Cx<float> acc[CH][3];
#pragma unroll
    for (int a = 0; a < CH; a++)
#pragma unroll
        for (int c = 0; c < 3; c++) acc[a][c] = {0.f, 0.f};
    for (int p = 1; p < 9; p++) {
        const Hop h = path_hop(g, x, s, p);   // h.sn: an even site, = its checkerboard index
        Cx<float> uu[9];
        load_link(U, h, p, uu);
        Cx<typename S::Real> v[12];
        in.load(h.sn * nr + r, v);
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const Cx<float> vb[3] = {{(float)v[b * 3].re, (float)v[b * 3].im},
                                     {(float)v[b * 3 + 1].re, (float)v[b * 3 + 1].im},
                                     {(float)v[b * 3 + 2].re, (float)v[b * 3 + 2].im}};
            Cx<float> t[3];
            matvec(uu, vb, t);
#pragma unroll
            for (int a = 0; a < CH; a++) {
                const float2 wv = w[((p - 1) * CH + a) * 4 + b];
                const Cx<float> wc{wv.x, wv.y};
#pragma unroll
                for (int c = 0; c < 3; c++) acc[a][c] = acc[a][c] + cmul(wc, t[c]);
            }
        }
    }
    const size_t stride = (size_t)g.Vh * nc;
#pragma unroll
    for (int a = 0; a < CH; a++)
#pragma unroll
        for (int c = 0; c < 3; c++) hid[(size_t)(a * 3 + c) * stride + u] = make_float2(acc[a][c].re, acc[a][c].im);
}

// Output at the even sites: out(x) = C0 in(x) + sum_{p>0} (T_p m_p)(x) with
// m_p = w[p] h at the odd neighbour (the channels are mixed first, then the
// four colour vectors transported).  One thread per (even site, RHS).
template <int CH, class S>
__global__ __launch_bounds__(kT) void eo_out_kernel(S in, S out, GaugeF18 U, Geom g, const float2 *__restrict__ wsrc,
                                                    const float2 *__restrict__ hid, int nr, int r0, int nc,
                                                    const int *skip)
{
    if (skip && *skip) return;   // uniform over the grid
    __shared__ float2 w[8 * 4 * CH + 16];
    for (int i = threadIdx.x; i < 8 * 4 * CH + 16; i += blockDim.x) w[i] = wsrc[i];
    __syncthreads();
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= g.Vh * nc) return;
    const int cb = u / nc, rc = u - cb * nc;
    const int e = cb * nr + r0 + rc;   // the unit's entry in the batch fields
    int x[4];
    coords_of(g, cb, x);
    Cx<float> acc[4][3];
    {
        Cx<typename S::Real> v[12];
        in.load(e, v);
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int c = 0; c < 3; c++) {
                Cx<float> z{0.f, 0.f};
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    const float2 wv = w[8 * 4 * CH + a * 4 + b];
                    z = z + cmul(Cx<float>{wv.x, wv.y}, Cx<float>{(float)v[b * 3 + c].re, (float)v[b * 3 + c].im});
                }
                acc[a][c] = z;
            }
    }
    const size_t stride = (size_t)g.Vh * nc;
    for (int p = 1; p < 9; p++) {
        const Hop h = path_hop(g, x, cb, p);   // h.sn: an odd site
        const size_t hu = (size_t)(h.sn - g.Vh) * nc + rc;
        Cx<float> m[4][3];
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int c = 0; c < 3; c++) m[a][c] = {0.f, 0.f};
#pragma unroll
        for (int k = 0; k < CH; k++) {
            Cx<float> hv[3];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const float2 t = hid[(size_t)(k * 3 + c) * stride + hu];
                hv[c] = {t.x, t.y};
            }
#pragma unroll
            for (int a = 0; a < 4; a++) {
                const float2 wv = w[((p - 1) * 4 + a) * CH + k];
                const Cx<float> wc{wv.x, wv.y};
#pragma unroll
                for (int c = 0; c < 3; c++) m[a][c] = m[a][c] + cmul(wc, hv[c]);
            }
        }
        Cx<float> uu[9];
        load_link(U, h, p, uu);
#pragma unroll
        for (int a = 0; a < 4; a++) {
            Cx<float> t[3];
            matvec(uu, m[a], t);
#pragma unroll
            for (int c = 0; c < 3; c++) acc[a][c] = acc[a][c] + t[c];
        }
    }
    using R = typename S::Real;
    Cx<R> o[12];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int c = 0; c < 3; c++) o[a * 3 + c] = {(R)acc[a][c].re, (R)acc[a][c].im};
    out.store(e, o);
}

int nblk(int n) { return (n + kT - 1) / kT; }

int net_buffers(qkx_net_s *n, const int L[4])
{
    if (n->fin && n->L[0] == L[0] && n->L[1] == L[1] && n->L[2] == L[2] && n->L[3] == L[3]) return QKX_SUCCESS;
    float2 **bufs[5] = {&n->fin, &n->fh, &n->fout, &n->gout, &n->gh};
    for (auto b : bufs) {
        if (*b) (void)hipFree(*b);
        *b = nullptr;
    }
    const size_t V = (size_t)L[0] * L[1] * L[2] * L[3];
    const int chans[5] = {4, n->ch, 4, 4, n->ch};
    for (int i = 0; i < 5; i++) QKX_HIP(hipMalloc(bufs[i], sizeof(float2) * 3 * chans[i] * V));
    for (int d = 0; d < 4; d++) n->L[d] = L[d];
    const int need = (int)(nblk((int)V) * 9 * 4 * n->ch * 2 + 1024);
    if (need > n->part_cap) {
        if (n->part) (void)hipFree(n->part);
        n->part = nullptr;
        QKX_HIP(hipMalloc(&n->part, sizeof(double) * need));
        n->part_cap = need;
    }
    return QKX_SUCCESS;
}

int check_net_gauge(const qkx_gauge_s *U)
{
    if (U->prec != QKX_PREC_SINGLE || U->recon != QKX_RECON_18) {
        qkx::set_error("net: the gauge field must be FP32 with 18 reals");
        return QKX_ERR_ARG;
    }
    return QKX_SUCCESS;
}

template <int CIN, int COUT, bool ADJ>
int launch_ptc(const float2 *in, float2 *out, const qkx_gauge_s *U, const void *w, bool bf)
{
    Geom g = qkx::make_geom(U->L, U->bc);
    GaugeF18 G = make_gauge<GaugeF18>(U);
    if (bf)
        hipLaunchKernelGGL((ptc_kernel<CIN, COUT, ADJ, true>), dim3(nblk(g.V)), dim3(kT), 0, qkx::g_stream, in, out,
                           G, g, w);
    else
        hipLaunchKernelGGL((ptc_kernel<CIN, COUT, ADJ, false>), dim3(nblk(g.V)), dim3(kT), 0, qkx::g_stream, in, out,
                           G, g, w);
    QKX_HIP(hipGetLastError());
    return QKX_SUCCESS;
}

// forward of both layers from n->fin into n->fh and n->fout
int net_forward(qkx_net_s *n, const qkx_gauge_s *U)
{
    const bool bf = n->bf16 != 0;
    const void *w1 = bf ? (const void *)n->wb : (const void *)n->w;
    const void *w2 = bf ? (const void *)(n->wb + n->n1) : (const void *)(n->w + n->n1);
    int rc;
    if (n->ch == 16) {
        if ((rc = launch_ptc<4, 16, false>(n->fin, n->fh, U, w1, bf))) return rc;
        return launch_ptc<16, 4, false>(n->fh, n->fout, U, w2, bf);
    }
    if (n->ch == 4) {
        if ((rc = launch_ptc<4, 4, false>(n->fin, n->fh, U, w1, bf))) return rc;
        return launch_ptc<4, 4, false>(n->fh, n->fout, U, w2, bf);
    }
    if ((rc = launch_ptc<4, 8, false>(n->fin, n->fh, U, w1, bf))) return rc;
    return launch_ptc<8, 4, false>(n->fh, n->fout, U, w2, bf);
}

template <int CIN, int COUT>
int launch_wgrad(const float2 *g, const float2 *in, const qkx_gauge_s *U, qkx_net_s *n, double2 *grad)
{
    Geom geo = qkx::make_geom(U->L, U->bc);
    const int nbx = std::min(nblk(geo.V), 128);
    hipLaunchKernelGGL((ptc_wgrad_kernel<CIN, COUT>), dim3(nbx, 9, (COUT / 4) * (CIN / 4)), dim3(kT), 0,
                       qkx::g_stream, g, in, make_gauge<GaugeF18>(U), geo, n->part);
    const int nw = 9 * COUT * CIN;
    hipLaunchKernelGGL(wgrad_finish_kernel, dim3((nw + 255) / 256), dim3(256), 0, qkx::g_stream, n->part, nbx, nw,
                       grad);
    QKX_HIP(hipGetLastError());
    return QKX_SUCCESS;
}

template <class F> int with_spinor(const qkx_spinor_s *f, F fn)
{
    return dispatch_prec(f->prec, [&](auto tag) { return fn(make_spinor<decltype(tag)>(f)); });
}

int check_spinor(const qkx_spinor_s *f, const qkx_gauge_s *U)
{
    if (f->subset != QKX_SUBSET_FULL || f->nrhs != 1 || (f->prec != QKX_PREC_DOUBLE && f->prec != QKX_PREC_SINGLE) ||
        f->L[0] != U->L[0] || f->L[1] != U->L[1] || f->L[2] != U->L[2] || f->L[3] != U->L[3]) {
        qkx::set_error("net: fields must be single full-lattice FP64/FP32 spinors on the gauge's lattice");
        return QKX_ERR_ARG;
    }
    return QKX_SUCCESS;
}

int to_feat(const qkx_spinor_s *f, float2 *out)
{
    const int V = f->V;
    return with_spinor(f, [&](auto S) {
        if constexpr (std::is_same_v<decltype(S), SpinorD> || std::is_same_v<decltype(S), SpinorF>) {
            hipLaunchKernelGGL((spinor_to_feat<decltype(S)>), dim3(nblk(V)), dim3(kT), 0, qkx::g_stream, S, out, V);
            QKX_HIP(hipGetLastError());
            return QKX_SUCCESS;
        } else {
            qkx::set_error("net: FP64/FP32 fields only");
            return QKX_ERR_ARG;
        }
    });
}

int from_feat(const float2 *in, qkx_spinor_s *f)
{
    const int V = f->V;
    return with_spinor(f, [&](auto S) {
        if constexpr (std::is_same_v<decltype(S), SpinorD> || std::is_same_v<decltype(S), SpinorF>) {
            hipLaunchKernelGGL((feat_to_spinor<decltype(S)>), dim3(nblk(V)), dim3(kT), 0, qkx::g_stream, in, S, V);
            QKX_HIP(hipGetLastError());
            return QKX_SUCCESS;
        } else {
            qkx::set_error("net: FP64/FP32 fields only");
            return QKX_ERR_ARG;
        }
    });
}

// right-hand sides per launch of the even-odd network: all of them while
// their hidden channels fit the bound (default kEoHiddenCap float2 entries,
// 4 GiB), else fewer
constexpr size_t kEoHiddenCap = size_t(1) << 29;
int eo_chunk(const qkx_net_s *n, const qkx_spinor_s *in)
{
    const size_t per_rhs = (size_t)n->ch * 3 * (size_t)in->Vh;
    const size_t cap = n->ehid_limit ? n->ehid_limit : kEoHiddenCap;
    return (int)std::max<size_t>(1, std::min<size_t>((size_t)in->nrhs, cap / per_rhs));
}

template <int CH, class S>
int launch_eo(qkx_net_s *n, qkx_spinor_s *out, const qkx_spinor_s *in, const qkx_gauge_s *U, int adjoint,
              const int *skip)
{
    const Geom g = qkx::make_geom(U->L, U->bc);
    const int nr = in->nrhs, nc = eo_chunk(n, in);
    const GaugeF18 G = make_gauge<GaugeF18>(U);
    const float2 *tab = n->weo + (adjoint ? eo_table(CH) : 0);
    // the right-hand sides in chunks of nc, so that the hidden channels of a
    // chunk fit the buffer (out may be in: each chunk reads only its own
    // columns before writing them)
    for (int r0 = 0; r0 < nr; r0 += nc) {
        const int c = std::min(nc, nr - r0), units = g.Vh * c;
        hipLaunchKernelGGL((eo_hidden_kernel<CH, S>), dim3(nblk(units)), dim3(kT), 0, qkx::g_stream,
                           make_spinor<S>(in), G, g, tab, n->ehid, nr, r0, c, skip);
        hipLaunchKernelGGL((eo_out_kernel<CH, S>), dim3(nblk(units)), dim3(kT), 0, qkx::g_stream, make_spinor<S>(in),
                           make_spinor<S>(out), G, g, tab + 8 * CH * 4, (const float2 *)n->ehid, nr, r0, c, skip);
    }
    QKX_HIP(hipGetLastError());
    return QKX_SUCCESS;
}

} // namespace

namespace qkx {
// M in -> out (FP64 or FP32 full spinors); used by the flexible solver
int net_apply(qkx_net_s *n, qkx_spinor_s *out, const qkx_spinor_s *in, const qkx_gauge_s *U)
{
    int rc;
    if ((rc = check_net_gauge(U)) || (rc = check_spinor(in, U)) || (rc = check_spinor(out, U))) return rc;
    if ((rc = net_buffers(n, U->L))) return rc;
    if ((rc = to_feat(in, n->fin))) return rc;
    if ((rc = net_forward(n, U))) return rc;
    return from_feat(n->fout, out);
}
// The even-odd weight tables from the current weights (BF16 copies when
// the network runs in BF16); a solve prepares them once.
int net_prep_eo(qkx_net_s *n)
{
    const int T = eo_table(n->ch);
    if (!n->weo) QKX_HIP(hipMalloc(&n->weo, sizeof(float2) * 2 * T));
    hipLaunchKernelGGL(eo_weights_kernel, dim3((2 * T + 255) / 256), dim3(256), 0, qkx::g_stream, n->w, n->wb,
                       n->bf16, n->ch, n->weo);
    QKX_HIP(hipGetLastError());
    return QKX_SUCCESS;
}

// out = N_e in (adjoint: N_e^dag in) on even-parity fields of one precision
// and batch size; net_prep_eo must have run since the weights last changed.
// out may be in.  skip (optional, device): both launches return at once
// while *skip != 0.
int net_apply_eo(qkx_net_s *n, qkx_spinor_s *out, const qkx_spinor_s *in, const qkx_gauge_s *U, int adjoint,
                 const int *skip)
{
    int rc;
    if ((rc = check_net_gauge(U))) return rc;
    if (in->subset != QKX_SUBSET_EVEN || out->subset != QKX_SUBSET_EVEN || in->prec != out->prec ||
        in->nrhs != out->nrhs || in->L[0] != U->L[0] || in->L[1] != U->L[1] || in->L[2] != U->L[2] ||
        in->L[3] != U->L[3] || out->L[0] != U->L[0] || out->L[1] != U->L[1] || out->L[2] != U->L[2] ||
        out->L[3] != U->L[3]) {
        qkx::set_error("net (even-odd): in and out must be even-parity fields of one precision and batch size on "
                       "the gauge's lattice");
        return QKX_ERR_ARG;
    }
    const size_t need = (size_t)n->ch * 3 * (size_t)in->Vh * eo_chunk(n, in);
    if (need > n->ehid_cap) {
        if (n->ehid) (void)hipFree(n->ehid);
        n->ehid = nullptr;
        n->ehid_cap = 0;
        QKX_HIP(hipMalloc(&n->ehid, sizeof(float2) * need));
        n->ehid_cap = need;
    }
    return dispatch_prec(in->prec, [&](auto tag) {
        using S = decltype(tag);
        if (n->ch == 16) return launch_eo<16, S>(n, out, in, U, adjoint, skip);
        if (n->ch == 4) return launch_eo<4, S>(n, out, in, U, adjoint, skip);
        return launch_eo<8, S>(n, out, in, U, adjoint, skip);
    });
}
} // namespace qkx

extern "C" {

int qkx_net_create(qkx_net **out, int ch)
{
    QKX_CHECK_ARG(out && (ch == 4 || ch == 8 || ch == 16), "qkx_net_create: hidden channels must be 4, 8 or 16");
    if (int rc = qkx::ensure_init()) return rc;
    qkx_net_s *n = new qkx_net_s{};
    n->ch = ch;
    n->n1 = 9 * ch * 4;
    n->n2 = 9 * 4 * ch;
    n->bf16 = 1;
    const int nt = n->n1 + n->n2;
    bool ok = hipMalloc(&n->w, sizeof(float2) * nt) == hipSuccess &&
              hipMalloc(&n->wb, sizeof(uint32_t) * nt) == hipSuccess &&
              hipMalloc(&n->grad, sizeof(double2) * nt) == hipSuccess &&
              hipMalloc(&n->m, sizeof(float2) * nt) == hipSuccess &&
              hipMalloc(&n->v, sizeof(float2) * nt) == hipSuccess &&
              hipMalloc(&n->loss, sizeof(double)) == hipSuccess;
    if (!ok) {
        qkx_net_destroy(n);
        qkx::set_error("qkx_net_create: device allocation failed");
        return QKX_ERR_NOMEM;
    }
    QKX_HIP(hipMemsetAsync(n->w, 0, sizeof(float2) * nt, qkx::g_stream));
    QKX_HIP(hipMemsetAsync(n->wb, 0, sizeof(uint32_t) * nt, qkx::g_stream));
    QKX_HIP(hipMemsetAsync(n->grad, 0, sizeof(double2) * nt, qkx::g_stream));
    QKX_HIP(hipMemsetAsync(n->m, 0, sizeof(float2) * nt, qkx::g_stream));
    QKX_HIP(hipMemsetAsync(n->v, 0, sizeof(float2) * nt, qkx::g_stream));
    QKX_HIP(hipMemsetAsync(n->loss, 0, sizeof(double), qkx::g_stream));
    QKX_HIP(hipStreamSynchronize(qkx::g_stream));
    *out = n;
    return QKX_SUCCESS;
}

int qkx_net_destroy(qkx_net *n)
{
    if (!n) return QKX_SUCCESS;
    void *ptrs[] = {n->w,    n->wb, n->grad, n->m,  n->v,    n->loss, n->fin,
                    n->fh,   n->fout, n->gout, n->gh, n->part, n->weo, n->ehid};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    delete n;
    return QKX_SUCCESS;
}

int qkx_net_release_buffers(qkx_net *n)
{
    QKX_CHECK_ARG(n, "qkx_net_release_buffers: null argument");
    QKX_HIP(hipStreamSynchronize(qkx::g_stream));   // no launch may still use them
    float2 **bufs[6] = {&n->fin, &n->fh, &n->fout, &n->gout, &n->gh, &n->ehid};
    for (auto b : bufs) {
        if (*b) (void)hipFree(*b);
        *b = nullptr;
    }
    n->ehid_cap = 0;
    n->L[0] = n->L[1] = n->L[2] = n->L[3] = 0;
    return QKX_SUCCESS;
}

int qkx_net_set_hidden_limit(qkx_net *n, unsigned long long bytes)
{
    QKX_CHECK_ARG(n, "qkx_net_set_hidden_limit: null argument");
    n->ehid_limit = (size_t)(bytes / sizeof(float2));
    return QKX_SUCCESS;
}

int qkx_net_nweights(const qkx_net *n, int *ncomplex)
{
    QKX_CHECK_ARG(n && ncomplex, "qkx_net_nweights: null argument");
    *ncomplex = n->n1 + n->n2;
    return QKX_SUCCESS;
}

int qkx_net_set_weights(qkx_net *n, const float *w)
{
    QKX_CHECK_ARG(n && w, "qkx_net_set_weights: null argument");
    const int nt = n->n1 + n->n2;
    QKX_HIP(hipMemcpyAsync(n->w, w, sizeof(float2) * nt, hipMemcpyHostToDevice, qkx::g_stream));
    hipLaunchKernelGGL(to_bf16_kernel, dim3((nt + 255) / 256), dim3(256), 0, qkx::g_stream, n->w, n->wb, nt);
    QKX_HIP(hipGetLastError());
    // a new parameter vector restarts the optimiser
    QKX_HIP(hipMemsetAsync(n->m, 0, sizeof(float2) * nt, qkx::g_stream));
    QKX_HIP(hipMemsetAsync(n->v, 0, sizeof(float2) * nt, qkx::g_stream));
    n->step = 0;
    QKX_HIP(hipStreamSynchronize(qkx::g_stream));
    return QKX_SUCCESS;
}

int qkx_net_get_weights(const qkx_net *n, float *w)
{
    QKX_CHECK_ARG(n && w, "qkx_net_get_weights: null argument");
    QKX_HIP(hipMemcpyAsync(w, n->w, sizeof(float2) * (n->n1 + n->n2), hipMemcpyDeviceToHost, qkx::g_stream));
    QKX_HIP(hipStreamSynchronize(qkx::g_stream));
    return QKX_SUCCESS;
}

int qkx_net_set_bf16(qkx_net *n, int on)
{
    QKX_CHECK_ARG(n, "qkx_net_set_bf16: null argument");
    n->bf16 = on ? 1 : 0;
    return QKX_SUCCESS;
}

int qkx_net_apply(qkx_net *n, qkx_spinor *out, const qkx_spinor *in, const qkx_gauge *U)
{
    QKX_CHECK_ARG(n && out && in && U, "qkx_net_apply: null argument");
    QKX_CHECK_ARG(out != in, "qkx_net_apply: out must differ from in");
    return qkx::net_apply(n, out, in, U);
}

int qkx_net_apply_eo(qkx_net *n, qkx_spinor *out, const qkx_spinor *in, const qkx_gauge *U, int adjoint)
{
    QKX_CHECK_ARG(n && out && in && U, "qkx_net_apply_eo: null argument");
    int rc = qkx::net_prep_eo(n);
    if (rc) return rc;
    return qkx::net_apply_eo(n, out, in, U, adjoint ? 1 : 0, nullptr);
}

int qkx_net_accumulate_grad(qkx_net *n, const qkx_spinor *r, const qkx_spinor *e, const qkx_gauge *U)
{
    QKX_CHECK_ARG(n && r && e && U, "qkx_net_accumulate_grad: null argument");
    int rc;
    if ((rc = check_net_gauge(U)) || (rc = check_spinor(r, U)) || (rc = check_spinor(e, U))) return rc;
    if ((rc = net_buffers(n, U->L))) return rc;
    const int V = U->V;
    if ((rc = to_feat(r, n->fin))) return rc;
    if ((rc = net_forward(n, U))) return rc;
    // target into gh's first 4 channels (free until the adjoint below)
    if ((rc = to_feat(e, n->gh))) return rc;
    const int ne = 12 * V, nb = std::min(nblk(ne), 1024);
    hipLaunchKernelGGL(loss_kernel, dim3(nb), dim3(kT), 0, qkx::g_stream, n->fout, n->gh, n->gout, ne, n->part);
    hipLaunchKernelGGL(loss_finish_kernel, dim3(1), dim3(64), 0, qkx::g_stream, n->part, nb, n->loss);
    QKX_HIP(hipGetLastError());
    const bool bf = n->bf16 != 0;
    const void *w2 = bf ? (const void *)(n->wb + n->n1) : (const void *)(n->w + n->n1);
    double2 *g1 = n->grad, *g2 = n->grad + n->n1;
    if (n->ch == 16) {
        if ((rc = launch_wgrad<16, 4>(n->gout, n->fh, U, n, g2))) return rc;
        if ((rc = launch_ptc<4, 16, true>(n->gout, n->gh, U, w2, bf))) return rc;
        return launch_wgrad<4, 16>(n->gh, n->fin, U, n, g1);
    }
    if (n->ch == 4) {
        if ((rc = launch_wgrad<4, 4>(n->gout, n->fh, U, n, g2))) return rc;
        if ((rc = launch_ptc<4, 4, true>(n->gout, n->gh, U, w2, bf))) return rc;
        return launch_wgrad<4, 4>(n->gh, n->fin, U, n, g1);
    }
    if ((rc = launch_wgrad<8, 4>(n->gout, n->fh, U, n, g2))) return rc;
    if ((rc = launch_ptc<4, 8, true>(n->gout, n->gh, U, w2, bf))) return rc;
    return launch_wgrad<4, 8>(n->gh, n->fin, U, n, g1);
}

int qkx_net_read_grad(qkx_net *n, float *g, double *loss)
{
    QKX_CHECK_ARG(n, "qkx_net_read_grad: null argument");
    const int nt = n->n1 + n->n2;
    std::vector<double2> h(nt);
    QKX_HIP(hipMemcpyAsync(h.data(), n->grad, sizeof(double2) * nt, hipMemcpyDeviceToHost, qkx::g_stream));
    if (loss) QKX_HIP(hipMemcpyAsync(loss, n->loss, sizeof(double), hipMemcpyDeviceToHost, qkx::g_stream));
    QKX_HIP(hipStreamSynchronize(qkx::g_stream));
    if (g)
        for (int i = 0; i < nt; i++) {
            g[2 * i] = (float)h[i].x;
            g[2 * i + 1] = (float)h[i].y;
        }
    return QKX_SUCCESS;
}

int qkx_net_zero_grad(qkx_net *n)
{
    QKX_CHECK_ARG(n, "qkx_net_zero_grad: null argument");
    QKX_HIP(hipMemsetAsync(n->grad, 0, sizeof(double2) * (n->n1 + n->n2), qkx::g_stream));
    QKX_HIP(hipMemsetAsync(n->loss, 0, sizeof(double), qkx::g_stream));
    return QKX_SUCCESS;
}

int qkx_net_adam(qkx_net *n, double lr, double beta1, double beta2, double eps, int count)
{
    QKX_CHECK_ARG(n && count > 0 && lr > 0, "qkx_net_adam: bad argument");
    n->step++;
    const float c1 = (float)(1.0 - std::pow(beta1, n->step)), c2 = (float)(1.0 - std::pow(beta2, n->step));
    const int nt = n->n1 + n->n2;
    hipLaunchKernelGGL(adam_kernel, dim3((nt + 255) / 256), dim3(256), 0, qkx::g_stream, n->w, n->wb, n->grad, n->m,
                       n->v, nt, (float)lr, (float)beta1, (float)beta2, (float)eps, c1, c2, 1.0f / (float)count);
    QKX_HIP(hipGetLastError());
    return QKX_SUCCESS;
}

} // extern "C"
