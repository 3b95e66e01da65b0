// qkx_internal.h -- device data layout, storage accessors and library state
// shared by the HIP sources of libqkx_hip.so (gfx950 only).
//
// Device layout ("parity-blocked struct of arrays"):
//   site index  s = parity * Vh + cb,  cb = lexicographic index / 2
//               (lexicographic order of the reference, include/qkxTM.h:130;
//                Lx must be even so that cb = x/2 + Lx/2*(y + Ly*(z + Lz*t)))
//   spinor      component-major: component k of site s at  base[k*stride + s]
//               DOUBLE: 12 x double2 (one complex each)
//               SINGLE:  6 x float4  (two complex each, k = spin*3+colour order)
//               HALF  :  6 x 4 fp16  (two complex each) + one float scale per site
//               BF16  :  6 x 4 bf16  (two complex each)
//   gauge       direction-major, link (mu, s) element k at base[(mu*NC + k)*V + s]
//               DOUBLE/18: 9 double2      DOUBLE/12: 6 double2
//               SINGLE/18: 9 float2       SINGLE/12: 3 float4
//               HALF/12  : 3 x 4 fp16     BF16/12  : 3 x 4 bf16
// Consecutive threads take consecutive s, so every load of a wavefront is one
// contiguous 512 B or 1 KiB run.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>
#include <cstdint>
#include <cstddef>
#include <string>
#include "qkx_hip.h"

// ---------------------------------------------------------------------------
// complex arithmetic
template <class T> struct Cx { T re, im; };

template <class T> __host__ __device__ __forceinline__ Cx<T> cx(T r, T i) { return Cx<T>{r, i}; }
template <class T> __device__ __forceinline__ Cx<T> operator+(Cx<T> a, Cx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <class T> __device__ __forceinline__ Cx<T> operator-(Cx<T> a, Cx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <class T> __device__ __forceinline__ Cx<T> operator*(T a, Cx<T> b) { return {a * b.re, a * b.im}; }
template <class T> __device__ __forceinline__ Cx<T> cmul(Cx<T> a, Cx<T> b)
{
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
// conj(a) * b
template <class T> __device__ __forceinline__ Cx<T> cjmul(Cx<T> a, Cx<T> b)
{
    return {a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re};
}
// i * a
template <class T> __device__ __forceinline__ Cx<T> timesi(Cx<T> a) { return {-a.im, a.re}; }

// ---------------------------------------------------------------------------
// geometry passed to kernels
struct Geom {
    int L[4];
    int Lh0;     // L[0]/2
    int V, Vh;
    int bc;      // sign applied to t-links of the last time slice (+1/-1)
};

// ---------------------------------------------------------------------------
// fp16 / bf16 pack helpers (4 values in a uint2)
__device__ __forceinline__ void unpack_h4(uint2 w, float f[4])
{
    __half2 a = *reinterpret_cast<__half2 *>(&w.x);
    __half2 b = *reinterpret_cast<__half2 *>(&w.y);
    float2 fa = __half22float2(a), fb = __half22float2(b);
    f[0] = fa.x; f[1] = fa.y; f[2] = fb.x; f[3] = fb.y;
}
__device__ __forceinline__ uint2 pack_h4(float a, float b, float c, float d)
{
    __half2 x = __floats2half2_rn(a, b), y = __floats2half2_rn(c, d);
    uint2 w;
    w.x = *reinterpret_cast<uint32_t *>(&x);
    w.y = *reinterpret_cast<uint32_t *>(&y);
    return w;
}
__device__ __forceinline__ float bf16_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf16_hi(uint32_t w) { return __uint_as_float(w & 0xffff0000u); }
__device__ __forceinline__ uint32_t f2bf16_bits(float f)
{
    // round to nearest even; values are finite here
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
}
__device__ __forceinline__ void unpack_b4(uint2 w, float f[4])
{
    f[0] = bf16_lo(w.x); f[1] = bf16_hi(w.x); f[2] = bf16_lo(w.y); f[3] = bf16_hi(w.y);
}
__device__ __forceinline__ uint2 pack_b4(float a, float b, float c, float d)
{
    uint2 w;
    w.x = f2bf16_bits(a) | (f2bf16_bits(b) << 16);
    w.y = f2bf16_bits(c) | (f2bf16_bits(d) << 16);
    return w;
}

// ---------------------------------------------------------------------------
// spinor storage accessors.  load_half<H>: H=0 -> spins 0,1 (components 0..5),
// H=1 -> spins 2,3 (components 6..11).
struct SpinorD {
    using Real = double;
    double2 *p;
    int stride;
    template <int H> __device__ __forceinline__ void load_half(int s, Cx<double> v[12]) const
    {
#pragma unroll
        for (int k = 0; k < 6; k++) {
            double2 d = p[(size_t)(6 * H + k) * stride + s];
            v[6 * H + k] = {d.x, d.y};
        }
    }
    __device__ __forceinline__ void load(int s, Cx<double> v[12]) const
    {
        load_half<0>(s, v);
        load_half<1>(s, v);
    }
    __device__ __forceinline__ void store(int s, const Cx<double> v[12]) const
    {
#pragma unroll
        for (int k = 0; k < 12; k++) p[(size_t)k * stride + s] = make_double2(v[k].re, v[k].im);
    }
};

struct SpinorF {
    using Real = float;
    float4 *p;
    int stride;
    template <int H> __device__ __forceinline__ void load_half(int s, Cx<float> v[12]) const
    {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            float4 d = p[(size_t)(3 * H + k) * stride + s];
            v[6 * H + 2 * k] = {d.x, d.y};
            v[6 * H + 2 * k + 1] = {d.z, d.w};
        }
    }
    __device__ __forceinline__ void load(int s, Cx<float> v[12]) const
    {
        load_half<0>(s, v);
        load_half<1>(s, v);
    }
    __device__ __forceinline__ void store(int s, const Cx<float> v[12]) const
    {
#pragma unroll
        for (int k = 0; k < 6; k++)
            p[(size_t)k * stride + s] = make_float4(v[2 * k].re, v[2 * k].im, v[2 * k + 1].re, v[2 * k + 1].im);
    }
};

// FP16 storage with one FP32 scale per site (max |component|).
struct SpinorH {
    using Real = float;
    uint2 *p;
    float *nrm;
    int stride;
    template <int H> __device__ __forceinline__ void load_half(int s, Cx<float> v[12]) const
    {
        float n = nrm[s];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            float f[4];
            unpack_h4(p[(size_t)(3 * H + k) * stride + s], f);
            v[6 * H + 2 * k] = {f[0] * n, f[1] * n};
            v[6 * H + 2 * k + 1] = {f[2] * n, f[3] * n};
        }
    }
    __device__ __forceinline__ void load(int s, Cx<float> v[12]) const
    {
        load_half<0>(s, v);
        load_half<1>(s, v);
    }
    __device__ __forceinline__ void store(int s, const Cx<float> v[12]) const
    {
        float m = 0.f;
#pragma unroll
        for (int k = 0; k < 12; k++) m = fmaxf(m, fmaxf(fabsf(v[k].re), fabsf(v[k].im)));
        float inv = m > 0.f ? 1.0f / m : 0.f;
        nrm[s] = m;
#pragma unroll
        for (int k = 0; k < 6; k++)
            p[(size_t)k * stride + s] = pack_h4(v[2 * k].re * inv, v[2 * k].im * inv,
                                                v[2 * k + 1].re * inv, v[2 * k + 1].im * inv);
    }
};

struct SpinorB {
    using Real = float;
    uint2 *p;
    int stride;
    template <int H> __device__ __forceinline__ void load_half(int s, Cx<float> v[12]) const
    {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            float f[4];
            unpack_b4(p[(size_t)(3 * H + k) * stride + s], f);
            v[6 * H + 2 * k] = {f[0], f[1]};
            v[6 * H + 2 * k + 1] = {f[2], f[3]};
        }
    }
    __device__ __forceinline__ void load(int s, Cx<float> v[12]) const
    {
        load_half<0>(s, v);
        load_half<1>(s, v);
    }
    __device__ __forceinline__ void store(int s, const Cx<float> v[12]) const
    {
#pragma unroll
        for (int k = 0; k < 6; k++)
            p[(size_t)k * stride + s] = pack_b4(v[2 * k].re, v[2 * k].im, v[2 * k + 1].re, v[2 * k + 1].im);
    }
};

// ---------------------------------------------------------------------------
// gauge storage accessors: load(mu, s, U) fills the 3x3 link U[row*3+col].
template <class T> __device__ __forceinline__ void recon_row2(Cx<T> U[9])
{
    // row2 = conj(row0 x row1)  (SU(3))
    Cx<T> a = cmul(U[1], U[5]) - cmul(U[2], U[4]);
    Cx<T> b = cmul(U[2], U[3]) - cmul(U[0], U[5]);
    Cx<T> c = cmul(U[0], U[4]) - cmul(U[1], U[3]);
    U[6] = {a.re, -a.im};
    U[7] = {b.re, -b.im};
    U[8] = {c.re, -c.im};
}

struct GaugeD18 {
    using Real = double;
    const double2 *p;
    int V;
    __device__ __forceinline__ void load(int mu, int s, Cx<double> U[9]) const
    {
#pragma unroll
        for (int k = 0; k < 9; k++) {
            double2 d = p[(size_t)(mu * 9 + k) * V + s];
            U[k] = {d.x, d.y};
        }
    }
};
struct GaugeD12 {
    using Real = double;
    const double2 *p;
    int V;
    __device__ __forceinline__ void load(int mu, int s, Cx<double> U[9]) const
    {
#pragma unroll
        for (int k = 0; k < 6; k++) {
            double2 d = p[(size_t)(mu * 6 + k) * V + s];
            U[k] = {d.x, d.y};
        }
        recon_row2(U);
    }
};
struct GaugeF18 {
    using Real = float;
    const float2 *p;
    int V;
    __device__ __forceinline__ void load(int mu, int s, Cx<float> U[9]) const
    {
#pragma unroll
        for (int k = 0; k < 9; k++) {
            float2 d = p[(size_t)(mu * 9 + k) * V + s];
            U[k] = {d.x, d.y};
        }
    }
};
struct GaugeF12 {
    using Real = float;
    const float4 *p;
    int V;
    __device__ __forceinline__ void load(int mu, int s, Cx<float> U[9]) const
    {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            float4 d = p[(size_t)(mu * 3 + k) * V + s];
            U[2 * k] = {d.x, d.y};
            U[2 * k + 1] = {d.z, d.w};
        }
        recon_row2(U);
    }
};
struct GaugeH12 {
    using Real = float;
    const uint2 *p;
    int V;
    __device__ __forceinline__ void load(int mu, int s, Cx<float> U[9]) const
    {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            float f[4];
            unpack_h4(p[(size_t)(mu * 3 + k) * V + s], f);
            U[2 * k] = {f[0], f[1]};
            U[2 * k + 1] = {f[2], f[3]};
        }
        recon_row2(U);
    }
};
struct GaugeB12 {
    using Real = float;
    const uint2 *p;
    int V;
    __device__ __forceinline__ void load(int mu, int s, Cx<float> U[9]) const
    {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            float f[4];
            unpack_b4(p[(size_t)(mu * 3 + k) * V + s], f);
            U[2 * k] = {f[0], f[1]};
            U[2 * k + 1] = {f[2], f[3]};
        }
        recon_row2(U);
    }
};

// ---------------------------------------------------------------------------
// host-side handle types (declared before the accessor factories below)
struct qkx_spinor_s {
    int L[4];
    int V, Vh;
    int prec;     // QKX_PREC_*
    int subset;   // QKX_SUBSET_*
    int nsites;   // V or Vh
    void *data;   // component-major storage
    float *nrm;   // HALF only
    size_t bytes;
};

struct qkx_gauge_s {
    int L[4];
    int V, Vh;
    int prec;
    int recon;
    int bc;
    void *data;
    size_t bytes;
    // cached low-precision copy for the mixed solvers (owned)
    qkx_gauge_s *inner;
};

// library state
namespace qkx {
extern hipStream_t g_stream;
extern int g_device;
extern int g_cu_count;
void set_error(const std::string &msg);
bool check_hip(hipError_t e, const char *what);
int ensure_init();
Geom make_geom(const int L[4], int bc);
// offsets of the two parities inside a field's storage (-1: parity absent)
void parity_offsets(const qkx_spinor_s *f, int off[2]);
size_t spinor_elem_bytes(int prec);   // bytes per component-chunk element
int spinor_nchunks(int prec);         // chunks per site
// device scratch for reductions: partial sums + counter + result
double *reduce_scratch();
unsigned int *reduce_counter();
double *reduce_result();
// profiling of hopping kernels
void prof_begin();
void prof_end(double bytes);
} // namespace qkx

#define QKX_CHECK_ARG(cond, msg)           \
    do {                                   \
        if (!(cond)) {                     \
            qkx::set_error(msg);           \
            return QKX_ERR_ARG;            \
        }                                  \
    } while (0)

#define QKX_HIP(call)                                  \
    do {                                               \
        hipError_t e__ = (call);                       \
        if (!qkx::check_hip(e__, #call)) return QKX_ERR_HIP; \
    } while (0)

// ---------------------------------------------------------------------------
// accessor factories and precision dispatch
template <class S> inline S make_spinor(const qkx_spinor_s *f);
template <> inline SpinorD make_spinor<SpinorD>(const qkx_spinor_s *f) { return SpinorD{(double2 *)f->data, f->nsites}; }
template <> inline SpinorF make_spinor<SpinorF>(const qkx_spinor_s *f) { return SpinorF{(float4 *)f->data, f->nsites}; }
template <> inline SpinorH make_spinor<SpinorH>(const qkx_spinor_s *f) { return SpinorH{(uint2 *)f->data, f->nrm, f->nsites}; }
template <> inline SpinorB make_spinor<SpinorB>(const qkx_spinor_s *f) { return SpinorB{(uint2 *)f->data, f->nsites}; }

template <class G> inline G make_gauge(const qkx_gauge_s *g);
template <> inline GaugeD18 make_gauge<GaugeD18>(const qkx_gauge_s *g) { return GaugeD18{(const double2 *)g->data, g->V}; }
template <> inline GaugeD12 make_gauge<GaugeD12>(const qkx_gauge_s *g) { return GaugeD12{(const double2 *)g->data, g->V}; }
template <> inline GaugeF18 make_gauge<GaugeF18>(const qkx_gauge_s *g) { return GaugeF18{(const float2 *)g->data, g->V}; }
template <> inline GaugeF12 make_gauge<GaugeF12>(const qkx_gauge_s *g) { return GaugeF12{(const float4 *)g->data, g->V}; }
template <> inline GaugeH12 make_gauge<GaugeH12>(const qkx_gauge_s *g) { return GaugeH12{(const uint2 *)g->data, g->V}; }
template <> inline GaugeB12 make_gauge<GaugeB12>(const qkx_gauge_s *g) { return GaugeB12{(const uint2 *)g->data, g->V}; }

// fn(tag) is called with a default-constructed accessor of the field's type.
template <class F> inline int dispatch_prec(int prec, F fn)
{
    switch (prec) {
    case QKX_PREC_DOUBLE: return fn(SpinorD{});
    case QKX_PREC_SINGLE: return fn(SpinorF{});
    case QKX_PREC_HALF: return fn(SpinorH{});
    case QKX_PREC_BF16: return fn(SpinorB{});
    }
    qkx::set_error("bad spinor precision");
    return QKX_ERR_ARG;
}
