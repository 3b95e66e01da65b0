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
//               HALF  :  3 x 8 fp16  (four complex each) + one float scale per site
//               BF16  :  3 x 8 bf16  (four complex each)
//   gauge       direction-major, link (mu, s) element k at base[(mu*NC + k)*V + s]
//               DOUBLE/18: 9 double2      DOUBLE/12: 6 double2
//               SINGLE/18: 9 float2       SINGLE/12: 3 float4
//               HALF/12, BF16/12: 8 x 16-bit ([4][V] uint4) then 4 x 16-bit ([4][V] uint2)
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
// Complex products written as explicit FMA chains (1 mul + 1 fma for the
// first product of a sum, 2 fma for each further one per real component):
// a*b, acc + a*b, conj(a)*b, acc + conj(a)*b.
template <class T> __device__ __forceinline__ Cx<T> cmul_f(Cx<T> a, Cx<T> b)
{
    return {fma(a.re, b.re, -a.im * b.im), fma(a.re, b.im, a.im * b.re)};
}
template <class T> __device__ __forceinline__ Cx<T> cmac(Cx<T> acc, Cx<T> a, Cx<T> b)
{
    return {fma(-a.im, b.im, fma(a.re, b.re, acc.re)), fma(a.im, b.re, fma(a.re, b.im, acc.im))};
}
template <class T> __device__ __forceinline__ Cx<T> cjmul_f(Cx<T> a, Cx<T> b)
{
    return {fma(a.re, b.re, a.im * b.im), fma(a.re, b.im, -a.im * b.re)};
}
template <class T> __device__ __forceinline__ Cx<T> cjmac(Cx<T> acc, Cx<T> a, Cx<T> b)
{
    return {fma(a.im, b.im, fma(a.re, b.re, acc.re)), fma(-a.im, b.re, fma(a.re, b.im, acc.im))};
}

// ---------------------------------------------------------------------------
// Compensated (double-double) sums for the global reductions of the solvers
// and BLAS: a sum is carried as hi + lo, and every addition keeps its exact
// rounding error (Knuth's TwoSum), so adding n per-workgroup partials keeps
// about twice FP64's precision instead of drifting by up to n ulps.  dd_add
// is symmetric in its operands (the two low parts are added together before
// they meet the error term), so an xor butterfly leaves the same bits in
// every lane and every workgroup that sums the same partials in the same
// order takes the same decisions.
struct DD {
    double hi, lo;
};
__device__ __forceinline__ DD two_sum(double a, double b)
{
    const double s = a + b;
    const double bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}
__device__ __forceinline__ DD dd_add(DD a, DD b)
{
    const DD s = two_sum(a.hi, b.hi);
    return two_sum(s.hi, s.lo + (a.lo + b.lo));
}
__device__ __forceinline__ DD dd_add(DD a, double b)
{
    const DD s = two_sum(a.hi, b);
    return two_sum(s.hi, s.lo + a.lo);
}
// xor-butterfly sum over the 64 lanes of a wavefront (same bits in every lane)
__device__ __forceinline__ DD dd_wave_sum(DD v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = dd_add(v, DD{__shfl_xor(v.hi, off, 64), __shfl_xor(v.lo, off, 64)});
    return v;
}
__device__ __forceinline__ double dd_value(DD v) { return v.hi + v.lo; }

// ---------------------------------------------------------------------------
// geometry passed to kernels
// Division of an index n < 2^30 by a fixed d as a multiply and a shift:
// n / d = (n * m) >> s with s = 30 + ceil(log2 d), m = floor(2^s / d) + 1
// (Granlund and Montgomery, PLDI 1994: exact because m d - 2^s <= 2^(s-30)).
struct FastDiv {
    uint32_t m;
    uint32_t s;
    int d;
};
__host__ __device__ __forceinline__ int fdiv(int n, const FastDiv &f)
{
    return (int)(((uint64_t)(uint32_t)n * f.m) >> f.s);
}
inline FastDiv make_fastdiv(int d)
{
    int l = 0;
    while ((1 << l) < d) l++;
    FastDiv f;
    f.s = 30u + (uint32_t)l;
    f.m = (uint32_t)((1ull << f.s) / (uint64_t)d + 1ull);
    f.d = d;
    return f;
}

struct Geom {
    int L[4];
    int Lh0;     // L[0]/2
    int V, Vh;
    int bc;      // sign applied to t-links of the last time slice (+1/-1)
    FastDiv dLh0, dL1, dL2;   // divisions by Lh0, L[1], L[2]
};

// full-lattice parity-blocked index s -> lexicographic index
__device__ __forceinline__ int blocked_to_lex(int s, const Geom &g)
{
    int p = (s >= g.Vh) ? 1 : 0;
    int cb = s - p * g.Vh;
    int xh = cb % g.Lh0;
    int rest = cb / g.Lh0;
    int y = rest % g.L[1];
    int zt = rest / g.L[1];
    int z = zt % g.L[2];
    int t = zt / g.L[2];
    int x = 2 * xh + ((p + y + z + t) & 1);
    return x + g.L[0] * rest;
}

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
// Loads/stores as (uniform base pointer, 32-bit byte offset): the base of each
// component lives in scalar registers and every site costs one 32-bit vector
// offset, which maps onto the saddr form of global_load/global_store.
template <class E> __device__ __forceinline__ E ld(const E *base, uint32_t idx)
{
    return *reinterpret_cast<const E *>(reinterpret_cast<const char *>(base) + idx * (uint32_t)sizeof(E));
}
// The same load marked non-temporal (streamed past the caches): for operands
// read exactly once per launch, such as the links of a half-lattice hop, so
// that they do not evict the neighbour spinors that are read up to 8 times.
#ifndef QKX_NT_LINKS
#define QKX_NT_LINKS 0
#endif
template <int N> struct NtVec;
template <> struct NtVec<8> { typedef __attribute__((ext_vector_type(2))) unsigned int type; };
template <> struct NtVec<16> { typedef __attribute__((ext_vector_type(4))) unsigned int type; };
template <class E> __device__ __forceinline__ E ld_link(const E *base, uint32_t idx)
{
    if constexpr (QKX_NT_LINKS && (sizeof(E) == 8 || sizeof(E) == 16)) {
        using V = typename NtVec<sizeof(E)>::type;
        const V v = __builtin_nontemporal_load(
            reinterpret_cast<const V *>(reinterpret_cast<const char *>(base) + idx * (uint32_t)sizeof(E)));
        E e;
        __builtin_memcpy(&e, &v, sizeof(E));
        return e;
    } else {
        return ld(base, idx);
    }
}
template <class E> __device__ __forceinline__ void st(E *base, uint32_t idx, E v)
{
    *reinterpret_cast<E *>(reinterpret_cast<char *>(base) + idx * (uint32_t)sizeof(E)) = v;
}

// 8 x 16-bit in a uint4
__device__ __forceinline__ void unpack_h8(uint4 w, float f[8])
{
    unpack_h4(make_uint2(w.x, w.y), f);
    unpack_h4(make_uint2(w.z, w.w), f + 4);
}
__device__ __forceinline__ uint4 pack_h8(const float f[8])
{
    uint2 a = pack_h4(f[0], f[1], f[2], f[3]), b = pack_h4(f[4], f[5], f[6], f[7]);
    return make_uint4(a.x, a.y, b.x, b.y);
}
__device__ __forceinline__ void unpack_b8(uint4 w, float f[8])
{
    unpack_b4(make_uint2(w.x, w.y), f);
    unpack_b4(make_uint2(w.z, w.w), f + 4);
}
__device__ __forceinline__ uint4 pack_b8(const float f[8])
{
    uint2 a = pack_b4(f[0], f[1], f[2], f[3]), b = pack_b4(f[4], f[5], f[6], f[7]);
    return make_uint4(a.x, a.y, b.x, b.y);
}

// ---------------------------------------------------------------------------
// spinor storage accessors.  Components are k = spin*3 + colour.
//   fetch<MASK>(s) returns the raw (unconverted) words, unpack<MASK> converts;
//   MASK bit 0 = spins 0,1 (components 0..5), bit 1 = spins 2,3 (6..11).
//   Splitting fetch from unpack lets the stencil keep one hop's loads in
//   flight while it computes the previous hop.
struct SpinorD {
    using Real = double;
    double2 *p;
    int stride;
    struct Raw {
        double2 c[12];
    };
    template <int MASK> __device__ __forceinline__ Raw fetch(int s) const
    {
        Raw r;
#pragma unroll
        for (int k = 0; k < 12; k++)
            if ((MASK >> (k / 6)) & 1) r.c[k] = ld(p + (size_t)k * stride, (uint32_t)s);
        return r;
    }
    template <int MASK> __device__ __forceinline__ void unpack(const Raw &r, Cx<double> v[12]) const
    {
#pragma unroll
        for (int k = 0; k < 12; k++)
            if ((MASK >> (k / 6)) & 1) v[k] = {r.c[k].x, r.c[k].y};
    }
    __device__ __forceinline__ void load(int s, Cx<double> v[12]) const { unpack<3>(fetch<3>(s), v); }
    __device__ __forceinline__ void store(int s, const Cx<double> v[12]) const
    {
#pragma unroll
        for (int k = 0; k < 12; k++) st(p + (size_t)k * stride, (uint32_t)s, make_double2(v[k].re, v[k].im));
    }
    // the same, marked non-temporal: the 192 B per site of a hop's output
    // stream to memory instead of displacing the neighbour spinors and links
    // that the launch is still re-reading from L2
    __device__ __forceinline__ void store_nt(int s, const Cx<double> v[12]) const
    {
        typedef __attribute__((ext_vector_type(2))) double d2v;
#pragma unroll
        for (int k = 0; k < 12; k++)
            __builtin_nontemporal_store(d2v{v[k].re, v[k].im}, reinterpret_cast<d2v *>(p + (size_t)k * stride + s));
    }
};

// FP32: 6 x float4, chunk j = components 2j, 2j+1
struct SpinorF {
    using Real = float;
    float4 *p;
    int stride;
    struct Raw {
        float4 c[6];
    };
    template <int MASK> __device__ __forceinline__ Raw fetch(int s) const
    {
        Raw r;
#pragma unroll
        for (int j = 0; j < 6; j++)
            if ((MASK >> (j / 3)) & 1) r.c[j] = ld(p + (size_t)j * stride, (uint32_t)s);
        return r;
    }
    template <int MASK> __device__ __forceinline__ void unpack(const Raw &r, Cx<float> v[12]) const
    {
#pragma unroll
        for (int j = 0; j < 6; j++)
            if ((MASK >> (j / 3)) & 1) {
                v[2 * j] = {r.c[j].x, r.c[j].y};
                v[2 * j + 1] = {r.c[j].z, r.c[j].w};
            }
    }
    __device__ __forceinline__ void load(int s, Cx<float> v[12]) const { unpack<3>(fetch<3>(s), v); }
    __device__ __forceinline__ void store(int s, const Cx<float> v[12]) const
    {
#pragma unroll
        for (int j = 0; j < 6; j++)
            st(p + (size_t)j * stride, (uint32_t)s,
               make_float4(v[2 * j].re, v[2 * j].im, v[2 * j + 1].re, v[2 * j + 1].im));
    }
};

// 16-bit storage: 3 x uint4, chunk j = components 4j..4j+3.  Spins 0,1 need
// chunks 0,1; spins 2,3 need chunks 1,2.
template <int MASK> __device__ __forceinline__ constexpr bool need_chunk16(int j)
{
    return (j == 0 && (MASK & 1)) || (j == 1) || (j == 2 && (MASK & 2));
}

// FP16 with one FP32 scale per site (max |component|)
struct SpinorH {
    using Real = float;
    uint4 *p;
    float *nrm;
    int stride;
    struct Raw {
        uint4 c[3];
        float n;
    };
    template <int MASK> __device__ __forceinline__ Raw fetch(int s) const
    {
        Raw r;
#pragma unroll
        for (int j = 0; j < 3; j++)
            if (need_chunk16<MASK>(j)) r.c[j] = ld(p + (size_t)j * stride, (uint32_t)s);
        r.n = ld(nrm, (uint32_t)s);
        return r;
    }
    template <int MASK> __device__ __forceinline__ void unpack(const Raw &r, Cx<float> v[12]) const
    {
#pragma unroll
        for (int j = 0; j < 3; j++)
            if (need_chunk16<MASK>(j)) {
                float f[8];
                unpack_h8(r.c[j], f);
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int k = 4 * j + i;
                    if ((MASK >> (k / 6)) & 1) v[k] = {f[2 * i] * r.n, f[2 * i + 1] * r.n};
                }
            }
    }
    // the stored values without the site scale, and the scale
    template <int MASK> __device__ __forceinline__ void unpack_raw(const Raw &r, Cx<float> v[12]) const
    {
#pragma unroll
        for (int j = 0; j < 3; j++)
            if (need_chunk16<MASK>(j)) {
                float f[8];
                unpack_h8(r.c[j], f);
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int k = 4 * j + i;
                    if ((MASK >> (k / 6)) & 1) v[k] = {f[2 * i], f[2 * i + 1]};
                }
            }
    }
    __device__ __forceinline__ float scale(const Raw &r) const { return r.n; }
    __device__ __forceinline__ void load(int s, Cx<float> v[12]) const { unpack<3>(fetch<3>(s), v); }
    __device__ __forceinline__ void store(int s, const Cx<float> v[12]) const
    {
        float m = 0.f;
#pragma unroll
        for (int k = 0; k < 12; k++) m = fmaxf(m, fmaxf(fabsf(v[k].re), fabsf(v[k].im)));
        // a site whose largest component is below the smallest normal float
        // (far from a point source) is stored as zero: 1/m would overflow
        const bool keep = m >= 1.17549435e-38f;
        if (!keep) m = 0.f;
        const float inv = keep ? 1.0f / m : 0.f;
        st(nrm, (uint32_t)s, m);
#pragma unroll
        for (int j = 0; j < 3; j++) {
            float f[8];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                f[2 * i] = v[4 * j + i].re * inv;
                f[2 * i + 1] = v[4 * j + i].im * inv;
            }
            st(p + (size_t)j * stride, (uint32_t)s, pack_h8(f));
        }
    }
};

// BF16 (no scale: bf16 has the FP32 exponent range)
struct SpinorB {
    using Real = float;
    uint4 *p;
    int stride;
    struct Raw {
        uint4 c[3];
    };
    template <int MASK> __device__ __forceinline__ Raw fetch(int s) const
    {
        Raw r;
#pragma unroll
        for (int j = 0; j < 3; j++)
            if (need_chunk16<MASK>(j)) r.c[j] = ld(p + (size_t)j * stride, (uint32_t)s);
        return r;
    }
    template <int MASK> __device__ __forceinline__ void unpack(const Raw &r, Cx<float> v[12]) const
    {
#pragma unroll
        for (int j = 0; j < 3; j++)
            if (need_chunk16<MASK>(j)) {
                float f[8];
                unpack_b8(r.c[j], f);
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int k = 4 * j + i;
                    if ((MASK >> (k / 6)) & 1) v[k] = {f[2 * i], f[2 * i + 1]};
                }
            }
    }
    __device__ __forceinline__ void load(int s, Cx<float> v[12]) const { unpack<3>(fetch<3>(s), v); }
    __device__ __forceinline__ void store(int s, const Cx<float> v[12]) const
    {
#pragma unroll
        for (int j = 0; j < 3; j++) {
            float f[8];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                f[2 * i] = v[4 * j + i].re;
                f[2 * i + 1] = v[4 * j + i].im;
            }
            st(p + (size_t)j * stride, (uint32_t)s, pack_b8(f));
        }
    }
};

// ---------------------------------------------------------------------------
// gauge storage accessors: fetch(mu, s) raw words, unpack -> U[row*3+col].
// row2 = conj(row0 x row1)  (SU(3)).  FMA: the products as explicit FMA
// chains (fewer instructions; used where the ALU, not memory, bounds the hop).
template <bool FMA = false, class T> __device__ __forceinline__ void recon_row2(Cx<T> U[9])
{
    Cx<T> a, b, c;
    if constexpr (FMA) {
        const Cx<T> n2 = {-U[2].re, -U[2].im}, n0 = {-U[0].re, -U[0].im}, n1 = {-U[1].re, -U[1].im};
        a = cmac(cmul_f(U[1], U[5]), n2, U[4]);
        b = cmac(cmul_f(U[2], U[3]), n0, U[5]);
        c = cmac(cmul_f(U[0], U[4]), n1, U[3]);
    } else {
        a = cmul(U[1], U[5]) - cmul(U[2], U[4]);
        b = cmul(U[2], U[3]) - cmul(U[0], U[5]);
        c = cmul(U[0], U[4]) - cmul(U[1], U[3]);
    }
    U[6] = {a.re, -a.im};
    U[7] = {b.re, -b.im};
    U[8] = {c.re, -c.im};
}

template <class E, int NC, class R> struct GaugeVec {
    using Real = R;
    const E *p;
    int V;
    struct Raw {
        E c[NC];
    };
    __device__ __forceinline__ Raw fetch(int mu, int s) const
    {
        Raw r;
#pragma unroll
        for (int k = 0; k < NC; k++) r.c[k] = ld_link(p + (size_t)(mu * NC + k) * V, (uint32_t)s);
        return r;
    }
};

// FP64, 18 reals: 9 double2
struct GaugeD18 : GaugeVec<double2, 9, double> {
    __device__ __forceinline__ void unpack(const Raw &r, Cx<double> U[9]) const
    {
#pragma unroll
        for (int k = 0; k < 9; k++) U[k] = {r.c[k].x, r.c[k].y};
    }
    __device__ __forceinline__ void load(int mu, int s, Cx<double> U[9]) const { unpack(fetch(mu, s), U); }
};
// FP64, 12 reals: 6 double2, row 2 reconstructed
struct GaugeD12 : GaugeVec<double2, 6, double> {
    __device__ __forceinline__ void unpack(const Raw &r, Cx<double> U[9]) const
    {
#pragma unroll
        for (int k = 0; k < 6; k++) U[k] = {r.c[k].x, r.c[k].y};
        recon_row2(U);
    }
    __device__ __forceinline__ void load(int mu, int s, Cx<double> U[9]) const { unpack(fetch(mu, s), U); }
};
// FP32, 18 reals: 9 float2
struct GaugeF18 : GaugeVec<float2, 9, float> {
    __device__ __forceinline__ void unpack(const Raw &r, Cx<float> U[9]) const
    {
#pragma unroll
        for (int k = 0; k < 9; k++) U[k] = {r.c[k].x, r.c[k].y};
    }
    __device__ __forceinline__ void load(int mu, int s, Cx<float> U[9]) const { unpack(fetch(mu, s), U); }
};
// FP32, 12 reals: 3 float4
struct GaugeF12 : GaugeVec<float4, 3, float> {
    __device__ __forceinline__ void unpack(const Raw &r, Cx<float> U[9]) const
    {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            U[2 * k] = {r.c[k].x, r.c[k].y};
            U[2 * k + 1] = {r.c[k].z, r.c[k].w};
        }
        recon_row2(U);
    }
    __device__ __forceinline__ void load(int mu, int s, Cx<float> U[9]) const { unpack(fetch(mu, s), U); }
};
// 16-bit, 12 reals: a uint4 (U0..U3) array [4][V] followed by a uint2 (U4, U5)
// array [4][V]; links are unitary so no scale is needed
template <bool BF> struct Gauge16 {
    using Real = float;
    const uint4 *p4;
    const uint2 *p2;
    int V;
    struct Raw {
        uint4 a;
        uint2 b;
    };
    __device__ __forceinline__ Raw fetch(int mu, int s) const
    {
        Raw r;
        r.a = ld_link(p4 + (size_t)mu * V, (uint32_t)s);
        r.b = ld_link(p2 + (size_t)mu * V, (uint32_t)s);
        return r;
    }
    __device__ __forceinline__ void unpack(const Raw &r, Cx<float> U[9]) const
    {
        float f[8], g[4];
        if constexpr (BF) {
            unpack_b8(r.a, f);
            unpack_b4(r.b, g);
        } else {
            unpack_h8(r.a, f);
            unpack_h4(r.b, g);
        }
#pragma unroll
        for (int k = 0; k < 4; k++) U[k] = {f[2 * k], f[2 * k + 1]};
        U[4] = {g[0], g[1]};
        U[5] = {g[2], g[3]};
        recon_row2<true>(U);
    }
    __device__ __forceinline__ void load(int mu, int s, Cx<float> U[9]) const { unpack(fetch(mu, s), U); }
};
using GaugeH12 = Gauge16<false>;
using GaugeB12 = Gauge16<true>;

// ---------------------------------------------------------------------------
// host-side handle types (declared before the accessor factories below)
struct qkx_spinor_s {
    int L[4];
    int V, Vh;
    int prec;     // QKX_PREC_*
    int subset;   // QKX_SUBSET_*
    int nsites;   // V or Vh
    int nrhs;     // right-hand sides stored innermost: entry (site, rhs) = site*nrhs + rhs
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
    // cached FP32 12-real copy for the reliable updates of the mixed solvers (owned)
    qkx_gauge_s *mid;
    // cached FP64 12-real copy for the solvers' internal FP64 operator, used
    // only while the links are SU(3) to 1e-14 (owned)
    qkx_gauge_s *outer;
    // max over links of |row2 - conj(row0 x row1)| (what a 12-real copy would
    // change), and the generation it was measured at (~0u: never)
    double recon_dev;
    double recon_dev_flip;   // the same with the last time slice's t-links negated
    unsigned recon_gen;
    // generation of the links: bumped by every write (upload, conversion into
    // this field, smearing, projection); a cached copy records the generation
    // of the field it was converted from and is refreshed when they differ
    unsigned gen;
    unsigned src_gen;
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
void prof_end(double bytes, int prec);
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
// storage entries of a field: sites x right-hand sides
inline int entries(const qkx_spinor_s *f) { return f->nsites * f->nrhs; }
template <> inline SpinorD make_spinor<SpinorD>(const qkx_spinor_s *f) { return SpinorD{(double2 *)f->data, entries(f)}; }
template <> inline SpinorF make_spinor<SpinorF>(const qkx_spinor_s *f) { return SpinorF{(float4 *)f->data, entries(f)}; }
template <> inline SpinorH make_spinor<SpinorH>(const qkx_spinor_s *f) { return SpinorH{(uint4 *)f->data, f->nrm, entries(f)}; }
template <> inline SpinorB make_spinor<SpinorB>(const qkx_spinor_s *f) { return SpinorB{(uint4 *)f->data, entries(f)}; }

template <class G> inline G make_gauge(const qkx_gauge_s *g);
template <> inline GaugeD18 make_gauge<GaugeD18>(const qkx_gauge_s *g) { GaugeD18 r; r.p = (const double2 *)g->data; r.V = g->V; return r; }
template <> inline GaugeD12 make_gauge<GaugeD12>(const qkx_gauge_s *g) { GaugeD12 r; r.p = (const double2 *)g->data; r.V = g->V; return r; }
template <> inline GaugeF18 make_gauge<GaugeF18>(const qkx_gauge_s *g) { GaugeF18 r; r.p = (const float2 *)g->data; r.V = g->V; return r; }
template <> inline GaugeF12 make_gauge<GaugeF12>(const qkx_gauge_s *g) { GaugeF12 r; r.p = (const float4 *)g->data; r.V = g->V; return r; }
// 16-bit links: uint4 part [4][V], then uint2 part [4][V]
template <> inline GaugeH12 make_gauge<GaugeH12>(const qkx_gauge_s *g)
{
    return GaugeH12{(const uint4 *)g->data, (const uint2 *)((const char *)g->data + (size_t)16 * 4 * g->V), g->V};
}
template <> inline GaugeB12 make_gauge<GaugeB12>(const qkx_gauge_s *g)
{
    return GaugeB12{(const uint4 *)g->data, (const uint2 *)((const char *)g->data + (size_t)16 * 4 * g->V), g->V};
}

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
