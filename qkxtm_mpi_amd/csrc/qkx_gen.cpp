// qkx_gen.cpp -- synthetic gauge fields and sources in the reference host
// layout (Complex U[V][4][3][3], Complex psi[V][4][3]).
//
// The reference reads its gauge fields from ILDG/LIME files
// (lib/read_conf_lime.cpp:111-329); this generator stands in for that input.
// Every random number is a pure function of (seed, global site, direction,
// draw index) through a splitmix64 counter hash, so a field does not depend on
// the number of threads or on the order sites are visited.
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>
#include <algorithm>

#include "qkx_hip.h"

namespace {

using cd = std::complex<double>;

inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Stream {
    uint64_t base;
    uint64_t k = 0;
    Stream(uint64_t seed, uint64_t site, uint64_t sub)
        : base(mix64(seed ^ mix64(site * 64 + sub + 1))) {}
    uint64_t next() { return mix64(base + 0xD1B54A32D192ED03ull * (++k)); }
    // uniform in (0,1): 53 random bits, offset by half an ulp
    double uniform() { return ((next() >> 11) + 0.5) * 0x1.0p-53; }
    // one standard normal (Box-Muller, cosine branch)
    double normal()
    {
        double u1 = uniform(), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

template <class F>
void parallel_sites(long V, F f)
{
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (V < 4096) nt = 1;
    std::vector<std::thread> th;
    long chunk = (V + nt - 1) / nt;
    for (unsigned t = 0; t < nt; t++) {
        long lo = t * chunk, hi = std::min(V, lo + chunk);
        if (lo >= hi) break;
        th.emplace_back([=] { for (long iv = lo; iv < hi; iv++) f(iv); });
    }
    for (auto &t : th) t.join();
}

typedef cd M3[3][3];

void mat_mul(const M3 a, const M3 b, M3 c)
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            cd s = 0;
            for (int k = 0; k < 3; k++) s += a[i][k] * b[k][j];
            c[i][j] = s;
        }
}

// exp(X) by scaling and squaring with a Taylor series (X anti-Hermitian,
// ||X|| = O(1)): exp(X) = exp(X/2^8)^(2^8), series to 1e-17 accuracy.
void mat_exp(const M3 X, M3 E)
{
    const int sq = 8;
    M3 Y, term, tmp;
    double s = std::ldexp(1.0, -sq);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            Y[i][j] = X[i][j] * s;
            E[i][j] = (i == j) ? 1.0 : 0.0;
            term[i][j] = (i == j) ? 1.0 : 0.0;
        }
    for (int n = 1; n <= 12; n++) {
        mat_mul(term, Y, tmp);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                term[i][j] = tmp[i][j] / double(n);
                E[i][j] += term[i][j];
            }
    }
    for (int k = 0; k < sq; k++) {
        mat_mul(E, E, tmp);
        std::memcpy(E, tmp, sizeof(M3));
    }
}

// Reunitarise to SU(3): normalise row 0, Gram-Schmidt row 1, row 2 =
// conj(row0 x row1).  Row 2 fixed this way gives det = 1.
void reunitarize(M3 U)
{
    double n0 = 0;
    for (int j = 0; j < 3; j++) n0 += std::norm(U[0][j]);
    n0 = 1.0 / std::sqrt(n0);
    for (int j = 0; j < 3; j++) U[0][j] *= n0;
    cd dot = 0;
    for (int j = 0; j < 3; j++) dot += std::conj(U[0][j]) * U[1][j];
    for (int j = 0; j < 3; j++) U[1][j] -= dot * U[0][j];
    double n1 = 0;
    for (int j = 0; j < 3; j++) n1 += std::norm(U[1][j]);
    n1 = 1.0 / std::sqrt(n1);
    for (int j = 0; j < 3; j++) U[1][j] *= n1;
    U[2][0] = std::conj(U[0][1] * U[1][2] - U[0][2] * U[1][1]);
    U[2][1] = std::conj(U[0][2] * U[1][0] - U[0][0] * U[1][2]);
    U[2][2] = std::conj(U[0][0] * U[1][1] - U[0][1] * U[1][0]);
}

int check_dims(const int L[4])
{
    for (int m = 0; m < 4; m++)
        if (L[m] <= 0) return -1;
    return 0;
}

} // namespace

extern "C" {

int qkx_gen_gauge(double *U_aos, const int L[4], double eps, unsigned long long seed)
{
    if (!U_aos || check_dims(L)) return QKX_ERR_ARG;
    long V = (long)L[0] * L[1] * L[2] * L[3];
    cd *U = reinterpret_cast<cd *>(U_aos);
    parallel_sites(V, [&](long iv) {
        for (int mu = 0; mu < 4; mu++) {
            Stream rng(seed, (uint64_t)iv, (uint64_t)mu);
            M3 A, X, E;
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    double re = rng.normal();
                    double im = rng.normal();
                    A[i][j] = cd(re, im);
                }
            // H = (A + A^dag)/2 - tr/3 ; X = i eps H
            cd H[3][3];
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) H[i][j] = 0.5 * (A[i][j] + std::conj(A[j][i]));
            cd tr = (H[0][0] + H[1][1] + H[2][2]) / 3.0;
            for (int i = 0; i < 3; i++) H[i][i] -= tr;
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) X[i][j] = cd(0.0, eps) * H[i][j];
            mat_exp(X, E);
            reunitarize(E);
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) U[((iv * 4 + mu) * 3 + i) * 3 + j] = E[i][j];
        }
    });
    return QKX_SUCCESS;
}

int qkx_gen_z2(double *psi_aos, const int L[4], unsigned long long seed)
{
    if (!psi_aos || check_dims(L)) return QKX_ERR_ARG;
    long V = (long)L[0] * L[1] * L[2] * L[3];
    const double r = 1.0 / std::sqrt(2.0);
    parallel_sites(V, [&](long iv) {
        Stream rng(seed, (uint64_t)iv, 7);
        uint64_t bits = rng.next();
        for (int k = 0; k < 12; k++) {
            psi_aos[iv * 24 + 2 * k + 0] = ((bits >> (2 * k)) & 1) ? -r : r;
            psi_aos[iv * 24 + 2 * k + 1] = ((bits >> (2 * k + 1)) & 1) ? -r : r;
        }
    });
    return QKX_SUCCESS;
}

int qkx_gen_gaussian(double *psi_aos, const int L[4], unsigned long long seed)
{
    if (!psi_aos || check_dims(L)) return QKX_ERR_ARG;
    long V = (long)L[0] * L[1] * L[2] * L[3];
    parallel_sites(V, [&](long iv) {
        Stream rng(seed, (uint64_t)iv, 11);
        for (int k = 0; k < 24; k++) psi_aos[iv * 24 + k] = rng.normal();
    });
    return QKX_SUCCESS;
}

int qkx_gen_point(double *psi_aos, const int L[4], const int x0[4], int spin, int colour)
{
    if (!psi_aos || check_dims(L) || spin < 0 || spin > 3 || colour < 0 || colour > 2)
        return QKX_ERR_ARG;
    for (int m = 0; m < 4; m++)
        if (x0[m] < 0 || x0[m] >= L[m]) return QKX_ERR_ARG;
    long V = (long)L[0] * L[1] * L[2] * L[3];
    std::memset(psi_aos, 0, sizeof(double) * 24 * V);
    long iv = (((long)x0[3] * L[2] + x0[2]) * L[1] + x0[1]) * L[0] + x0[0];
    psi_aos[iv * 24 + (spin * 3 + colour) * 2] = 1.0;
    return QKX_SUCCESS;
}

int qkx_gen_site(int x0[4], const int L[4], unsigned long long seed)
{
    if (check_dims(L)) return QKX_ERR_ARG;
    Stream rng(seed, 0, 13);
    for (int m = 0; m < 4; m++) x0[m] = (int)(rng.next() % (uint64_t)L[m]);
    return QKX_SUCCESS;
}

} // extern "C"
