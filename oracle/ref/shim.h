/* Force-included when compiling the reference sources.  ISO C++ has no
 * operator*(int, std::complex<double>); three lines of the reference that are
 * not on the operator/CG path (lib/lattice_util.cpp:2173,2288,3477) use it and
 * otherwise stop the file from compiling.  Nothing on the hot path calls it. */
#include <complex>
inline std::complex<double> operator*(int a, const std::complex<double> &b)
{
    return std::complex<double>((double)a * b.real(), (double)a * b.imag());
}
