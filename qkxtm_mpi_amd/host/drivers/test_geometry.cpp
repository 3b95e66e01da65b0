// test_geometry.cpp -- counterpart of the reference's bin/test_geometry.cpp:
// build the lattice geometry and check its neighbour tables.  Every site's
// pointPlus / pointMinus entry in each direction must be the site one step
// away (coordinates compared, periodic wrap), the two tables must invert
// each other, and fields of one lattice size must share one set of tables.
// Prints "geometry ok V=<sites>" and the five gamma matrices of the basis
// ("gamma g a b re im" lines).  Makes no GPU call.
//   usage: test_geometry Lx Ly Lz Lt
#include <cstdio>
#include <cstdlib>

#include "qkxtm/qkxTM.h"
#include "qkxtm/lattice_util.h"

using namespace qkxTM;

int main(int argc, char *argv[])
{
    LatticeInfo latInfo;
    for (int m = 0; m < 4; m++) latInfo.L[m] = argc > 4 ? atoi(argv[1 + m]) : 4 + 2 * m;
    Vector *a = new Vector(&latInfo);
    Vector *b = new Vector(&latInfo);
    const int *L = a->L;
    long bad = 0;
    for (long iv = 0; iv < a->V4d; iv++) {
        int c[4];
        long r = iv;
        for (int m = 0; m < 4; m++) {
            c[m] = (int)(r % L[m]);
            r /= L[m];
        }
        for (int mu = 0; mu < 4; mu++) {
            for (int sgn = -1; sgn <= 1; sgn += 2) {
                const long nb = sgn > 0 ? a->pointPlus[mu][iv] : a->pointMinus[mu][iv];
                long s = nb;
                for (int m = 0; m < 4; m++) {
                    const int cm = (int)(s % L[m]);
                    s /= L[m];
                    const int want = m == mu ? (c[m] + sgn + L[m]) % L[m] : c[m];
                    if (cm != want) bad++;
                }
            }
            if (a->pointMinus[mu][a->pointPlus[mu][iv]] != iv) bad++;
        }
    }
    for (int mu = 0; mu < 4; mu++)
        if (a->pointPlus[mu] != b->pointPlus[mu] || a->pointMinus[mu] != b->pointMinus[mu]) bad++;
    if (bad) {
        printf("geometry FAILED: %ld wrong entries\n", bad);
        return 1;
    }
    printf("geometry ok V=%ld\n", a->V4d);
    for (int g = 0; g < 5; g++)
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++)
                printf("gamma %d %d %d %+.1f %+.1f\n", g, i, j, a->gammaM[g][i][j].real(), a->gammaM[g][i][j].imag());
    delete b;
    delete a;
    return 0;
}
