/*
 * lattice_util.h -- drop-in counterpart of the reference's
 * include/lattice_util.h (classes Vector and Gauge, hot-path methods only),
 * backed by libqkx_hip.so.
 *
 * The host arrays M keep the reference layout (Vector::M Complex[V][4][3],
 * Gauge::M Complex[V][4][3][3]) and stay authoritative at the API boundary:
 * every operator uploads its inputs, runs on the GPU and downloads its
 * output into this->M.  Drivers may therefore read and write M directly, as
 * the reference drivers do (bin/test_invert.cpp:89).  The gauge is uploaded
 * on every operator call unless Gauge::constGauge is true and it was uploaded
 * before (the reference sets constGauge once the links are final,
 * bin/test_invert.cpp:110).  For the fast path keep fields on the device with
 * the C API in qkx_hip.h.
 */
#ifndef QKXTM_DROPIN_LATTICE_UTIL_H
#define QKXTM_DROPIN_LATTICE_UTIL_H

#include "qkxTM.h"

struct qkx_spinor_s;
struct qkx_gauge_s;

namespace qkxTM {

class Gauge;
class Vector;

enum PARITY { EO, OE, OO, EE };   /* reference include/lattice_util.h:23 */

/* LatticeGeometry, reference include/lattice_util.h:25-76, lib/lattice_util.cpp:44-378,
 * for one process holding the whole lattice.  pointPlus/pointMinus are the
 * neighbour tables of lib/lattice_util.cpp:86-99 (periodic; on one process the
 * ghost-zone redirection of :111-202 never applies).  Fields of one lattice size
 * share one set of tables, built by the first and freed with the last. */
class LatticeGeometry {
  public:
    long int V4d;
    long int V3d;
    long int lV4d;
    long int lV3d;
    int L[NDIM];
    int P[NDIM];
    int lL[NDIM];
    int procPos[NDIM];
    int rank;
    int Nprocs;

    int procPlus[NDIM];      /* next / previous process in each direction: 0 */
    int procMinus[NDIM];
    int *pointPlus[NDIM];    /* lexicographic index of x + mu, x - mu (periodic) */
    int *pointMinus[NDIM];

    /* ghost-zone bookkeeping, lib/lattice_util.cpp:102-207, 392-397, at its
     * one-process values: no ghost zone (start 0, NtotalBoundary 0, pointers
     * null; the reference's Vector, Gauge and Propagator each declare the two
     * pointer arrays, here they sit in the common base) */
    int placeBoundaryStartPlus[NDIM];
    int placeBoundaryStartMinus[NDIM];
    int lSurfaceBoundary[NDIM];
    int NtotalBoundary;
    void *ptr_boundaryPlusStart[NDIM];
    void *ptr_boundaryMinusStart[NDIM];

    bool initialized;

    LatticeGeometry(LatticeInfo *latInfo);
    virtual ~LatticeGeometry();

    /* gamma_1 .. gamma_4, gamma_5 in the UKQCD basis, lib/lattice_util.cpp:273-282 */
    Complex gammaM[5][4][4];
};

class Vector : public LatticeGeometry {
  public:
    long int size;
    long int size_total;
    Complex (*M)[NSPINS][NCOLORS];

    Vector(LatticeInfo *latInfo);
    virtual ~Vector();

    void zero();
    void copy(Vector &src);
    double norm2();
    void normalize();
    void communicateToPlus();
    void communicateToMinus();

    void applyDslash(Vector &psi, Gauge &u);
    void applyDslash_chiral(Vector &psi, Gauge &u);
    void applyDslash_xx(Vector &psi, Gauge &u, PARITY parity);
    void applyDslash_xx_chiral(Vector &psi, Gauge &u, PARITY parity);
    void applyDeltaAddDslash(Vector &dslash, Vector &psi, LatticeInfo *latInfo);
    void applyTwistAddDslash(Vector &dslash, Vector &psi, LatticeInfo *latInfo);
    void applyTwistAddDslash_chiral(Vector &dslash, Vector &psi, LatticeInfo *latInfo);
    void applyTwistAddDslash_PC(Vector &dslash, Vector &psi, LatticeInfo *latInfo, PARITY parity);
    void applyTwistAddDslash_PC_chiral(Vector &dslash, Vector &psi, LatticeInfo *latInfo, PARITY parity);

    void applyDslashDag(Vector &psi, Gauge &u);
    void applyTwistDagAddDslashDag(Vector &dslashDag, Vector &psi, LatticeInfo *latInfo);
    void applyTwist(Vector &psi, LatticeInfo *latInfo);
    void applyTwistDag(Vector &psi, LatticeInfo *latInfo);
    void applyTwistInv(Vector &psi, LatticeInfo *latInfo);
    void applyTwistInv_chiral(Vector &psi, LatticeInfo *latInfo);
    void applyTwistDagInv(Vector &psi, LatticeInfo *latInfo);
    void applyGamma5(Vector &psi, LatticeInfo *latInfo);
    void applyGamma5_chiral(Vector &psi, LatticeInfo *latInfo);

    void tm(Vector &psi, Gauge &u, LatticeInfo *latInfo);
    void tm_chiral(Vector &psi, Gauge &u, LatticeInfo *latInfo);
    void tmDag(Vector &psi, Gauge &u, LatticeInfo *latInfo);
    void tmDagtm(Vector &psi, Vector &vec_tmp, Gauge &u, LatticeInfo *latInfo);
    void Wilson(Vector &psi, Gauge &u, LatticeInfo *latInfo);

    void tm_xx(Vector &psi, Gauge &u, LatticeInfo *latInfo, PARITY parity);
    void tm_xxDagTm_xx(Vector &psi, Gauge &u, LatticeInfo *latInfo, PARITY parity);
    void tm_xx_chiral(Vector &psi, Gauge &u, LatticeInfo *latInfo, PARITY parity);
    void tm_xxDagTm_xx_chiral(Vector &psi, Gauge &u, LatticeInfo *latInfo, PARITY parity);

    /* source preparation: include/lattice_util.h:101 (NsmearGauss, alphaGauss
     * from latInfo; this != &vec_tmp) */
    void Gaussian_smearing(Vector &vec_tmp, Gauge &gauge, LatticeInfo *latInfo);
    /* include/lattice_util.h:138 */
    void rescale(double a);

    /* device mirror (library internal) */
    qkx_spinor_s *dev();
    void upload();
    void download();

  private:
    qkx_spinor_s *dev_;
    Vector(const Vector &);
    Vector &operator=(const Vector &);
};

class Gauge : public LatticeGeometry {
  public:
    long int size;
    long int size_total;
    bool constGauge;
    Complex (*M)[NDIM][NCOLORS][NCOLORS];

    Gauge(LatticeInfo *latInfo);
    virtual ~Gauge();

    void zero();
    void copy(Gauge &src);
    double norm2();
    void communicateToPlus();
    void communicateToMinus();
    void applyBoundaryCondition(LatticeInfo *latInfo);
    /* include/lattice_util.h:209-210 (NsmearAPE, alphaAPE from latInfo; tmp is
     * not modified) */
    double calculatePlaquette();
    void APE_smearing(Gauge &tmp, LatticeInfo *latInfo);

    /* device mirror, uploaded as needed (see constGauge above); hostModified()
     * marks it stale after M was rewritten outside the class (library internal) */
    qkx_gauge_s *dev();
    void hostModified();

  private:
    qkx_gauge_s *dev_;
    bool uploaded_;
    Gauge(const Gauge &);
    Gauge &operator=(const Gauge &);
};

/* include/lattice_util.h:158-185.  M (host, reference layout) is
 * authoritative: zero, copy, copyVector and absorbVector are host copies like
 * the reference's; norm2 and the rotations run on the GPU. */
class Propagator : public LatticeGeometry {
  public:
    long int size;
    long int size_total;
    Complex (*M)[NSPINS][NSPINS][NCOLORS][NCOLORS];

    Propagator(LatticeInfo *latInfo);
    virtual ~Propagator();

    void zero();
    void copy(Propagator &src);
    void communicateToPlus();
    void communicateToMinus();
    void copyVector(Vector &v, int nu, int c2);
    void absorbVector(Vector &phi, int nu, int c2);
    double norm2();
    void rotateToPhysicalBasePlus();
    void rotateToPhysicalBaseMinus();

  private:
    qkx_spinor_s *dev_;
    void rotate(int sign);
    Propagator(const Propagator &);
    Propagator &operator=(const Propagator &);
};

} // namespace qkxTM

/* include/lattice_util.h:325: SU(3) projection of the spatial links, in place */
void qcd_projectSU33d(qkxTM::Gauge &gf);

/* include/lattice_util.h:319-320: ILDG/LIME gauge files (returns 0; errors
 * abort like errorQKXTM) */
int qkxTM_MPI_getGaugeLime(char *fname, qkxTM::Gauge &u);
int qkxTM_writeGaugeLime(char *fname, qkxTM::Gauge &u, char *message);

#endif
