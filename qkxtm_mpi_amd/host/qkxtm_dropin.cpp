// qkxtm_dropin.cpp -- the reference's Vector / Gauge / CG / BLAS interface
// (include/lattice_util.h, include/solvers.h, include/blas_qkxTM.h) over the
// C ABI of libqkx_hip.so.  C++98 (gnu++98), like the reference.
//
// Each operator uploads its inputs from the host arrays M (reference layout),
// runs on the GPU and downloads the result into this->M, so drivers that read
// or write M directly keep working.  Errors that the reference turns into
// errorQKXTM (bad parity, bad process grid) abort the same way.
#include "qkxtm/qkxTM.h"
#include "qkxtm/lattice_util.h"
#include "qkxtm/solvers.h"
#include "qkxtm/blas_qkxTM.h"
#include "qkx_hip.h"

#include <cstdlib>
#include <vector>

using namespace qkxTM;

namespace {

void check(int rc, const char *what)
{
    if (rc != QKX_SUCCESS) errorQKXTM("%s: %s\n", what, qkx_last_error());
}

qkx_params params_of(const LatticeInfo *li)
{
    qkx_params p;
    for (int m = 0; m < 4; m++) p.L[m] = li->L[m];
    p.kappa = li->kappa;
    p.mu = li->mu;
    p.twistSign = li->twistSign;
    p.boundaryT = li->boundaryT;
    p.tol = li->tol;
    p.maxIter = (int)li->maxIter;
    p.reliableDelta = li->reliableDelta;
    return p;
}

int parity_code(PARITY p)
{
    switch (p) {
    case EO: return QKX_PARITY_EO;
    case OE: return QKX_PARITY_OE;
    case OO: return QKX_PARITY_OO;
    case EE: return QKX_PARITY_EE;
    }
    return -99;
}

// Neighbour tables shared by the fields of one lattice size (the reference
// builds 8 V ints for every field: 21 MB per Vector at 24^3x48).
struct NbTables {
    int L[NDIM];
    int *plus[NDIM], *minus[NDIM];
    long refs;
};
std::vector<NbTables> g_tables;

bool same_dims(const int a[NDIM], const int b[NDIM])
{
    for (int i = 0; i < NDIM; i++)
        if (a[i] != b[i]) return false;
    return true;
}

void tables_acquire(const int L[NDIM], int *plus[NDIM], int *minus[NDIM])
{
    for (size_t k = 0; k < g_tables.size(); k++)
        if (same_dims(g_tables[k].L, L)) {
            g_tables[k].refs++;
            for (int m = 0; m < NDIM; m++) {
                plus[m] = g_tables[k].plus[m];
                minus[m] = g_tables[k].minus[m];
            }
            return;
        }
    NbTables t;
    const long V = (long)L[0] * L[1] * L[2] * L[3];
    for (int m = 0; m < NDIM; m++) {
        t.L[m] = L[m];
        t.plus[m] = (int *)malloc(V * sizeof(int));
        t.minus[m] = (int *)malloc(V * sizeof(int));
        if (!t.plus[m] || !t.minus[m]) errorQKXTM("Error out of memory");
    }
    t.refs = 1;
    // stride of direction m in the lexicographic index x + L0 (y + L1 (z + L2 t))
    long stride[NDIM];
    stride[0] = 1;
    for (int m = 1; m < NDIM; m++) stride[m] = stride[m - 1] * L[m - 1];
    for (long iv = 0; iv < V; iv++)
        for (int m = 0; m < NDIM; m++) {
            const int c = (int)((iv / stride[m]) % L[m]);
            const long up = (c + 1 == L[m]) ? -(long)(L[m] - 1) * stride[m] : stride[m];
            const long dn = (c == 0) ? (long)(L[m] - 1) * stride[m] : -stride[m];
            t.plus[m][iv] = (int)(iv + up);
            t.minus[m][iv] = (int)(iv + dn);
        }
    g_tables.push_back(t);
    for (int m = 0; m < NDIM; m++) {
        plus[m] = t.plus[m];
        minus[m] = t.minus[m];
    }
}

void tables_release(const int L[NDIM])
{
    for (size_t k = 0; k < g_tables.size(); k++)
        if (same_dims(g_tables[k].L, L)) {
            if (--g_tables[k].refs == 0) {
                for (int m = 0; m < NDIM; m++) {
                    free(g_tables[k].plus[m]);
                    free(g_tables[k].minus[m]);
                }
                g_tables.erase(g_tables.begin() + k);
            }
            return;
        }
}

} // namespace

// ---- LatticeGeometry: lib/lattice_util.cpp:44-99, 273-282 (one process) ------
LatticeGeometry::LatticeGeometry(LatticeInfo *latInfo)
{
    rank = latInfo->rank < 0 ? 0 : latInfo->rank;
    for (int i = 0; i < NDIM; i++) {
        L[i] = latInfo->L[i];
        P[i] = latInfo->P[i];
    }
    if (P[0] * P[1] * P[2] * P[3] != 1 || latInfo->Nprocs > 1)
        errorQKXTM("Error one GPU holds the whole lattice: the process grid must be 1,1,1,1\n");
    for (int i = 0; i < NDIM; i++) {
        if (L[i] <= 0 || (L[i] & 1)) errorQKXTM("Error every lattice extent must be even and positive\n");
        lL[i] = L[i];
        procPos[i] = 0;
    }
    V4d = (long)L[0] * L[1] * L[2] * L[3];
    V3d = (long)L[0] * L[1] * L[2];
    lV4d = V4d;
    lV3d = V3d;
    Nprocs = 1;
    for (int i = 0; i < NDIM; i++) {
        procPlus[i] = procMinus[i] = 0;
        placeBoundaryStartPlus[i] = placeBoundaryStartMinus[i] = 0;
        lSurfaceBoundary[i] = (int)(V4d / L[i]);
        ptr_boundaryPlusStart[i] = ptr_boundaryMinusStart[i] = 0;
    }
    NtotalBoundary = 0;
    tables_acquire(L, pointPlus, pointMinus);
    // UKQCD basis: gammaM[0..3] = gamma_1..gamma_4, gammaM[4] = gamma_5
    // = gamma_1 gamma_2 gamma_3 gamma_4 (the basis the hopping kernels use)
    for (int g = 0; g < 5; g++)
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) gammaM[g][a][b] = Complex(0, 0);
    const Complex I(0, 1), one(1, 0);
    gammaM[0][0][3] = I;
    gammaM[0][1][2] = I;
    gammaM[0][2][1] = -I;
    gammaM[0][3][0] = -I;
    gammaM[1][0][3] = one;
    gammaM[1][1][2] = -one;
    gammaM[1][2][1] = -one;
    gammaM[1][3][0] = one;
    gammaM[2][0][2] = I;
    gammaM[2][1][3] = -I;
    gammaM[2][2][0] = -I;
    gammaM[2][3][1] = I;
    gammaM[3][0][0] = one;
    gammaM[3][1][1] = one;
    gammaM[3][2][2] = -one;
    gammaM[3][3][3] = -one;
    gammaM[4][0][2] = one;
    gammaM[4][1][3] = one;
    gammaM[4][2][0] = one;
    gammaM[4][3][1] = one;
    initialized = true;
}

LatticeGeometry::~LatticeGeometry() { tables_release(L); }

// ---- Vector: lib/lattice_util.cpp:382-477 ----------------------------------
Vector::Vector(LatticeInfo *latInfo) : LatticeGeometry(latInfo), dev_(0)
{
    size = lV4d;
    size_total = lV4d;
    M = (Complex(*)[NSPINS][NCOLORS])malloc(NSPINOR * size_total * sizeof(double));
    if (!M) errorQKXTM("Error out of memory");
}

Vector::~Vector()
{
    if (dev_) qkx_spinor_destroy(dev_);
    free(M);
}

qkx_spinor_s *Vector::dev()
{
    if (!dev_) check(qkx_spinor_create(&dev_, L, QKX_PREC_DOUBLE, QKX_SUBSET_FULL), "Vector");
    return dev_;
}
void Vector::upload() { check(qkx_spinor_upload(dev(), &(M[0][0][0].real())), "Vector upload"); }
void Vector::download() { check(qkx_spinor_download(dev(), &(M[0][0][0].real())), "Vector download"); }

void Vector::zero() { memset(&(M[0][0][0].real()), 0, NSPINOR * size_total * sizeof(double)); }
void Vector::copy(Vector &src) { memcpy(&(M[0][0][0].real()), &(src.M[0][0][0].real()), lV4d * NSPINOR * sizeof(double)); }

double Vector::norm2()
{
    upload();
    double r = 0;
    check(qkx_norm2(&r, dev()), "norm2");
    return r;
}

// Vector::normalize :1915-1923: divide by the square root of norm2()
void Vector::normalize()
{
    const double n2 = norm2();   // uploads
    check(qkx_spinor_rescale(dev(), 1.0 / sqrt(n2)), "normalize");
    download();
}

// Ghost-zone exchange, Vector :479-526, Propagator :602-650, Gauge :726-778.
// The reference exchanges a face only in a direction whose local extent is
// below the global one; one device holds the whole lattice and its kernels
// wrap periodically, so there is never a face to exchange.
void Vector::communicateToPlus() {}
void Vector::communicateToMinus() {}
void Propagator::communicateToPlus() {}
void Propagator::communicateToMinus() {}
void Gauge::communicateToPlus() {}
void Gauge::communicateToMinus() {}

// hopping terms: applyDslash :1158, _chiral :1215, _xx :1272, _xx_chiral :1346, Dag :1643
void Vector::applyDslash(Vector &psi, Gauge &u)
{
    psi.upload();
    check(qkx_dslash(dev(), psi.dev(), u.dev(), QKX_BASIS_UKQCD, 0, QKX_PARITY_ALL), "applyDslash");
    download();
}
void Vector::applyDslash_chiral(Vector &psi, Gauge &u)
{
    psi.upload();
    check(qkx_dslash(dev(), psi.dev(), u.dev(), QKX_BASIS_CHIRAL, 0, QKX_PARITY_ALL), "applyDslash_chiral");
    download();
}
void Vector::applyDslash_xx(Vector &psi, Gauge &u, PARITY parity)
{
    psi.upload();
    check(qkx_dslash(dev(), psi.dev(), u.dev(), QKX_BASIS_UKQCD, 0, parity_code(parity)), "applyDslash_xx");
    download();
}
void Vector::applyDslash_xx_chiral(Vector &psi, Gauge &u, PARITY parity)
{
    psi.upload();
    check(qkx_dslash(dev(), psi.dev(), u.dev(), QKX_BASIS_CHIRAL, 0, parity_code(parity)), "applyDslash_xx_chiral");
    download();
}
void Vector::applyDslashDag(Vector &psi, Gauge &u)
{
    psi.upload();
    check(qkx_dslash(dev(), psi.dev(), u.dev(), QKX_BASIS_UKQCD, 1, QKX_PARITY_ALL), "applyDslashDag");
    download();
}

// site terms added to a hop: :1420, :1444, :1468, :1510, :1552, :1701
static void xpay(Vector &out, Vector &d, Vector &psi, LatticeInfo *li, int basis, int kind, int parity,
                 const char *what)
{
    d.upload();
    if (&psi != &d) psi.upload();
    qkx_params p = params_of(li);
    check(qkx_twist_add_dslash(out.dev(), d.dev(), psi.dev(), &p, basis, kind, parity), what);
    out.download();
}
void Vector::applyDeltaAddDslash(Vector &d, Vector &psi, LatticeInfo *li)
{
    xpay(*this, d, psi, li, QKX_BASIS_UKQCD, QKX_XPAY_WILSON, QKX_PARITY_ALL, "applyDeltaAddDslash");
}
void Vector::applyTwistAddDslash(Vector &d, Vector &psi, LatticeInfo *li)
{
    xpay(*this, d, psi, li, QKX_BASIS_UKQCD, QKX_XPAY_TM, QKX_PARITY_ALL, "applyTwistAddDslash");
}
void Vector::applyTwistAddDslash_chiral(Vector &d, Vector &psi, LatticeInfo *li)
{
    xpay(*this, d, psi, li, QKX_BASIS_CHIRAL, QKX_XPAY_TM, QKX_PARITY_ALL, "applyTwistAddDslash_chiral");
}
void Vector::applyTwistAddDslash_PC(Vector &d, Vector &psi, LatticeInfo *li, PARITY parity)
{
    xpay(*this, d, psi, li, QKX_BASIS_UKQCD, QKX_XPAY_PC, parity_code(parity), "applyTwistAddDslash_PC");
}
void Vector::applyTwistAddDslash_PC_chiral(Vector &d, Vector &psi, LatticeInfo *li, PARITY parity)
{
    xpay(*this, d, psi, li, QKX_BASIS_CHIRAL, QKX_XPAY_PC, parity_code(parity), "applyTwistAddDslash_PC_chiral");
}
void Vector::applyTwistDagAddDslashDag(Vector &d, Vector &psi, LatticeInfo *li)
{
    xpay(*this, d, psi, li, QKX_BASIS_UKQCD, QKX_XPAY_TM_DAG, QKX_PARITY_ALL, "applyTwistDagAddDslashDag");
}

// site-local twist family: :1563, :1726, :1588, :1615, :1750, :1778, :1800
static void twist(Vector &out, Vector &psi, LatticeInfo *li, int basis, int kind, const char *what)
{
    psi.upload();
    qkx_params p = params_of(li);
    check(qkx_twist(out.dev(), psi.dev(), &p, basis, kind), what);
    out.download();
}
void Vector::applyTwist(Vector &psi, LatticeInfo *li) { twist(*this, psi, li, QKX_BASIS_UKQCD, QKX_TWIST, "applyTwist"); }
void Vector::applyTwistDag(Vector &psi, LatticeInfo *li) { twist(*this, psi, li, QKX_BASIS_UKQCD, QKX_TWIST_DAG, "applyTwistDag"); }
void Vector::applyTwistInv(Vector &psi, LatticeInfo *li) { twist(*this, psi, li, QKX_BASIS_UKQCD, QKX_TWIST_INV, "applyTwistInv"); }
void Vector::applyTwistInv_chiral(Vector &psi, LatticeInfo *li) { twist(*this, psi, li, QKX_BASIS_CHIRAL, QKX_TWIST_INV, "applyTwistInv_chiral"); }
void Vector::applyTwistDagInv(Vector &psi, LatticeInfo *li) { twist(*this, psi, li, QKX_BASIS_UKQCD, QKX_TWIST_DAG_INV, "applyTwistDagInv"); }
void Vector::applyGamma5(Vector &psi, LatticeInfo *li) { twist(*this, psi, li, QKX_BASIS_UKQCD, QKX_GAMMA5, "applyGamma5"); }
void Vector::applyGamma5_chiral(Vector &psi, LatticeInfo *li) { twist(*this, psi, li, QKX_BASIS_CHIRAL, QKX_GAMMA5, "applyGamma5_chiral"); }

// composites: tm :1925, tm_chiral :1930, tmDag :1935, tmDagtm :1940, Wilson :1949
static void tm_op(Vector &out, Vector &psi, Gauge &u, LatticeInfo *li, int basis, int op, const char *what)
{
    psi.upload();
    qkx_params p = params_of(li);
    check(qkx_tm(out.dev(), psi.dev(), u.dev(), &p, basis, op), what);
    out.download();
}
void Vector::tm(Vector &psi, Gauge &u, LatticeInfo *li) { tm_op(*this, psi, u, li, QKX_BASIS_UKQCD, QKX_OP_TM, "tm"); }
void Vector::tm_chiral(Vector &psi, Gauge &u, LatticeInfo *li) { tm_op(*this, psi, u, li, QKX_BASIS_CHIRAL, QKX_OP_TM, "tm_chiral"); }
void Vector::tmDag(Vector &psi, Gauge &u, LatticeInfo *li) { tm_op(*this, psi, u, li, QKX_BASIS_UKQCD, QKX_OP_TM_DAG, "tmDag"); }
void Vector::Wilson(Vector &psi, Gauge &u, LatticeInfo *li) { tm_op(*this, psi, u, li, QKX_BASIS_UKQCD, QKX_OP_WILSON, "Wilson"); }
void Vector::tmDagtm(Vector &psi, Vector &vec_tmp, Gauge &u, LatticeInfo *li)
{
    // as in the reference, vec_tmp ends up holding D psi
    psi.upload();
    qkx_params p = params_of(li);
    check(qkx_tm(vec_tmp.dev(), psi.dev(), u.dev(), &p, QKX_BASIS_UKQCD, QKX_OP_TM), "tmDagtm");
    check(qkx_tm(dev(), vec_tmp.dev(), u.dev(), &p, QKX_BASIS_UKQCD, QKX_OP_TM_DAG), "tmDagtm");
    vec_tmp.download();
    download();
}

// even-odd Schur operator: tm_xx :2001, tm_xxDagTm_xx :2024, chiral :2037, :2060
static void tm_xx_op(Vector &out, Vector &psi, Gauge &u, LatticeInfo *li, int basis, PARITY parity, int normal,
                     const char *what)
{
    psi.upload();
    qkx_params p = params_of(li);
    check(qkx_tm_xx(out.dev(), psi.dev(), u.dev(), &p, basis, parity_code(parity), normal), what);
    out.download();
}
void Vector::tm_xx(Vector &psi, Gauge &u, LatticeInfo *li, PARITY parity)
{
    tm_xx_op(*this, psi, u, li, QKX_BASIS_UKQCD, parity, 0, "tm_xx");
}
void Vector::tm_xxDagTm_xx(Vector &psi, Gauge &u, LatticeInfo *li, PARITY parity)
{
    tm_xx_op(*this, psi, u, li, QKX_BASIS_UKQCD, parity, 1, "tm_xxDagTm_xx");
}
void Vector::tm_xx_chiral(Vector &psi, Gauge &u, LatticeInfo *li, PARITY parity)
{
    tm_xx_op(*this, psi, u, li, QKX_BASIS_CHIRAL, parity, 0, "tm_xx_chiral");
}
void Vector::tm_xxDagTm_xx_chiral(Vector &psi, Gauge &u, LatticeInfo *li, PARITY parity)
{
    tm_xx_op(*this, psi, u, li, QKX_BASIS_CHIRAL, parity, 1, "tm_xxDagTm_xx_chiral");
}

// ---- Gauge: lib/lattice_util.cpp:649-724 -----------------------------------
Gauge::Gauge(LatticeInfo *latInfo) : LatticeGeometry(latInfo), constGauge(false), dev_(0), uploaded_(false)
{
    size = lV4d;
    size_total = lV4d;
    M = (Complex(*)[NDIM][NCOLORS][NCOLORS])malloc(NLINKS * size_total * sizeof(double));
    if (!M) errorQKXTM("Error out of memory");
}

Gauge::~Gauge()
{
    if (dev_) qkx_gauge_destroy(dev_);
    free(M);
}

qkx_gauge_s *Gauge::dev()
{
    // the host links carry the boundary sign (applyBoundaryCondition), so the
    // kernels apply none of their own
    if (!dev_) check(qkx_gauge_create(&dev_, L, QKX_PREC_DOUBLE, QKX_RECON_18, +1), "Gauge");
    if (!uploaded_ || !constGauge) {
        check(qkx_gauge_upload(dev_, &(M[0][0][0][0].real())), "Gauge upload");
        uploaded_ = true;
    }
    return dev_;
}

void Gauge::hostModified() { uploaded_ = false; }

void Gauge::zero()
{
    memset(&(M[0][0][0][0].real()), 0, NLINKS * size_total * sizeof(double));
    uploaded_ = false;
}
void Gauge::copy(Gauge &src)
{
    memcpy(&(M[0][0][0][0].real()), &(src.M[0][0][0][0].real()), lV4d * NLINKS * sizeof(double));
    uploaded_ = false;
}
double Gauge::norm2()
{
    double s = 0;
    for (long iv = 0; iv < lV4d; iv++)
        for (int d = 0; d < NDIM; d++)
            for (int a = 0; a < NCOLORS; a++)
                for (int b = 0; b < NCOLORS; b++) s += std::norm(M[iv][d][a][b]);
    return s;
}
// Gauge::applyBoundaryCondition :708-724: t-links of the last slice * -1
void Gauge::applyBoundaryCondition(LatticeInfo *latInfo)
{
    if (latInfo->boundaryT != -1) return;
    printfQKXTM("Flip sign in gauge link\n");
    for (long iv = lL[0] * lL[1] * lL[2] * (lL[3] - 1); iv < lV4d; iv++)
        for (int a = 0; a < NCOLORS; a++)
            for (int b = 0; b < NCOLORS; b++) M[iv][3][a][b] = -M[iv][3][a][b];
    uploaded_ = false;
}

// Gauge::calculatePlaquette :923-946
double Gauge::calculatePlaquette()
{
    double r = 0;
    check(qkx_plaquette(&r, dev()), "calculatePlaquette");
    return r;
}

// Gauge::APE_smearing :950-1078: this <- NsmearAPE steps of tmp's spatial
// links, tmp's temporal links; nothing happens for NsmearAPE = 0
void Gauge::APE_smearing(Gauge &tmp, LatticeInfo *latInfo)
{
    if (latInfo->NsmearAPE == 0) return;
    printfQKXTM("Perform APE smearing\n");
    qkx_gauge_s *in = tmp.dev();
    if (!dev_) check(qkx_gauge_create(&dev_, L, QKX_PREC_DOUBLE, QKX_RECON_18, +1), "Gauge");
    check(qkx_ape_smear(dev_, in, latInfo->NsmearAPE, latInfo->alphaAPE), "APE_smearing");
    check(qkx_gauge_download(dev_, &(M[0][0][0][0].real())), "APE_smearing download");
    uploaded_ = true;
}

// qcd_projectSU33d lib/blas.cpp:6-422
void qcd_projectSU33d(qkxTM::Gauge &gf)
{
    qkx_gauge_s *g = gf.dev();
    check(qkx_gauge_project_su3(g), "qcd_projectSU33d");
    check(qkx_gauge_download(g, &(gf.M[0][0][0][0].real())), "qcd_projectSU33d download");
}

// qkxTM_MPI_getGaugeLime lib/read_conf_lime.cpp:111-264 into the host links
int qkxTM_MPI_getGaugeLime(char *fname, qkxTM::Gauge &u)
{
    if (u.initialized == false) errorQKXTM("gauge not initialized\n");
    check(qkx_lime_read_gauge(fname, u.L, &(u.M[0][0][0][0].real())), "qkxTM_MPI_getGaugeLime");
    u.hostModified();
    return 0;
}

// qkxTM_writeGaugeLime lib/write_conf_lime.cpp:49-199 from the host links
int qkxTM_writeGaugeLime(char *fname, qkxTM::Gauge &u, char *message)
{
    check(qkx_lime_write_gauge(fname, u.L, &(u.M[0][0][0][0].real()), message), "qkxTM_writeGaugeLime");
    return 0;
}

// Vector::Gaussian_smearing :1080-1146 (whole lattice, as the reference);
// nothing happens for NsmearGauss = 0.  The reference alternates its steps
// between *this and vec_tmp, so vec_tmp ends with the smeared field (even
// step count) or the iterate before the last step (odd count); both are
// reproduced: n - 1 steps into one field, the last step into the other.
void Vector::Gaussian_smearing(Vector &vec_tmp, Gauge &gauge, LatticeInfo *latInfo)
{
    const int n = latInfo->NsmearGauss;
    const double alpha = latInfo->alphaGauss;
    if (n == 0) return;
    printfQKXTM("Perform Gaussian smearing\n");
    vec_tmp.upload();
    if (n % 2 == 0) {
        check(qkx_gauss_smear(dev(), vec_tmp.dev(), gauge.dev(), n, alpha, -1), "Gaussian_smearing");
        download();
        vec_tmp.copy(*this);
    } else if (n == 1) {
        check(qkx_gauss_smear(dev(), vec_tmp.dev(), gauge.dev(), 1, alpha, -1), "Gaussian_smearing");
        download();
    } else {
        // dev() <- n - 1 steps, vec_tmp.dev() <- the last step
        check(qkx_gauss_smear(dev(), vec_tmp.dev(), gauge.dev(), n - 1, alpha, -1), "Gaussian_smearing");
        check(qkx_gauss_smear(vec_tmp.dev(), dev(), gauge.dev(), 1, alpha, -1), "Gaussian_smearing");
        check(qkx_spinor_download(dev(), &(vec_tmp.M[0][0][0].real())), "Gaussian_smearing");
        check(qkx_spinor_download(vec_tmp.dev(), &(M[0][0][0].real())), "Gaussian_smearing");
    }
}

// Vector::rescale :2314-2321
void Vector::rescale(double a)
{
    upload();
    check(qkx_spinor_rescale(dev(), a), "rescale");
    download();
}

// ---- Propagator: lib/lattice_util.cpp:528-598, 2073-2082, 2323-2453 ---------
Propagator::Propagator(LatticeInfo *latInfo) : LatticeGeometry(latInfo), dev_(0)
{
    size = lV4d;
    size_total = lV4d;
    M = (Complex(*)[NSPINS][NSPINS][NCOLORS][NCOLORS])malloc(NPROPS * size_total * sizeof(double));
    if (!M) errorQKXTM("Error out of memory");
}

Propagator::~Propagator()
{
    if (dev_) qkx_spinor_destroy(dev_);
    free(M);
}

void Propagator::zero() { memset(&(M[0][0][0][0][0].real()), 0, NPROPS * size_total * sizeof(double)); }
void Propagator::copy(Propagator &src)
{
    memcpy(&(M[0][0][0][0][0].real()), &(src.M[0][0][0][0][0].real()), lV4d * NPROPS * sizeof(double));
}
void Propagator::copyVector(Vector &v, int nu, int c2)
{
    for (long iv = 0; iv < lV4d; iv++)
        for (int mu = 0; mu < NSPINS; mu++)
            for (int c1 = 0; c1 < NCOLORS; c1++) M[iv][mu][nu][c1][c2] = v.M[iv][mu][c1];
}
void Propagator::absorbVector(Vector &phi, int nu, int c2) { copyVector(phi, nu, c2); }

double Propagator::norm2()
{
    if (!dev_) check(qkx_spinor_create_batch(&dev_, L, QKX_PREC_DOUBLE, QKX_SUBSET_FULL, 12), "Propagator");
    check(qkx_prop_upload(dev_, &(M[0][0][0][0][0].real())), "Propagator upload");
    double r = 0;
    check(qkx_norm2(&r, dev_), "Propagator norm2");
    return r;
}

void Propagator::rotate(int sign)
{
    if (!dev_) check(qkx_spinor_create_batch(&dev_, L, QKX_PREC_DOUBLE, QKX_SUBSET_FULL, 12), "Propagator");
    check(qkx_prop_upload(dev_, &(M[0][0][0][0][0].real())), "Propagator upload");
    check(qkx_prop_rotate_physical(dev_, sign), "rotateToPhysicalBase");
    check(qkx_prop_download(dev_, &(M[0][0][0][0][0].real())), "Propagator download");
}
void Propagator::rotateToPhysicalBasePlus() { rotate(+1); }
void Propagator::rotateToPhysicalBaseMinus() { rotate(-1); }

// ---- CG: lib/solvers.cpp:8-81 ----------------------------------------------
CG::CG(LatticeInfo *param, qkxTM::Gauge &gauge) : u(gauge), lastIterations(0)
{
    maxIter = (int)param->maxIter;
    tol = param->tol;
    reliableDelta = param->reliableDelta;
}

void CG::operator()(qkxTM::Vector &out, qkxTM::Vector &in, LatticeInfo *param)
{
    int rank = out.rank;
    out.upload();
    in.upload();
    qkx_params p = params_of(param);
    p.tol = tol;
    p.maxIter = maxIter;
    qkx_solver_stats st;
    int rc = qkx_cg(out.dev(), in.dev(), u.dev(), &p, &st);
    if (rc != QKX_SUCCESS && rc != QKX_NOT_CONVERGED) check(rc, "CG");
    lastIterations = st.iterations;
    out.download();
    if (rc == QKX_NOT_CONVERGED) printfQKXTM("Warning solver reached number of max iterations \n");
    else printfQKXTM("CG converged after iteration %d\n", st.iterations);
}

// ---- BLAS: lib/blas_qkxTM.cpp:4-56 -----------------------------------------
double reCdotXdagY(qkxTM::Vector &x, qkxTM::Vector &y)
{
    x.upload();
    if (&y != &x) y.upload();
    double r = 0;
    check(qkx_redot(&r, x.dev(), y.dev()), "reCdotXdagY");
    return r;
}
void X_eq_X_p_aY(qkxTM::Vector &x, double a, qkxTM::Vector &y)
{
    x.upload();
    if (&y != &x) y.upload();
    check(qkx_axpy(x.dev(), a, y.dev()), "X_eq_X_p_aY");
    x.download();
}
void X_eq_aX_p_Y(qkxTM::Vector &x, double a, qkxTM::Vector &y)
{
    x.upload();
    if (&y != &x) y.upload();
    check(qkx_xpay(x.dev(), a, y.dev()), "X_eq_aX_p_Y");
    x.download();
}
