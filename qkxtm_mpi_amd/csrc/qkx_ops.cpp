// qkx_ops.cpp -- C API of the Dirac operators: argument checks and the
// mapping of every reference operator onto the fused hop / site kernels.
//
// Reference operators (lib/lattice_util.cpp): twist coefficient
// a = 2 kappa mu twistSign (:1424-1425), A = 1 + i a g5, A^-1 =
// (1 - i a g5)/(1 + (2 kappa mu)^2) (:1588-1613).
#include "qkx_internal.h"
#include "qkx_core.h"

namespace {

double twist_a(const qkx_params *p) { return 2.0 * p->kappa * p->mu * p->twistSign; }
double twist_norm(const qkx_params *p)
{
    double b = 2.0 * p->kappa * p->mu;
    return 1.0 + b * b;
}

bool same_lattice(const qkx_spinor_s *a, const int L[4])
{
    return a->L[0] == L[0] && a->L[1] == L[1] && a->L[2] == L[2] && a->L[3] == L[3];
}

int check_basis(int basis)
{
    if (basis != QKX_BASIS_UKQCD && basis != QKX_BASIS_CHIRAL) {
        qkx::set_error("bad basis");
        return QKX_ERR_ARG;
    }
    return QKX_SUCCESS;
}

// subset that holds parity p, or FULL
bool holds_parity(const qkx_spinor_s *f, int p)
{
    return f->subset == QKX_SUBSET_FULL || (p == 0 && f->subset == QKX_SUBSET_EVEN) ||
           (p == 1 && f->subset == QKX_SUBSET_ODD);
}

} // namespace

namespace qkx {

// Schur operator on parity pe (0 even / 1 odd):
//   out_pe = [A in - kappa^2 H A^-1 H in]_pe  (then g5 if g5out)
// reference tm_xx lib/lattice_util.cpp:2001-2022.
int schur(qkx_spinor_s *out, const qkx_spinor_s *in, const qkx_gauge_s *U, double kappa, double a, double nrm,
          int basis, int pe, int g5out, qkx_spinor_s *tmp_other, double *norm_partial, int *nblocks)
{
    int po = 1 - pe;
    int rc = hop(tmp_other, in, nullptr, U, basis, 0, po, 1.0 / nrm, -a / nrm, 0.0, 0.0, 0);
    if (rc) return rc;
    return hop(out, tmp_other, in, U, basis, 0, pe, -kappa * kappa, 0.0, 1.0, a, g5out, norm_partial, nblocks);
}

// Argument checks of qkx_tm_xx; *pe receives the output parity (0 even, 1 odd).
int tm_xx_check(const qkx_spinor_s *out, const qkx_spinor_s *in, const qkx_gauge_s *U, const qkx_params *p,
                int basis, int parity, int *pe)
{
    QKX_CHECK_ARG(out && in && U && p, "qkx_tm_xx: null argument");
    QKX_CHECK_ARG(out != in, "qkx_tm_xx: out must differ from in");
    QKX_CHECK_ARG(parity == QKX_PARITY_EE || parity == QKX_PARITY_OO, "Error wrong parity choise");
    if (int rc = check_basis(basis)) return rc;
    QKX_CHECK_ARG(same_lattice(out, U->L) && same_lattice(in, U->L), "qkx_tm_xx: lattices differ");
    *pe = parity == QKX_PARITY_EE ? 0 : 1;
    QKX_CHECK_ARG(holds_parity(out, *pe) && holds_parity(in, *pe), "qkx_tm_xx: field subsets do not match parity");
    return QKX_SUCCESS;
}

// qkx_tm_xx on parity pe with the caller's intermediate fields: t on the
// other parity, y (normal only) on parity pe; both of in's precision and
// right-hand-side count.
int tm_xx_apply(qkx_spinor_s *out, const qkx_spinor_s *in, const qkx_gauge_s *U, const qkx_params *p, int basis,
                int pe, int normal, qkx_spinor_s *t, qkx_spinor_s *y)
{
    const double k = p->kappa, a = twist_a(p), n = twist_norm(p);
    int rc;
    if (!normal) {
        rc = schur(out, in, U, k, a, n, basis, pe, 0, t);
    } else {
        // tm_xxDagTm_xx :2024-2035: g5 M(-s) g5 M(s)
        rc = schur(y, in, U, k, a, n, basis, pe, 1, t);
        if (rc) return rc;
        rc = schur(out, y, U, k, -a, n, basis, pe, 1, t);
    }
    if (rc) return rc;
    return zero_parity(out, 1 - pe);
}

} // namespace qkx

extern "C" {

int qkx_dslash(qkx_spinor *out, const qkx_spinor *in, const qkx_gauge *U, int basis, int dagger, int parity)
{
    QKX_CHECK_ARG(out && in && U, "qkx_dslash: null argument");
    QKX_CHECK_ARG(out != in, "qkx_dslash: out must differ from in");
    QKX_CHECK_ARG(same_lattice(out, U->L) && same_lattice(in, U->L), "qkx_dslash: lattices differ");
    if (int rc = check_basis(basis)) return rc;
    QKX_CHECK_ARG(!(dagger && basis == QKX_BASIS_CHIRAL) || in->prec == QKX_PREC_DOUBLE, "qkx_dslash: bad basis");
    if (parity == QKX_PARITY_ALL) {
        QKX_CHECK_ARG(in->subset == QKX_SUBSET_FULL && out->subset == QKX_SUBSET_FULL,
                      "qkx_dslash: the full-lattice hop needs full fields");
        return qkx::hop(out, in, nullptr, U, basis, dagger, -1, 1.0, 0.0, 0.0, 0.0, 0);
    }
    // reference applyDslash_xx (:1272): EO -> even outputs, OE -> odd outputs
    QKX_CHECK_ARG(parity == QKX_PARITY_EO || parity == QKX_PARITY_OE, "Error wrong parity for hopping term");
    int p = parity == QKX_PARITY_EO ? 0 : 1;
    QKX_CHECK_ARG(holds_parity(out, p) && holds_parity(in, 1 - p), "qkx_dslash: field subsets do not match parity");
    int rc = qkx::hop(out, in, nullptr, U, basis, dagger, p, 1.0, 0.0, 0.0, 0.0, 0);
    if (rc) return rc;
    return qkx::zero_parity(out, 1 - p);
}

int qkx_twist(qkx_spinor *out, const qkx_spinor *in, const qkx_params *p, int basis, int kind)
{
    QKX_CHECK_ARG(out && in && p, "qkx_twist: null argument");
    if (int rc = check_basis(basis)) return rc;
    double a = twist_a(p), n = twist_norm(p);
    switch (kind) {
    case QKX_TWIST: return qkx::site(out, in, nullptr, basis, -1, 1.0, a, 0, 0, 0);
    case QKX_TWIST_DAG: return qkx::site(out, in, nullptr, basis, -1, 1.0, -a, 0, 0, 0);
    case QKX_TWIST_INV: return qkx::site(out, in, nullptr, basis, -1, 1.0 / n, -a / n, 0, 0, 0);
    case QKX_TWIST_DAG_INV: return qkx::site(out, in, nullptr, basis, -1, 1.0 / n, a / n, 0, 0, 0);
    case QKX_GAMMA5: return qkx::site(out, in, nullptr, basis, -1, 1.0, 0, 0, 0, 1);
    }
    qkx::set_error("qkx_twist: bad kind");
    return QKX_ERR_ARG;
}

int qkx_twist_add_dslash(qkx_spinor *out, const qkx_spinor *dslash, const qkx_spinor *psi, const qkx_params *p,
                         int basis, int kind, int parity)
{
    QKX_CHECK_ARG(out && dslash && psi && p, "qkx_twist_add_dslash: null argument");
    if (int rc = check_basis(basis)) return rc;
    double k = p->kappa, a = twist_a(p);
    switch (kind) {
    case QKX_XPAY_TM: return qkx::site(out, dslash, psi, basis, -1, -k, 0, 1.0, a, 0);
    case QKX_XPAY_TM_DAG: return qkx::site(out, dslash, psi, basis, -1, -k, 0, 1.0, -a, 0);
    case QKX_XPAY_WILSON: return qkx::site(out, dslash, psi, basis, -1, -k, 0, 1.0, 0, 0);
    case QKX_XPAY_PC: {
        QKX_CHECK_ARG(parity == QKX_PARITY_OO || parity == QKX_PARITY_EE, "Error wrong parity for hopping term");
        int keep = out->subset == QKX_SUBSET_FULL ? (parity == QKX_PARITY_OO ? 1 : 0) : -1;
        return qkx::site(out, dslash, psi, basis, keep, -k * k, 0, 1.0, a, 0);
    }
    }
    qkx::set_error("qkx_twist_add_dslash: bad kind");
    return QKX_ERR_ARG;
}

int qkx_tm(qkx_spinor *out, const qkx_spinor *in, const qkx_gauge *U, const qkx_params *p, int basis, int op)
{
    QKX_CHECK_ARG(out && in && U && p, "qkx_tm: null argument");
    QKX_CHECK_ARG(out != in, "qkx_tm: out must differ from in");
    QKX_CHECK_ARG(in->subset == QKX_SUBSET_FULL && out->subset == QKX_SUBSET_FULL, "qkx_tm: full fields required");
    QKX_CHECK_ARG(same_lattice(out, U->L) && same_lattice(in, U->L), "qkx_tm: lattices differ");
    if (int rc = check_basis(basis)) return rc;
    // FP64 fields: the stencil runs on the cached 12-real copy of U while its
    // links are SU(3) to 1e-14 (row 2 rebuilt as conj(row0 x row1), so the
    // result moves by about 1e-15 relative; 384 instead of 576 link bytes per
    // site); other links, and the other precisions, use U as given
    if (in->prec == QKX_PREC_DOUBLE)
        if (int rc = qkx::gauge_fp64_operator(U, &U)) return rc;
    double k = p->kappa, a = twist_a(p);
    switch (op) {
    case QKX_OP_TM: return qkx::hop(out, in, in, U, basis, 0, -1, -k, 0, 1.0, a, 0);
    case QKX_OP_TM_DAG: return qkx::hop(out, in, in, U, basis, 1, -1, -k, 0, 1.0, -a, 0);
    case QKX_OP_WILSON: return qkx::hop(out, in, in, U, basis, 0, -1, -k, 0, 1.0, 0, 0);
    case QKX_OP_TM_DAG_TM: {
        qkx::Temp t(in->L, in->prec, QKX_SUBSET_FULL, in->nrhs);
        if (!t.f) return QKX_ERR_NOMEM;
        int rc = qkx::hop(t.f, in, in, U, basis, 0, -1, -k, 0, 1.0, a, 0);
        if (rc) return rc;
        return qkx::hop(out, t.f, t.f, U, basis, 1, -1, -k, 0, 1.0, -a, 0);
    }
    }
    qkx::set_error("qkx_tm: bad op");
    return QKX_ERR_ARG;
}

int qkx_tm_xx(qkx_spinor *out, const qkx_spinor *in, const qkx_gauge *U, const qkx_params *p, int basis, int parity,
              int normal)
{
    int pe = 0;
    if (int rc = qkx::tm_xx_check(out, in, U, p, basis, parity, &pe)) return rc;
    const int sub_other = pe == 0 ? QKX_SUBSET_ODD : QKX_SUBSET_EVEN;
    qkx::Temp t(in->L, in->prec, sub_other, in->nrhs);
    if (!t.f) return QKX_ERR_NOMEM;
    if (!normal) return qkx::tm_xx_apply(out, in, U, p, basis, pe, 0, t.f, nullptr);
    qkx::Temp y(in->L, in->prec, pe == 0 ? QKX_SUBSET_EVEN : QKX_SUBSET_ODD, in->nrhs);
    if (!y.f) return QKX_ERR_NOMEM;
    return qkx::tm_xx_apply(out, in, U, p, basis, pe, 1, t.f, y.f);
}

} // extern "C"
