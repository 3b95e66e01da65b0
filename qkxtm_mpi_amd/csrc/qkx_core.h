// qkx_core.h -- internal entry points shared by the library's translation units.
#pragma once
#include <cstddef>
#include <string>
#include "qkx_hip.h"

struct qkx_spinor_s;
struct qkx_gauge_s;
struct qkx_net_s;

namespace qkx {

// hopping kernel with fused epilogue (qkx_dslash.hip):
//   out(par_out) = cR R + gR i g5 R + cX aux + gX i g5 aux, R = H in or H^dag in
//   norm_partial (optional): per-workgroup partial sums of |out|^2, count in *nblocks_out
int hop(qkx_spinor_s *out, const qkx_spinor_s *in, const qkx_spinor_s *aux, const qkx_gauge_s *U,
        int basis, int dagger, int par_out, double cR, double gR, double cX, double gX, int g5out,
        double *norm_partial = nullptr, int *nblocks_out = nullptr);
// the same with per-right-hand-side coefficients coef[r] = {cR, gR, cX, gX}
int hop_batch(qkx_spinor_s *out, const qkx_spinor_s *in, const qkx_spinor_s *aux, const qkx_gauge_s *U,
              int basis, int dagger, int par_out, const double coef[][4], int g5out,
              double *norm_partial = nullptr, int *nblocks_out = nullptr);
// n <= 8 independent single-right-hand-side hops in one launch, instance i
// on U[i]'s lattice: out[i] = epilogue_i(H in[i] (or H^dag), aux[i]) with
// coef[i] = {cR, gR, cX, gX}; aux may be null (for all); norm_partial[i]
// (optional) receives instance i's per-workgroup |out|^2 partials, count in
// nblocks_out[i]; skip[i] (optional) makes instance i's workgroups return
// when *skip[i] != 0.  Fields of one precision, links of one storage.
int hop_multi(int n, qkx_spinor_s *const *out, const qkx_spinor_s *const *in, const qkx_spinor_s *const *aux,
              const qkx_gauge_s *const *U, int dagger, int par_out, const double (*coef)[4], int g5out,
              double *const *norm_partial = nullptr, int *nblocks_out = nullptr, const int *const *skip = nullptr);
// site-local combination (qkx_dslash.hip)
int site(qkx_spinor_s *out, const qkx_spinor_s *y, const qkx_spinor_s *x, int basis, int keep_par,
         double cR, double gR, double cX, double gX, int g5out);
int site_batch(qkx_spinor_s *out, const qkx_spinor_s *y, const qkx_spinor_s *x, int basis, int keep_par,
               const double coef[][4], int g5out);
// The last hop of a CG normal operator with the CG update fused in:
// out = epilogue(hop) is not stored; instead, with alpha = *r2cur / sum(pap),
// x += alpha p and r -= alpha out at every output site, and per-workgroup
// |r|^2 partials go to r2_part (count in *nblocks_out).
struct CgFuse {
    qkx_spinor_s *x, *r, *p;
    const double *r2cur;
    const double *pap_part;
    int npap;
    double *r2_part;
    bool x_in_dir = false;   // update r only; the direction kernel updates x
};
int hop_cg(const qkx_spinor_s *in, const qkx_spinor_s *aux, const qkx_gauge_s *U, int dagger, int par_out,
           double cR, double gR, double cX, double gX, int g5out, const CgFuse &f, int *nblocks_out);
// The CG update fused into a batched hop (FP16 batches of 3-16 RHS on FP16
// links, shared-link kernel): x += alpha p, r -= alpha (hop result), alpha_r
// = r2cur[r] / pap[r] or 0 where !active[r] or r2cur[r] <= stop2[r]; per-
// workgroup per-RHS |r|^2 partials into r2_part.  Device pointers.
struct CgBatchFuse {
    qkx_spinor_s *x, *r, *p;
    const double *r2cur, *stop2;
    const int *active;
    const double *pap;
    double *r2_part;
    bool x_in_dir = false;   // update r only; the direction kernel updates x
};
// max over links of |row2 - conj(row0 x row1)| of an FP64 18-real field,
// measured once per generation of its links
int gauge_recon_deviation(const qkx_gauge_s *U, double *dev);
// the field the solvers use for their internal FP64 operator: U itself, or
// its cached FP64 12-real copy while U's links are SU(3) to 1e-14
int gauge_fp64_operator(const qkx_gauge_s *U, const qkx_gauge_s **op);
int cg_fuse_on();   // qkx_set_cg_fuse: 1 when the CG update is fused into the last hop
bool cg_graph_batched();   // qkx_set_cg_graph: batched solves replay captured iterations
bool hop_cg_batch_available(const qkx_spinor_s *in, const qkx_gauge_s *U, const qkx_spinor_s *x,
                            const qkx_spinor_s *r);
int hop_cg_batch(const qkx_spinor_s *in, const qkx_spinor_s *aux, const qkx_gauge_s *U, int dagger, int par_out,
                 const double coef[][4], int g5out, const CgBatchFuse &f, int *nblocks_out);
// the learned network on even-parity batched fields (qkx_learned.hip):
// net_prep_eo builds the forward and adjoint weight tables from the current
// weights; net_apply_eo: out = N_e in (adjoint: N_e^dag in), out may be in;
// both launches return at once while *skip != 0 (skip optional)
int net_prep_eo(qkx_net_s *n);
int net_apply_eo(qkx_net_s *n, qkx_spinor_s *out, const qkx_spinor_s *in, const qkx_gauge_s *U, int adjoint,
                 const int *skip = nullptr);
void set_hop_tuning(int xcd_remap, int block, int depth64, int depth32);
void set_hop_skip(const int *flag);
void set_rhs_tile(int rt);
void set_site_order(int blocked);
void set_neighbour_staging(int on);
void set_batch_link_sharing(int on);
// hop autotuner (qkx_dslash.hip): 0 off, 1 batched launches, 2 every launch;
// set_no_tune(1) keeps the calling thread's launches on cached or default
// shapes (no scratch fields); tune_report lists the tuned launch kinds
void set_autotune(int mode);
void set_no_tune(int on);
std::string tune_report();
// even-odd Schur operator on parity pe (qkx_ops.cpp):
//   out_pe = [A in - kappa^2 H A^-1 H in]_pe (then g5 if g5out); the optional
//   norm_partial collects |out|^2 per workgroup of the second hop
int schur(qkx_spinor_s *out, const qkx_spinor_s *in, const qkx_gauge_s *U, double kappa, double a, double nrm,
          int basis, int pe, int g5out, qkx_spinor_s *tmp_other, double *norm_partial = nullptr,
          int *nblocks = nullptr);

// qkx_tm_xx in two parts (qkx_ops.cpp): the argument checks (*pe = output
// parity) and the application with caller-provided intermediates (t: the
// other parity; y: parity pe, normal form only)
int tm_xx_check(const qkx_spinor_s *out, const qkx_spinor_s *in, const qkx_gauge_s *U, const qkx_params *p,
                int basis, int parity, int *pe);
int tm_xx_apply(qkx_spinor_s *out, const qkx_spinor_s *in, const qkx_gauge_s *U, const qkx_params *p, int basis,
                int pe, int normal, qkx_spinor_s *t, qkx_spinor_s *y);
// device address of the no-allocation scratch buffer on the library's device
// (qkx_noalloc.hip): taken once per device by qkx_init
int noalloc_setup(int device);

// the batched solver's compensated per-RHS row sums (qkx_solvers_batch.hip)
int sum_rows_dev(const double *part, int nblocks, int nr, double *dst);

// BLAS (qkx_blas.hip): device-resident results, no host synchronisation
int redot_dev(const qkx_spinor_s *x, const qkx_spinor_s *y, double *result);
// *result = sum of partial[0..n) in a fixed order (double-double), one launch
int finish_sum_dev(const double *partial, int n, double *result);
// x = a x + b y; sc_mode 1: b = sc[0], 2: a = sc[0], 3: b = -sc[0]
int axpby_dev(qkx_spinor_s *x, const qkx_spinor_s *y, double a, double b, const double *sc, int sc_mode);
// s = b - s and *s2 = |s|^2 (host value, the bits of qkx_norm2) in one pass
// over FP64 fields; with h (FP16, same sites) also h = s rounded as a copy
// rounds it (qkx_blas.hip)
int resid_norm(qkx_spinor_s *s, const qkx_spinor_s *b, qkx_spinor_s *h, double *s2);

// instrumentation on (qkx_core.cpp)
bool prof_active();

// smearing (qkx_smear.hip): frees the cached APE link workspace
void smear_release();

// fields (qkx_fields.hip)
int zero_parity(qkx_spinor_s *f, int parity);
// device staging buffer for host transfers (grows; reused between calls)
void *stage_buffer(size_t bytes);
// the links of g changed: cached low-precision copies made from g are stale
void gauge_modified(qkx_gauge_s *g);
// *cache holds a conversion of U at precision prec / recon, current with U's links
int gauge_cached(qkx_gauge_s **cache, const qkx_gauge_s *U, int prec, int recon);

// core (qkx_core.cpp)
double *result_slot(int k);
qkx_spinor_s *pool_get(const int L[4], int prec, int subset);
qkx_spinor_s *pool_get_batch(const int L[4], int prec, int subset, int nrhs);
void pool_put(qkx_spinor_s *f);
// destroys a pool field (it must not be in use elsewhere)
void pool_release(qkx_spinor_s *f);

// RAII temporary from the pool
struct Temp {
    qkx_spinor_s *f;
    Temp(const int L[4], int prec, int subset, int nrhs = 1) : f(pool_get_batch(L, prec, subset, nrhs)) {}
    ~Temp() { if (f) pool_put(f); }
    Temp(const Temp &) = delete;
    Temp &operator=(const Temp &) = delete;
};

} // namespace qkx
