/*
 * qkx_hip.h -- C ABI of libqkx_hip.so: the twisted-mass Wilson operator,
 * its CG solvers and BLAS on one MI355X (gfx950).
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes.  No HIP
 * or C++ types appear here, so the header can be included from the C++98
 * reference-compatible classes (include/qkxtm/) and bound from ctypes.
 *
 * Every entry point names the reference interface it replaces
 * (paths relative to the reference repository root).
 *
 * Conventions (identical to the reference):
 *   site order   iv = ((t*Lz + z)*Ly + y)*Lx + x          include/qkxTM.h:130
 *   host spinor  Complex M[V][4][3]   (re,im interleaved)  include/lattice_util.h:90
 *   host gauge   Complex M[V][4][3][3], mu 0=x 1=y 2=z 3=t include/lattice_util.h:196
 *   parity       EO=0 OE=1 OO=2 EE=3                      include/lattice_util.h:23
 * Device storage is the library's own (parity-blocked struct-of-arrays); the
 * upload/download calls convert.
 *
 * Threading: all calls run on one library stream per process and are
 * serialised by the caller (the reference is single-threaded per rank).
 */
#ifndef QKX_HIP_H
#define QKX_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
/* The reference aborts through errorQKXTM (include/qkxTM.h:137-141) and only
 * prints a warning when CG hits maxIter (lib/solvers.cpp:68-70).  This
 * library returns a status instead; qkx_last_error() gives the message. */
#define QKX_SUCCESS        0
#define QKX_NOT_CONVERGED  1   /* solver hit maxIter (reference: warning only) */
#define QKX_ERR_ARG       -1   /* bad argument (shape, parity, handle)         */
#define QKX_ERR_HIP       -2   /* HIP runtime error                             */
#define QKX_ERR_NOMEM     -3   /* device allocation failed                      */
#define QKX_ERR_NONFINITE -4   /* NaN/Inf met in a solver                       */
#define QKX_ERR_NODEVICE  -5   /* no usable GPU                                 */
#define QKX_ERR_BREAKDOWN -6   /* solver breakdown (p^dag A p <= 0)             */

/* ---- enums ------------------------------------------------------------- */
#define QKX_PARITY_ALL  -1
#define QKX_PARITY_EO    0     /* hopping: even output sites (odd zeroed)      */
#define QKX_PARITY_OE    1     /* hopping: odd output sites (even zeroed)      */
#define QKX_PARITY_OO    2     /* twist-PC / Schur: odd sites                  */
#define QKX_PARITY_EE    3     /* twist-PC / Schur: even sites                 */

#define QKX_BASIS_UKQCD  0     /* include/hopping_term.h                        */
#define QKX_BASIS_CHIRAL 1     /* include/hopping_term_chiral.h                 */

#define QKX_PREC_DOUBLE  0
#define QKX_PREC_SINGLE  1
#define QKX_PREC_HALF    2     /* FP16 storage + one FP32 scale per site        */
#define QKX_PREC_BF16    3     /* BF16 storage (gauge and spinor)               */

#define QKX_RECON_18    18     /* full 3x3 complex link                         */
#define QKX_RECON_12    12     /* two rows stored, third = conj(row0 x row1)    */

#define QKX_SUBSET_FULL  0     /* all V sites                                   */
#define QKX_SUBSET_EVEN  1     /* V/2 even sites                                */
#define QKX_SUBSET_ODD   2     /* V/2 odd sites                                 */

/* site-local twist family (qkx_twist) */
#define QKX_TWIST          0   /* A      = 1 + i a g5     applyTwist        */
#define QKX_TWIST_DAG      1   /* A^dag  = 1 - i a g5     applyTwistDag     */
#define QKX_TWIST_INV      2   /* A^-1                     applyTwistInv     */
#define QKX_TWIST_DAG_INV  3   /* (A^dag)^-1               applyTwistDagInv  */
#define QKX_GAMMA5         4   /* g5                       applyGamma5       */

/* hop + site term (qkx_twist_add_dslash) */
#define QKX_XPAY_TM        0   /* -k D + psi + i a g5 psi   applyTwistAddDslash       */
#define QKX_XPAY_TM_DAG    1   /* -k D + psi - i a g5 psi   applyTwistDagAddDslashDag */
#define QKX_XPAY_WILSON    2   /* psi - k D                 applyDeltaAddDslash       */
#define QKX_XPAY_PC        3   /* -k^2 D + (psi + i a g5 psi) on one parity  applyTwistAddDslash_PC */

/* full-lattice operators (qkx_tm) */
#define QKX_OP_TM          0   /* Vector::tm / tm_chiral */
#define QKX_OP_TM_DAG      1   /* Vector::tmDag          */
#define QKX_OP_TM_DAG_TM   2   /* Vector::tmDagtm        */
#define QKX_OP_WILSON      3   /* Vector::Wilson         */

/* ---- parameters (hot-path subset of LatticeInfo, include/qkxTM.h:156-204) */
typedef struct {
    int    L[4];           /* LatticeInfo::L[4]          qkxTM.h:166          */
    double kappa;          /* LatticeInfo::kappa         qkxTM.h:170          */
    double mu;             /* LatticeInfo::mu            qkxTM.h:172          */
    int    twistSign;      /* LatticeInfo::twistSign     qkxTM.h:173 (+1/-1)  */
    int    boundaryT;      /* LatticeInfo::boundaryT     qkxTM.h:184; see qkx_gauge_create */
    double tol;            /* LatticeInfo::tol           qkxTM.h:174          */
    int    maxIter;        /* LatticeInfo::maxIter       qkxTM.h:175          */
    double reliableDelta;  /* LatticeInfo::reliableDelta qkxTM.h:176          */
} qkx_params;

typedef struct qkx_gauge_s  qkx_gauge;
typedef struct qkx_spinor_s qkx_spinor;

/* ---- library --------------------------------------------------------------- */
/* The entries of this section (device, stream, timing, instrumentation, launch
 * tuning, memory) have no reference counterpart: the reference has no device.
 * Selects the device and creates the library stream.  Idempotent. */
int  qkx_init(int device);
int  qkx_finalize(void);
const char *qkx_last_error(void);
int  qkx_device_count(int *n);
int  qkx_sync(void);
/* Device memory of the current device, in bytes: in use (by every allocation
 * on it, the library's field pool included) and total. */
int  qkx_mem_info(double *used_bytes, double *total_bytes);
/* HIP events on the library stream, for timing (ms between start and stop). */
int  qkx_timer_start(void);
int  qkx_timer_stop(double *ms);
/* Per-kernel instrumentation: when on, every hopping-kernel launch is
 * bracketed by events; qkx_prof_read returns launches and summed ms of the
 * hopping kernels since the last reset. */
int  qkx_prof_enable(int on);
int  qkx_prof_read(int *launches, double *ms, double *bytes);
/* the same restricted to hopping kernels on spinors of one precision */
int  qkx_prof_read_prec(int prec, int *launches, double *ms, double *bytes);
int  qkx_prof_reset(void);
/* Launch shape of the hopping kernel: workgroup size (64/128/256), whether
 * workgroups are remapped so that each XCD walks a contiguous site range, and
 * the software-pipeline depth (1 or 2 hops in flight) of the FP64 and of the
 * FP32/FP16/BF16 kernels. */
int  qkx_set_tuning(int xcd_remap, int block, int depth64, int depth32);
/* Order in which the hopping kernels walk the sites of a parity.  1: blocked
 * -- x, then blocks of up to 8 y rows and up to 8 t slices, then z -- so that
 * neighbour spinors and backward links are reused from L2; 0: checkerboard
 * index order; 2 (default): blocked for the full-lattice FP64 hops (tm and
 * its family, the full-lattice CG), checkerboard order for the rest.  Results
 * per site are identical; reductions sum in another order. */
int  qkx_set_site_order(int blocked);
/* Single-RHS FP16 hops on one parity (FP16 12-real links, checkerboard site
 * order, x row of at most 64 sites): 1 stages each workgroup's x and y
 * neighbour spinors in LDS with one coalesced read of the input range they
 * span, 2 does the same per wave (no workgroup barrier); 0 (default): every
 * thread gathers its own.  The results are identical. */
int  qkx_set_neighbour_staging(int mode);
/* Hops on batches of 3-16 FP16 right-hand sides with FP16 links can fetch
 * and reconstruct each site's 8 links once per workgroup, into LDS, for all
 * the site's right-hand sides (and fuse the batched CG update into the last
 * hop).  1 (default): when the launch holds at least 4 waves per SIMD of
 * (site, RHS) units; 2: always; 0: never (every thread fetches its links).
 * The hop results are identical. */
int  qkx_set_batch_link_sharing(int on);
/* The hop autotuner: on the first launch of a kind (spinor and link
 * storage, right-hand sides, lattice sites, full or half lattice, dagger,
 * xpay input, basis) it times the candidate launch shapes -- workgroup size
 * (256 or 128 threads), links shared through LDS (batches of 3-16 RHS),
 * pipeline depth -- on scratch fields of that kind and keeps the fastest for
 * every later launch of the kind.  Each shape computes every output site
 * with the same arithmetic: the same bits for FP16 and FP64 fields; FP32 and
 * BF16 batches with shared links may round a contracted product-sum
 * differently in their last bit.  The per-workgroup norm partials group
 * sites differently.  mode 0: off (the fixed defaults), 1 (default):
 * batched launches (nrhs > 1), 2: every launch.  Setting a mode clears the
 * cache.  An explicit qkx_set_batch_link_sharing(0 or 2) wins over a tuned
 * choice.  No tuning happens under graph capture or instrumentation. */
int qkx_set_autotune(int mode);
/* The tuned kinds, one line each ("... : block=B share=S depth=D T us
 * (default T0 us)"), NUL-terminated into buf (len bytes); *needed receives
 * the full length + 1. */
int qkx_autotune_report(char *buf, int len, int *needed);
/* CG iterations enqueued between host reads of the solver state (default 16;
 * 1 gives no launches past the stopping iteration). */
int  qkx_set_cg_batch(int iterations);
/* 1: the CG update x += alpha p, r -= alpha Ap is fused into the last hop
 * of the normal operator (no Ap field, one launch less per iteration);
 * 2 (default): as 1, but where x is FP32 (the inner CG of the mixed-precision
 * solves, single and batched) x += alpha p moves into the direction kernel,
 * which reads p anyway (the same bits, one read of p less per iteration);
 * 0: separate update kernel. */
int  qkx_set_cg_fuse(int mode);
/* CG iterations replayed two at a time from captured HIP graphs (one launch
 * per pair) instead of every kernel launched on its own; the results are
 * identical.  mode bit 0: the single-RHS solvers (qkx_solve, qkx_cg), bit 1:
 * the batched solvers (qkx_solve_batch, qkx_solve_learned).  Default 0: on
 * MI355X the eager launches already run ahead of the GPU, and the replay
 * measured no faster. */
int  qkx_set_cg_graph(int mode);
/* Right-hand sides per thread in batched hops (1..4, each link fetched once
 * per tile; a tile that does not divide nrhs falls back to 1; 0 = automatic,
 * which is 1: wider tiles were measured slower on MI355X). */
int  qkx_set_rhs_tile(int rt);

/* ---- fields (Gauge / Vector, include/lattice_util.h:80-212) ---------------- */
/* Gauge::Gauge lib/lattice_util.cpp:649-678.  boundaryT = -1 multiplies the
 * t-links of the last time slice by -1 inside the kernels, which is what
 * Gauge::applyBoundaryCondition (lib/lattice_util.cpp:708-724) does to the
 * stored links; use +1 when the uploaded links already carry the sign.
 * RECON_12 needs SU(3) links. */
int qkx_gauge_create(qkx_gauge **g, const int L[4], int prec, int recon, int boundaryT);
int qkx_gauge_destroy(qkx_gauge *g);
/* Host AoS (reference layout) -> device.  A RECON_12 field rebuilds row 2 as
 * conj(row0 x row1): links that carry the antiperiodic sign (a field created
 * with boundaryT = +1) are stored with the last time slice's t-links negated
 * and the field takes boundaryT = -1 (the same operator); links that are not
 * SU(3) (to 1e-14 for FP64, 1e-6 below) are refused with QKX_ERR_ARG. */
int qkx_gauge_upload(qkx_gauge *g, const double *U_aos);
/* Device -> device precision/compression change (for mixed-precision
 * solvers); src is FP64 recon 18.  A RECON_12 destination follows the rule
 * of qkx_gauge_upload (sign-carrying links flipped, dst takes boundaryT = -1;
 * non-SU(3) links refused). */
int qkx_gauge_convert(qkx_gauge *dst, const qkx_gauge *src);
/* Device -> host AoS (reference layout), FP64 recon-18 fields; the links as
 * stored (the boundaryT = -1 sign is applied by the kernels, not stored). */
int qkx_gauge_download(const qkx_gauge *g, double *U_aos);
/* The largest |row2 - conj(row0 x row1)| over all links of an FP64 18-real
 * field: how far a 12-real copy (row 2 rebuilt) would move any link.  Below
 * 1e-14 the solvers (qkx_solve, qkx_solve_batch) run their internal FP64
 * operator on such a copy (25% fewer bytes per hop); the reported true
 * residual always uses the field itself.  Links that carry the boundary sign
 * or are not SU(3) give O(1) (sign-carrying links are copied with the sign
 * removed and boundaryT = -1 instead).  Measured once per change of the
 * links.  No reference counterpart. */
int qkx_gauge_su3_deviation(const qkx_gauge *g, double *dev);

/* Vector::Vector lib/lattice_util.cpp:382-413 */
int qkx_spinor_create(qkx_spinor **s, const int L[4], int prec, int subset);
/* A batch of nrhs (<= QKX_MAX_RHS) fields on one lattice, stored with the
 * right-hand side innermost so one stencil launch serves all of them and each
 * link is read once per site (no reference counterpart: the reference solves
 * the 12 propagator columns one after another, bin/test_invert_prop.cpp:83-97).
 * Every single-field call also accepts a batch and acts on all its members. */
#define QKX_MAX_RHS 16
int qkx_spinor_create_batch(qkx_spinor **s, const int L[4], int prec, int subset, int nrhs);
int qkx_spinor_destroy(qkx_spinor *s);
/* Host AoS Complex[V][4][3] <-> device.  A parity subset takes its own sites
 * from the full host array; download writes zeros on the other parity. */
int qkx_spinor_upload(qkx_spinor *s, const double *psi_aos);
int qkx_spinor_download(const qkx_spinor *s, double *psi_aos);
/* one member of a batch */
int qkx_spinor_upload_rhs(qkx_spinor *s, int rhs, const double *psi_aos);
int qkx_spinor_download_rhs(const qkx_spinor *s, int rhs, double *psi_aos);
/* Vector::zero :443-447, Vector::copy :449-453 (copy also converts precision
 * and moves a parity between full and subset fields). */
int qkx_spinor_zero(qkx_spinor *s);
/* Gaussian noise, variance 1 per real component, generated on the device
 * from (seed, entry, component) by a counter-based hash: the learned
 * preconditioner's random training vectors.  No reference counterpart. */
int qkx_spinor_gaussian(qkx_spinor *s, unsigned long long seed);
int qkx_spinor_copy(qkx_spinor *dst, const qkx_spinor *src);

/* ---- Dirac operators (Vector methods, include/lattice_util.h:102-137) ------ */
/* applyDslash :1158, applyDslash_chiral :1215, applyDslashDag :1643,
 * applyDslash_xx :1272, applyDslash_xx_chiral :1346 (lib/lattice_util.cpp).
 * parity: QKX_PARITY_ALL, _EO (even outputs) or _OE (odd outputs); on a full
 * output field the other parity is zeroed, as in the reference.  out != in. */
int qkx_dslash(qkx_spinor *out, const qkx_spinor *in, const qkx_gauge *U,
               int basis, int dagger, int parity);

/* applyTwist :1563, applyTwistDag :1726, applyTwistInv :1588,
 * applyTwistInv_chiral :1615, applyTwistDagInv :1750, applyGamma5 :1778,
 * applyGamma5_chiral :1800.  out may alias in. */
int qkx_twist(qkx_spinor *out, const qkx_spinor *in, const qkx_params *p,
              int basis, int kind);

/* applyTwistAddDslash :1420 (+_chiral :1444), applyTwistDagAddDslashDag :1701,
 * applyDeltaAddDslash :1552, applyTwistAddDslash_PC :1468 (+_chiral :1510;
 * parity OO/EE).  out may alias dslash. */
int qkx_twist_add_dslash(qkx_spinor *out, const qkx_spinor *dslash,
                         const qkx_spinor *psi, const qkx_params *p,
                         int basis, int kind, int parity);

/* tm :1925, tm_chiral :1930, tmDag :1935, tmDagtm :1940, Wilson :1949.
 * The hop and the site term run as one fused kernel.  out != in.  FP64 fields
 * on FP64 18-real links run on the library's cached 12-real copy while the
 * links are SU(3) to 1e-14 (the result moves by about 1e-15 relative); see
 * qkx_gauge_su3_deviation. */
int qkx_tm(qkx_spinor *out, const qkx_spinor *in, const qkx_gauge *U,
           const qkx_params *p, int basis, int op);

/* tm_xx :2001, tm_xx_chiral :2037 (normal = 0) and tm_xxDagTm_xx :2024,
 * tm_xxDagTm_xx_chiral :2060 (normal = 1): the even-odd Schur operator
 * A - k^2 H A^-1 H on parity EE or OO.  Full fields in and out (other parity
 * of out zeroed), or parity-subset fields of the chosen parity. */
int qkx_tm_xx(qkx_spinor *out, const qkx_spinor *in, const qkx_gauge *U,
              const qkx_params *p, int basis, int parity, int normal);

/* No reference counterpart.  For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): qkx_tm_xx with the same arguments and results, but its
 * intermediate fields (the other parity, and for normal = 1 the first Schur
 * application) live in a static device buffer of QKX_NOALLOC_SCRATCH_BYTES
 * that each GPU holds from qkx_init until the process ends, instead of the
 * library's field pool.  The call allocates nothing; it needs qkx_init first
 * and instrumentation off, and refuses (QKX_ERR_ARG) fields whose
 * intermediates exceed the buffer: (normal ? 2 : 1) x V/2 x nrhs x the
 * bytes per site of in's precision (FP64 192, FP32 96, FP16 52, BF16 48).
 * Calls that overlap on one GPU share the buffer and race. */
#define QKX_NOALLOC_SCRATCH_BYTES (64u << 20)
int qkx_tm_xx_noalloc(qkx_spinor *out, const qkx_spinor *in, const qkx_gauge *U,
                      const qkx_params *p, int basis, int parity, int normal);

/* ---- BLAS (include/blas_qkxTM.h:3-5, lib/blas_qkxTM.cpp) -------------------- */
int qkx_redot(double *res, const qkx_spinor *x, const qkx_spinor *y); /* reCdotXdagY :4  */
int qkx_axpy(qkx_spinor *x, double a, const qkx_spinor *y);          /* X_eq_X_p_aY :31 */
int qkx_xpay(qkx_spinor *x, double a, const qkx_spinor *y);          /* X_eq_aX_p_Y :45 */
int qkx_norm2(double *res, const qkx_spinor *x);                     /* Vector::norm2 lattice_util.cpp:464 */

/* ---- solvers ---------------------------------------------------------------- */
typedef struct {
    int    iterations;        /* inner iterations (all precisions)             */
    int    reliable_updates;  /* FP64 residual recomputations                  */
    int    converged;         /* 1 if the stopping test passed                 */
    double iterated_r2;       /* final iterated |r|^2 (normal equations)       */
    double true_residual;     /* |b - D x| / |b| recomputed in FP64 (-1: n/a)  */
    double seconds;           /* device wall time of the solve                 */
    int    fallback_iterations; /* qkx_solve_gcr: CG iterations after a stall  */
} qkx_solver_stats;

/* CG::CG + CG::operator() (include/solvers.h:12,21; lib/solvers.cpp:8-81):
 * CG on D^dag D x = b over the full lattice, UKQCD basis, FP64, x accumulated
 * onto its input value with r0 = b (the reference assumes x = 0), stop when
 * |r|^2 <= tol^2 |b|^2 or after maxIter iterations.  true_residual is left at
 * -1 (the reference computes none). */
int qkx_cg(qkx_spinor *x, const qkx_spinor *b, const qkx_gauge *U,
           const qkx_params *p, qkx_solver_stats *st);

/* qkx_cg that also copies the iterate x into store[i] after every `every`
 * iterations (i < nstore; *nstored receives the count).  No reference
 * counterpart: it produces the (residual, error) training pairs of the
 * learned preconditioner from plain-CG iterates. */
int qkx_cg_record(qkx_spinor *x, const qkx_spinor *b, const qkx_gauge *U,
                  const qkx_params *p, int every, qkx_spinor **store,
                  int nstore, int *nstored, qkx_solver_stats *st);

/* Solve D x = b (UKQCD basis) to true relative residual |b - D x|/|b| <= tol.
 * No reference counterpart: the reference has no even-odd, mixed-precision
 * or true-residual solver.  Options: */
typedef struct {
    int    eo;               /* 1: even-odd Schur system on the even sites      */
    int    inner_prec;       /* QKX_PREC_* of the inner CG (DOUBLE = pure FP64) */
    int    inner_recon;      /* gauge compression of the inner operator         */
    double delta;            /* reliable-update trigger (residual drop)         */
    int    check_every;      /* FP64 true-residual check period (pure FP64)     */
    int    maxIter;          /* iteration cap over all restarts                 */
    double tol;              /* target true relative residual                   */
} qkx_solve_opts;

/* U must be FP64 RECON_18.  U_inner (may be NULL -> derived internally and
 * cached) is the gauge used by the inner operator. */
int qkx_solve(qkx_spinor *x, const qkx_spinor *b, const qkx_gauge *U,
              const qkx_gauge *U_inner, const qkx_params *p,
              const qkx_solve_opts *o, qkx_solver_stats *st);

/* Solve D_r x_r = b_r for every right-hand side r of a batch (x, b from
 * qkx_spinor_create_batch, FP64, full lattice) in lockstep: one stencil launch
 * serves the whole batch and reads each link once per site.  This replaces the
 * reference's loop of single solves over the 12 spin-colour sources of a
 * propagator (bin/test_invert_prop.cpp:83-97).  mu (may be NULL -> p->mu for
 * all) gives each right-hand side its own twisted mass on the shared gauge
 * field.  Right-hand sides that meet tol are frozen while the others go on;
 * reliable updates are taken for the whole batch.  st (may be NULL) receives
 * nrhs entries; iterations is the batch's iteration count. */
int qkx_solve_batch(qkx_spinor *x, const qkx_spinor *b, const qkx_gauge *U,
                    const qkx_gauge *U_inner, const qkx_params *p,
                    const double *mu, const qkx_solve_opts *o,
                    qkx_solver_stats *st);

/* Solve several independent systems D_i x_i = b_i -- each on its own
 * lattice (uneven volumes allowed), gauge field U[i] and parameters p[i]
 * (kappa, mu, twistSign, boundaryT) -- in lockstep, even-odd preconditioned
 * (o->eo must be 1) with the inner precision, reliable-update trigger,
 * tolerance and iteration cap of o.  Every CG step is one launch of each
 * kernel for all n <= 8 instances: the four hops of the normal operator, the
 * update and the direction step, each workgroup working for one instance.
 * Each instance keeps its own scalars and stops on its own true residual;
 * reliable updates are taken per instance.  x[i], b[i]: single full FP64
 * fields on U[i]'s lattice; U[i]: FP64 recon 18 (all SU(3) or none, so that
 * their working copies share one storage); st (may be NULL) receives n
 * entries.  No reference counterpart: the reference makes one solve per call
 * (lib/solvers.cpp:16-81). */
int qkx_solve_multi(int n, qkx_spinor *const *x, const qkx_spinor *const *b, const qkx_gauge *const *U,
                    const qkx_params *p, const qkx_solve_opts *o, qkx_solver_stats *st);

/* ---- learned preconditioner ----------------------------------------------- */
/* No reference counterpart.  M = L2 L1, two gauge-equivariant
 * parallel-transport convolutions (Lehner and Wettig, Phys. Rev. D 108,
 * 034503 (2023)) with paths of length 0 and 1:
 *   out_a(x) = sum_p sum_b W[p][a][b] (T_p in_b)(x),  T_0 = 1,
 *   T_{1+2mu} f(x) = U_mu(x) f(x+mu),  T_{2+2mu} f(x) = U_mu(x-mu)^dag f(x-mu),
 * on 4 -> ch -> 4 colour-vector channels (the 4 spin components of a spinor).
 * Weights: W1[9][ch][4] then W2[9][4][ch], complex, as interleaved FP32 on
 * the host (2 * qkx_net_nweights floats).  The network's gauge field is FP32
 * RECON_18 (its boundaryT as for the operators). */
typedef struct qkx_net_s qkx_net;
int qkx_net_create(qkx_net **n, int ch);                  /* ch = 4, 8 or 16 */
int qkx_net_destroy(qkx_net *n);
int qkx_net_nweights(const qkx_net *n, int *ncomplex);
/* Frees the network's work buffers (training features on its last lattice,
 * the even-odd hidden channels); the next call that needs them allocates
 * them again.  Weights and optimiser state stay. */
int qkx_net_release_buffers(qkx_net *n);
/* Bound on the even-odd hidden-channel buffer (bytes; 0: 4 GiB): batches
 * whose hidden channels do not fit are processed a few right-hand sides at a
 * time (at least one), with the same results. */
int qkx_net_set_hidden_limit(qkx_net *n, unsigned long long bytes);
int qkx_net_set_weights(qkx_net *n, const float *w);      /* resets the optimiser */
int qkx_net_get_weights(const qkx_net *n, float *w);
/* forward passes read BF16 weights (1, default; FP32 accumulation) or FP32 (0) */
int qkx_net_set_bf16(qkx_net *n, int on);
/* out = M in; single full-lattice FP64 or FP32 spinors */
int qkx_net_apply(qkx_net *n, qkx_spinor *out, const qkx_spinor *in,
                  const qkx_gauge *U);
/* out = N_e in (adjoint: N_e^dag in), N_e = P_e M P_e the network on the even
 * sites: in and out even-parity fields (single or batched, any precision,
 * the same for both; out may be in).  With the input zero on the odd sites,
 * the hidden channels are needed at the odd sites only, and C0 = W2[0] W1[0]
 * carries the rest. */
int qkx_net_apply_eo(qkx_net *n, qkx_spinor *out, const qkx_spinor *in,
                     const qkx_gauge *U, int adjoint);
/* Adds the gradient of |M r - e|^2 with respect to the weights
 * (dL/dRe W + i dL/dIm W) and the loss to the network's accumulators. */
int qkx_net_accumulate_grad(qkx_net *n, const qkx_spinor *r,
                            const qkx_spinor *e, const qkx_gauge *U);
/* accumulated gradient (2 * nweights floats, may be NULL) and loss (may be NULL) */
int qkx_net_read_grad(qkx_net *n, float *g, double *loss);
int qkx_net_zero_grad(qkx_net *n);
/* Adam on the FP32 master weights with the accumulated gradient / count;
 * clears the gradient (the loss keeps accumulating until qkx_net_zero_grad)
 * and refreshes the BF16 copy. */
int qkx_net_adam(qkx_net *n, double lr, double beta1, double beta2, double eps,
                 int count);

typedef struct {
    int    restart;   /* GCR cycle length (1..16)                               */
    int    maxIter;   /* iteration cap (GCR + fallback CG)                       */
    double tol;       /* target true relative residual                          */
    double stall;     /* a cycle reducing |r| by less than this factor falls back
                         to the even-odd mixed-precision CG (0: never)          */
} qkx_gcr_opts;

/* Flexible GCR on D x = b (full lattice, UKQCD, FP64 vectors) with the
 * network as right preconditioner (net may be NULL: plain GCR).  U_net is
 * the network's FP32 gauge field.  No reference counterpart. */
int qkx_solve_gcr(qkx_spinor *x, const qkx_spinor *b, const qkx_gauge *U,
                  const qkx_gauge *U_net, qkx_net *net, const qkx_params *p,
                  const qkx_gcr_opts *o, qkx_solver_stats *st);

/* The learned, mixed-precision, multi-RHS solver.  Solves D_r x_r = b_r for
 * every right-hand side r of a batch (as qkx_solve_batch: x, b full FP64
 * batches, one mu) on the even-odd Schur system Mh, with the network as
 * right preconditioner: the inner CG (low precision, o->inner_prec) runs on
 * (Mh N)^dag (Mh N) y = (Mh N)^dag bh with N = P_e M P_e, the network
 * restricted to the even sites (its hidden channels live on the odd sites),
 * and x = N y is accumulated in FP64 at every reliable update (o->delta).  The
 * FP64 true residual judges: when a reliable-update cycle cuts the worst true
 * residual of the batch by less than `stall` (0 < stall <= 1), the solve goes
 * on as the plain CG from its current solution (st->fallback_iterations
 * counts those iterations).  net = NULL is qkx_solve_batch.  o->eo must be 1
 * and o->inner_prec FP32, FP16 or BF16; U_net is the network's FP32 gauge
 * field.  No reference counterpart: it replaces the reference's loop of
 * single CG solves over a propagator's 12 sources
 * (bin/test_invert_prop.cpp:83-97, lib/solvers.cpp:16-81). */
int qkx_solve_learned(qkx_spinor *x, const qkx_spinor *b, const qkx_gauge *U,
                      const qkx_gauge *U_net, qkx_net *net, const qkx_params *p,
                      const qkx_solve_opts *o, double stall, qkx_solver_stats *st);

/* No reference counterpart: the solvers' global reduction, exposed so that
 * its compensation can be checked (the reference sums with MPI_Allreduce,
 * lib/blas_qkxTM.cpp:24).  *res = the sum of the n doubles at vals (host),
 * added on the device the way the CG kernels add their per-workgroup
 * partials: in double-double, in a fixed order, every consuming workgroup
 * alike.  mode 0: eight workgroups each summing all n (the consumer side of
 * the single-RHS CG; they must agree bit for bit), 1: the one-workgroup
 * pre-sum used when partials are many, 2: the batched solver's row sum
 * (two stages beyond 256 rows). */
int qkx_reduce_sum(const double *vals, int n, int mode, double *res);

/* Relative true residual |b - D x| / |b| in FP64, full lattice, UKQCD. */
int qkx_true_residual(double *res, const qkx_spinor *x, const qkx_spinor *b,
                      const qkx_gauge *U, const qkx_params *p);

/* ---- source preparation (include/lattice_util.h:101, 209-210, 325) ---------
 * FP64 like the reference; gauge fields FP64 recon 18.  A write to a gauge
 * field here (or by upload/convert) refreshes the low-precision copies the
 * mixed solvers cache from it. */
/* Gauge::APE_smearing lib/lattice_util.cpp:950-1078: the spatial links of out
 * become nsmear APE steps of in's, U_i -> Proj_SU3[U_i + alpha * (sum of the
 * four spatial staples)]; the temporal links of out become in's.  nsmear = 0
 * leaves out untouched, as the reference does.  out != in, same lattice and
 * boundaryT; in is not modified (the reference uses its U_tmp as scratch). */
int qkx_ape_smear(qkx_gauge *out, const qkx_gauge *in, int nsmear, double alpha);
/* qcd_projectSU33d lib/blas.cpp:6-422: the spatial links of g, in place. */
int qkx_gauge_project_su3(qkx_gauge *g);
/* Gauge::calculatePlaquette lib/lattice_util.cpp:923-946: mean of
 * Re tr(plaquette)/3 over sites and the six planes. */
int qkx_plaquette(double *plaq, const qkx_gauge *g);
/* Vector::Gaussian_smearing lib/lattice_util.cpp:1080-1146:
 *   out <- nsmear steps of psi -> (psi + alpha H3 psi) / (1 + 6 alpha),
 *   H3 psi(x) = sum_{i<3} [U_i(x) psi(x+i) + U_i(x-i)^dag psi(x-i)].
 * in/out: full FP64 fields (batches smear every member); out != in.
 * t = -1 smears the whole lattice (the reference); t >= 0 smears time slice t
 * only and zeroes out elsewhere (the same result when in vanishes off slice
 * t, as a point source does, at 1/L_t of the cost).  nsmear = 0 leaves out
 * untouched, as the reference does. */
int qkx_gauss_smear(qkx_spinor *out, const qkx_spinor *in, const qkx_gauge *U, int nsmear, double alpha,
                    int t);

/* ---- propagators (Propagator, include/lattice_util.h:158-185) ---------------
 * A propagator is a batch field of 12 members (qkx_spinor_create_batch, FP64,
 * full lattice): member 3 nu + c2 is the column of source spin nu and source
 * colour c2, i.e. Propagator::M[iv][mu][nu][c1][c2] = member(3 nu + c2)[iv][mu][c1].
 * The batched solver's solution of the 12 spin-colour point sources (member
 * 3 nu + c2 = point source (nu, c2)) is such a field.  Propagator::zero, copy
 * and norm2 are qkx_spinor_zero, qkx_spinor_copy and qkx_norm2 on it. */
/* Host Complex[V][4][4][3][3] (reference layout) <-> device. */
int qkx_prop_upload(qkx_spinor *prop, const double *M_aos);
int qkx_prop_download(const qkx_spinor *prop, double *M_aos);
/* Propagator::absorbVector lib/lattice_util.cpp:2073-2082 (= copyVector
 * :574-581): column (nu, c2) <- phi, a single full FP64 field. */
int qkx_prop_absorb(qkx_spinor *prop, const qkx_spinor *phi, int nu, int c2);
/* Propagator::rotateToPhysicalBasePlus / Minus :2323-2453, sign +1 / -1:
 * S -> (1 + sign i g5) S (1 + sign i g5) / 2 per site and colour pair, UKQCD g5. */
int qkx_prop_rotate_physical(qkx_spinor *prop, int sign);
/* Vector::rescale :2314-2321: x <- a x (every member of a batch). */
int qkx_spinor_rescale(qkx_spinor *x, double a);

/* ---- ILDG/LIME gauge files (include/lattice_util.h:319-320) -----------------
 * Our own LIME record parser (the reference uses c-lime).  Host arrays in the
 * reference layout Complex[V][4][3][3]; the file's big-endian order t,z,y,x,
 * mu,row,col is that layout byte-swapped. */
/* qkxTM_MPI_getGaugeLime lib/read_conf_lime.cpp:111-264: the links of the
 * "ildg-binary-data" record, precision from "ildg-format" (64, as the
 * reference; 32 also accepted and widened).  The lattice in the XML and the
 * payload size must match L. */
int qkx_lime_read_gauge(const char *path, const int L[4], double *U_aos);
/* qkxTM_writeGaugeLime lib/write_conf_lime.cpp:49-199: the reference's three
 * records ("ildg-format" with its XML text, "xlf-info" = message,
 * "ildg-binary-data" in 64-bit), byte for byte. */
int qkx_lime_write_gauge(const char *path, const int L[4], const double *U_aos, const char *message);
/* The same into / from a device gauge field (FP64 recon 18 for writing). */
int qkx_gauge_read_lime(qkx_gauge *g, const char *path);
int qkx_gauge_write_lime(const qkx_gauge *g, const char *path, const char *message);

/* ---- synthetic inputs (seeded stand-ins for configurations and sources) ---- */
/* Counter-hash RNG keyed by (seed, site, direction): thread-order independent.
 * Gauge: U = exp(i eps H), H Gaussian Hermitian traceless, reunitarised to
 * SU(3).  Z2: (+-1 +-i)/sqrt2 per component.  All in host reference layout. */
int qkx_gen_gauge(double *U_aos, const int L[4], double eps, unsigned long long seed);
int qkx_gen_z2(double *psi_aos, const int L[4], unsigned long long seed);
int qkx_gen_gaussian(double *psi_aos, const int L[4], unsigned long long seed);
int qkx_gen_point(double *psi_aos, const int L[4], const int x0[4], int spin, int colour);
int qkx_gen_site(int x0[4], const int L[4], unsigned long long seed);

#ifdef __cplusplus
}
#endif
#endif /* QKX_HIP_H */
