// qkx_noalloc.hip -- qkx_tm_xx_noalloc: the even-odd Schur operator (and its
// normal form) for callers that cannot allocate device memory: code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc.
//
// qkx_tm_xx takes its intermediate fields (the other parity's A^-1 H in, and
// for the normal form the first Schur application) from the library's field
// pool, which allocates on first use.  This variant keeps them in `scratch`,
// a file-scope device array of QKX_NOALLOC_SCRATCH_BYTES that every GPU the
// library runs on holds until the process ends.  Nothing in it has to start
// at zero: each intermediate is written in full before it is read.  Calls that
// overlap on one GPU share the array and race.
#include "qkx_internal.h"
#include "qkx_core.h"

// In host code `scratch` is a host-side stand-in, not a device address; the
// device address comes from hipGetSymbolAddress, once per device
// (g_scratch_dev), and only that address is handed to kernels.
static __device__ __align__(256) unsigned char scratch[QKX_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
unsigned char *g_scratch_dev[kMaxDevices] = {};

constexpr size_t kAlign = 256;
size_t align_up(size_t n) { return (n + kAlign - 1) / kAlign * kAlign; }

// Bytes a parity-subset field of `entries` (site, RHS) entries takes in the
// scratch: the chunk arrays, then (FP16) the per-entry scales.
size_t field_bytes(int prec, size_t entries)
{
    size_t b = align_up(qkx::spinor_elem_bytes(prec) * qkx::spinor_nchunks(prec) * entries);
    if (prec == QKX_PREC_HALF) b += align_up(sizeof(float) * entries);
    return b;
}

// A parity-subset field laid out like qkx_spinor_create_batch's, at `base`.
qkx_spinor_s field_at(unsigned char *base, const qkx_spinor_s *like, int subset)
{
    qkx_spinor_s f{};
    for (int m = 0; m < 4; m++) f.L[m] = like->L[m];
    f.V = like->V;
    f.Vh = like->Vh;
    f.prec = like->prec;
    f.subset = subset;
    f.nsites = like->Vh;
    f.nrhs = like->nrhs;
    const size_t entries = (size_t)f.nsites * f.nrhs;
    f.data = base;
    f.nrm = nullptr;
    if (f.prec == QKX_PREC_HALF)
        f.nrm = reinterpret_cast<float *>(
            base + align_up(qkx::spinor_elem_bytes(f.prec) * qkx::spinor_nchunks(f.prec) * entries));
    f.bytes = qkx::spinor_elem_bytes(f.prec) * qkx::spinor_nchunks(f.prec) * entries;
    return f;
}

} // namespace

namespace qkx {

int noalloc_setup(int device)
{
    if (device < 0 || device >= kMaxDevices) {
        set_error("qkx_init: device index beyond the no-allocation table");
        return QKX_ERR_ARG;
    }
    if (g_scratch_dev[device]) return QKX_SUCCESS;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        set_error(std::string("qkx_init: device address of the no-allocation scratch: ") + hipGetErrorString(e));
        return QKX_ERR_HIP;
    }
    g_scratch_dev[device] = static_cast<unsigned char *>(p);
    return QKX_SUCCESS;
}

} // namespace qkx

extern "C" {

int qkx_tm_xx_noalloc(qkx_spinor *out, const qkx_spinor *in, const qkx_gauge *U, const qkx_params *p, int basis,
                      int parity, int normal)
{
    const int dev = qkx::g_device;
    QKX_CHECK_ARG(dev >= 0 && dev < kMaxDevices && g_scratch_dev[dev],
                  "qkx_tm_xx_noalloc: call qkx_init first (it takes the scratch address)");
    QKX_CHECK_ARG(!qkx::prof_active(), "qkx_tm_xx_noalloc: per-kernel instrumentation (qkx_prof_enable) creates "
                                       "events; switch it off");
    int pe = 0;
    if (int rc = qkx::tm_xx_check(out, in, U, p, basis, parity, &pe)) return rc;
    const size_t entries = (size_t)in->Vh * in->nrhs;
    const size_t one = field_bytes(in->prec, entries);
    const size_t need = normal ? 2 * one : one;
    if (need > (size_t)QKX_NOALLOC_SCRATCH_BYTES) {
        qkx::set_error("qkx_tm_xx_noalloc: the intermediate fields need " + std::to_string(need) +
                       " bytes, more than the " + std::to_string((size_t)QKX_NOALLOC_SCRATCH_BYTES) +
                       "-byte scratch buffer");
        return QKX_ERR_ARG;
    }
    unsigned char *base = g_scratch_dev[dev];
    qkx_spinor_s t = field_at(base, in, pe == 0 ? QKX_SUBSET_ODD : QKX_SUBSET_EVEN);
    qkx_spinor_s y = field_at(base + one, in, pe == 0 ? QKX_SUBSET_EVEN : QKX_SUBSET_ODD);
    // the hop autotuner would allocate scratch fields for a kind it has not
    // timed yet: such launches keep the cached or default shape here
    qkx::set_no_tune(1);
    const int rc = qkx::tm_xx_apply(out, in, U, p, basis, pe, normal, &t, normal ? &y : nullptr);
    qkx::set_no_tune(0);
    return rc;
}

} // extern "C"
