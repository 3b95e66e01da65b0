// qkx_core.cpp -- library state: device selection, the library stream, error
// reporting, timers, hopping-kernel instrumentation, reduction scratch and a
// pool of temporary fields.
#include "qkx_internal.h"
#include "qkx_core.h"
#include <algorithm>
#include <cstring>
#include <string>
#include <cstdio>
#include <mutex>
#include <vector>

namespace qkx {

extern int cg_batch_size;
hipStream_t g_stream = nullptr;
int g_device = -1;
int g_cu_count = 0;

namespace {
std::string g_err;
bool g_init = false;
double *g_partial = nullptr;     // reduction partials
unsigned int *g_counter = nullptr;
double *g_result = nullptr;      // result slots
hipEvent_t g_t0 = nullptr, g_t1 = nullptr;

// instrumentation
bool g_prof = false;
struct ProfRec {
    hipEvent_t a, b;
    double bytes;
    int prec;
};
std::vector<ProfRec> g_prof_recs;
size_t g_prof_used = 0;

struct PoolEntry {
    qkx_spinor_s *f;
    bool busy;
};
std::vector<PoolEntry> g_pool;
} // namespace

void set_error(const std::string &msg) { g_err = msg; }

bool check_hip(hipError_t e, const char *what)
{
    if (e == hipSuccess) return true;
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return false;
}

int ensure_init()
{
    if (g_init) return QKX_SUCCESS;
    return qkx_init(0);
}

double *reduce_scratch() { return g_partial; }
unsigned int *reduce_counter() { return g_counter; }
double *reduce_result() { return g_result; }
double *result_slot(int k) { return g_result + k; }

bool prof_active() { return g_prof; }

void prof_begin()
{
    if (!g_prof) return;
    if (g_prof_used == g_prof_recs.size()) {
        ProfRec r;
        (void)hipEventCreate(&r.a);
        (void)hipEventCreate(&r.b);
        r.bytes = 0;
        g_prof_recs.push_back(r);
    }
    (void)hipEventRecord(g_prof_recs[g_prof_used].a, g_stream);
}

void prof_end(double bytes, int prec)
{
    if (!g_prof) return;
    (void)hipEventRecord(g_prof_recs[g_prof_used].b, g_stream);
    g_prof_recs[g_prof_used].bytes = bytes;
    g_prof_recs[g_prof_used].prec = prec;
    g_prof_used++;
}

qkx_spinor_s *pool_get_batch(const int L[4], int prec, int subset, int nrhs)
{
    for (auto &e : g_pool) {
        if (!e.busy && e.f->prec == prec && e.f->subset == subset && e.f->nrhs == nrhs && e.f->L[0] == L[0] &&
            e.f->L[1] == L[1] && e.f->L[2] == L[2] && e.f->L[3] == L[3]) {
            e.busy = true;
            return e.f;
        }
    }
    qkx_spinor_s *f = nullptr;
    if (qkx_spinor_create_batch(&f, L, prec, subset, nrhs) != QKX_SUCCESS) return nullptr;
    g_pool.push_back({f, true});
    return f;
}

qkx_spinor_s *pool_get(const int L[4], int prec, int subset) { return pool_get_batch(L, prec, subset, 1); }

void pool_release(qkx_spinor_s *f)
{
    for (size_t i = 0; i < g_pool.size(); i++)
        if (g_pool[i].f == f) {
            qkx_spinor_destroy(f);
            g_pool.erase(g_pool.begin() + i);
            return;
        }
}

void pool_put(qkx_spinor_s *f)
{
    for (auto &e : g_pool)
        if (e.f == f) e.busy = false;
}

void pool_clear()
{
    for (auto &e : g_pool) qkx_spinor_destroy(e.f);
    g_pool.clear();
}

} // namespace qkx

extern "C" {

int qkx_init(int device)
{
    using namespace qkx;
    if (g_init) {
        if (device == g_device) return QKX_SUCCESS;
        set_error("qkx_init: already initialised on another device");
        return QKX_ERR_ARG;
    }
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        set_error("qkx_init: no HIP device available");
        return QKX_ERR_NODEVICE;
    }
    if (device < 0 || device >= n) {
        set_error("qkx_init: device index out of range");
        return QKX_ERR_ARG;
    }
    QKX_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    QKX_HIP(hipGetDeviceProperties(&prop, device));
    g_cu_count = prop.multiProcessorCount;
    QKX_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    QKX_HIP(hipMalloc(&g_partial, sizeof(double) * 8 * 4096));
    QKX_HIP(hipMalloc(&g_counter, sizeof(unsigned int) * 16));
    QKX_HIP(hipMalloc(&g_result, sizeof(double) * 64));
    QKX_HIP(hipMemset(g_counter, 0, sizeof(unsigned int) * 16));
    QKX_HIP(hipMemset(g_result, 0, sizeof(double) * 64));
    QKX_HIP(hipEventCreate(&g_t0));
    QKX_HIP(hipEventCreate(&g_t1));
    QKX_HIP(hipDeviceSynchronize());
    if (int rc = noalloc_setup(device)) return rc;
    g_device = device;
    g_init = true;
    return QKX_SUCCESS;
}

int qkx_finalize(void)
{
    using namespace qkx;
    if (!g_init) return QKX_SUCCESS;
    (void)hipStreamSynchronize(g_stream);
    pool_clear();
    smear_release();
    for (auto &r : g_prof_recs) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    g_prof_recs.clear();
    g_prof_used = 0;
    (void)hipFree(g_partial);
    (void)hipFree(g_counter);
    (void)hipFree(g_result);
    (void)hipEventDestroy(g_t0);
    (void)hipEventDestroy(g_t1);
    (void)hipStreamDestroy(g_stream);
    g_stream = nullptr;
    g_init = false;
    g_device = -1;
    return QKX_SUCCESS;
}

const char *qkx_last_error(void) { return qkx::g_err.c_str(); }

int qkx_device_count(int *n)
{
    QKX_CHECK_ARG(n, "qkx_device_count: null argument");
    *n = 0;
    if (hipGetDeviceCount(n) != hipSuccess) *n = 0;
    return QKX_SUCCESS;
}

int qkx_mem_info(double *used_bytes, double *total_bytes)
{
    QKX_CHECK_ARG(used_bytes && total_bytes, "qkx_mem_info: null argument");
    size_t fr = 0, tot = 0;
    QKX_HIP(hipMemGetInfo(&fr, &tot));
    *used_bytes = (double)(tot - fr);
    *total_bytes = (double)tot;
    return QKX_SUCCESS;
}

int qkx_sync(void)
{
    int rc = qkx::ensure_init();
    if (rc) return rc;
    QKX_HIP(hipStreamSynchronize(qkx::g_stream));
    return QKX_SUCCESS;
}

int qkx_timer_start(void)
{
    int rc = qkx::ensure_init();
    if (rc) return rc;
    QKX_HIP(hipEventRecord(qkx::g_t0, qkx::g_stream));
    return QKX_SUCCESS;
}

int qkx_timer_stop(double *ms)
{
    QKX_CHECK_ARG(ms, "qkx_timer_stop: null argument");
    QKX_HIP(hipEventRecord(qkx::g_t1, qkx::g_stream));
    QKX_HIP(hipEventSynchronize(qkx::g_t1));
    float f = 0;
    QKX_HIP(hipEventElapsedTime(&f, qkx::g_t0, qkx::g_t1));
    *ms = f;
    return QKX_SUCCESS;
}

int qkx_prof_enable(int on)
{
    int rc = qkx::ensure_init();
    if (rc) return rc;
    QKX_HIP(hipStreamSynchronize(qkx::g_stream));
    qkx::g_prof = on != 0;
    qkx::g_prof_used = 0;
    return QKX_SUCCESS;
}

int qkx_prof_read(int *launches, double *ms, double *bytes)
{
    return qkx_prof_read_prec(-1, launches, ms, bytes);
}

int qkx_prof_read_prec(int prec, int *launches, double *ms, double *bytes)
{
    QKX_CHECK_ARG(launches && ms && bytes, "qkx_prof_read: null argument");
    QKX_HIP(hipStreamSynchronize(qkx::g_stream));
    double tot = 0, by = 0;
    int n = 0;
    for (size_t i = 0; i < qkx::g_prof_used; i++) {
        if (prec >= 0 && qkx::g_prof_recs[i].prec != prec) continue;
        float f = 0;
        QKX_HIP(hipEventElapsedTime(&f, qkx::g_prof_recs[i].a, qkx::g_prof_recs[i].b));
        tot += f;
        by += qkx::g_prof_recs[i].bytes;
        n++;
    }
    *launches = n;
    *ms = tot;
    *bytes = by;
    return QKX_SUCCESS;
}

int qkx_prof_reset(void)
{
    QKX_HIP(hipStreamSynchronize(qkx::g_stream));
    qkx::g_prof_used = 0;
    return QKX_SUCCESS;
}

int qkx_set_rhs_tile(int rt)
{
    QKX_CHECK_ARG(rt >= 0 && rt <= 4, "qkx_set_rhs_tile: tile must be 0 (automatic) or 1..4");
    qkx::set_rhs_tile(rt);
    return QKX_SUCCESS;
}

int qkx_set_cg_batch(int iterations)
{
    QKX_CHECK_ARG(iterations >= 1 && iterations <= 1024, "qkx_set_cg_batch: 1..1024");
    qkx::cg_batch_size = iterations;
    return QKX_SUCCESS;
}

int qkx_set_batch_link_sharing(int on)
{
    QKX_CHECK_ARG(on >= 0 && on <= 2, "qkx_set_batch_link_sharing: 0, 1 or 2");
    qkx::set_batch_link_sharing(on);
    return QKX_SUCCESS;
}

int qkx_set_autotune(int mode)
{
    QKX_CHECK_ARG(mode >= 0 && mode <= 2, "qkx_set_autotune: 0, 1 or 2");
    qkx::set_autotune(mode);
    return QKX_SUCCESS;
}

int qkx_autotune_report(char *buf, int len, int *needed)
{
    const std::string r = qkx::tune_report();
    if (needed) *needed = (int)r.size() + 1;
    if (buf && len > 0) {
        const size_t n = std::min((size_t)len - 1, r.size());
        std::memcpy(buf, r.data(), n);
        buf[n] = 0;
    }
    return QKX_SUCCESS;
}

int qkx_set_site_order(int blocked)
{
    QKX_CHECK_ARG(blocked >= 0 && blocked <= 2, "qkx_set_site_order: 0, 1 or 2");
    qkx::set_site_order(blocked);
    return QKX_SUCCESS;
}

int qkx_set_neighbour_staging(int on)
{
    QKX_CHECK_ARG(on >= 0 && on <= 2, "qkx_set_neighbour_staging: 0, 1 or 2");
    qkx::set_neighbour_staging(on);
    return QKX_SUCCESS;
}

int qkx_set_tuning(int xcd_remap, int block, int depth64, int depth32)
{
    QKX_CHECK_ARG(block == 64 || block == 128 || block == 256, "qkx_set_tuning: block must be 64, 128 or 256");
    QKX_CHECK_ARG((depth64 == 1 || depth64 == 2) && (depth32 == 1 || depth32 == 2),
                  "qkx_set_tuning: pipeline depth must be 1 or 2");
    qkx::set_hop_tuning(xcd_remap != 0, block, depth64, depth32);
    return QKX_SUCCESS;
}

} // extern "C"
