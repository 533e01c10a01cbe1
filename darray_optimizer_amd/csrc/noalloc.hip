// Allocation-free variant of the fused FP64 K*Omega pass, for callers that
// cannot allocate device memory (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc).
//
// The only scratch of dh_rbf_project_f64 is the vector of squared norms of
// Xi and Xj (m + n doubles).  Here it lives in a static device array of
// DH_NOALLOC_MAX_POINTS doubles (8 MiB; every GPU the library runs on holds
// its own copy until the process ends).  Host code never passes the array's
// name to a kernel (in host code it is only a stand-in for the symbol): its
// device address is looked up once per device with hipGetSymbolAddress when
// dh_create runs with that device current, and that address is what the
// kernels receive.  The norms are fully rewritten by every call, so nothing
// needs clearing; calls that overlap on one GPU (any handles, any streams)
// race on the array.
#include <mutex>

#include "dh_common.h"

int dh_project_f64_impl(dh_handle_t h, int m, int n, int d, int l, const double* Xi, int ldxi,
                        const double* Xj, int ldxj, const double* ni, const double* nj,
                        const double* yi, const double* yj, double gamma, const double* O,
                        int ldo, double beta, double* Y, int ldy, bool pack);

static __device__ __align__(256) double scratch[DH_NOALLOC_MAX_POINTS];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};  // device address of `scratch`, per device

double* scratch_of(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return (device >= 0 && device < kMaxDevices) ? g_scratch_dev[device] : nullptr;
}

}  // namespace

int dh_noalloc_setup(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= kMaxDevices)
        return dh_fail(DH_ERR_UNSUPPORTED, "dh_create: device %d beyond the %d supported", device,
                       kMaxDevices);
    if (g_scratch_dev[device]) return 0;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return dh_fail(DH_ERR_HIP, "dh_create: hipGetSymbolAddress(scratch) on device %d: %s",
                       device, hipGetErrorString(e));
    g_scratch_dev[device] = static_cast<double*>(p);
    return 0;
}

extern "C" int dh_rbf_project_f64_noalloc(dh_handle_t h, int m, int n, int d, int l,
                                          const double* Xi, int ldxi, const double* Xj, int ldxj,
                                          const double* yi, const double* yj, double gamma,
                                          const double* O, int ldo, double beta, double* Y,
                                          int ldy) {
    const char* name = "dh_rbf_project_f64_noalloc";
    if (!h) return dh_fail(-1, "%s: NULL handle", name);
    if (m < 0) return dh_fail(-2, "%s: m < 0", name);
    if (n < 0) return dh_fail(-3, "%s: n < 0", name);
    if (d < 0) return dh_fail(-4, "%s: d < 0", name);
    if (l < 0) return dh_fail(-5, "%s: l < 0", name);
    if (ldxi < max(1, d)) return dh_fail(-7, "%s: ldxi < d", name);
    if (ldxj < max(1, d)) return dh_fail(-9, "%s: ldxj < d", name);
    if (ldo < max(1, n)) return dh_fail(-14, "%s: ldo < n", name);
    if (ldy < max(1, m)) return dh_fail(-17, "%s: ldy < m", name);
    if (d > 256)
        return dh_fail(DH_ERR_UNSUPPORTED, "%s: d = %d > 256 needs a K block workspace", name, d);
    if ((long)m + n > DH_NOALLOC_MAX_POINTS)
        return dh_fail(DH_ERR_UNSUPPORTED,
                       "%s: m + n = %ld exceeds the %d points of the static norm buffer", name,
                       (long)m + n, DH_NOALLOC_MAX_POINTS);
    if (m == 0 || l == 0) return 0;
    double* nrm = scratch_of(h->device);
    if (!nrm)
        return dh_fail(DH_ERR_HIP, "%s: scratch address unknown on device %d", name, h->device);
    DH_CHECK_STATUS(dh_sqnorms_f64(h, m, d, Xi, ldxi, nrm));
    DH_CHECK_STATUS(dh_sqnorms_f64(h, n, d, Xj, ldxj, nrm + m));
    // The kernel-timing hooks may create events: off for this launch (a handle
    // is used by one host thread at a time).
    void* timing = h->timing;
    h->timing = nullptr;
    const int st = dh_project_f64_impl(h, m, n, d, l, Xi, ldxi, Xj, ldxj, nrm, nrm + m, yi, yj,
                                       gamma, O, ldo, beta, Y, ldy, false);
    h->timing = timing;
    return st;
}
