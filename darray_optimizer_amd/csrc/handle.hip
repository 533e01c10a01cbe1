// Handle, stream and workspace management; error reporting.
//
// Replaces cuda_init / cuda_finalize (dkernel.cu:28-39), which created process
// globals (one cuBLAS handle, one cuSOLVER handle, one stream) and launched
// the custom kernels on the legacy default stream.  Here every call runs on
// the handle's one stream, and device workspaces are allocated once and
// grown on demand instead of the per-call cudaMallocManaged / cudaFree of
// rbf_kernel and tc_gemm (dkernel.cu:135-147, 196-198).
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <new>

#include <utility>
#include <vector>

#include "dh_common.h"

static thread_local char g_last_error[1024] = "";

int dh_fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
    return code;
}

extern "C" const char* dh_last_error_string(void) { return g_last_error; }

void* dh_ws(dh_handle_t h, dh_slot slot, size_t bytes) {
    if (bytes == 0) bytes = 256;
    if (h->ws_size[slot] >= bytes) return h->ws[slot];
    // Grow: the old buffer may still be read by queued work.
    if (h->ws[slot]) {
        if (hipStreamSynchronize(h->stream) != hipSuccess || hipFree(h->ws[slot]) != hipSuccess) {
            dh_fail(DH_ERR_HIP, "workspace release failed");
            return nullptr;
        }
        h->ws[slot] = nullptr;
        h->ws_size[slot] = 0;
    }
    size_t want = bytes + bytes / 8;  // headroom for slightly larger later calls
    want = (want + 4095) & ~size_t(4095);
    void* p = nullptr;
    if (hipMalloc(&p, want) != hipSuccess) {
        (void)hipGetLastError();
        dh_fail(DH_ERR_ALLOC, "hipMalloc of %zu bytes (workspace slot %d) failed", want, (int)slot);
        return nullptr;
    }
    h->ws[slot] = p;
    h->ws_size[slot] = want;
    return p;
}

extern "C" int dh_create(int device, void* stream, dh_handle_t* out) {
    if (!out) return dh_fail(-3, "dh_create: out is NULL");
    *out = nullptr;
    int ndev = 0;
    DH_CHECK_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev)
        return dh_fail(-1, "dh_create: device %d out of range (%d devices)", device, ndev);
    DH_CHECK_HIP(hipSetDevice(device));
    DH_CHECK_STATUS(dh_noalloc_setup(device));
    dh_handle_s* h = new (std::nothrow) dh_handle_s();
    if (!h) return dh_fail(DH_ERR_ALLOC, "dh_create: out of host memory");
    h->device = device;
    if (stream) {
        h->stream = (hipStream_t)stream;
        h->own_stream = false;
    } else {
        hipError_t e = hipStreamCreate(&h->stream);  // default priority
        if (e != hipSuccess) {
            delete h;
            return dh_fail(DH_ERR_HIP, "hipStreamCreate failed: %s", hipGetErrorString(e));
        }
        h->own_stream = true;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) h->num_cu = prop.multiProcessorCount;
    if (h->num_cu <= 0) h->num_cu = 256;
    *out = h;
    return 0;
}

extern "C" int dh_destroy(dh_handle_t h) {
    if (!h) return 0;
    (void)hipSetDevice(h->device);
    (void)hipStreamSynchronize(h->stream);
    dh_batch_release(h);
    dh_tune_release(h);
    dh_timing_enable(h, 0);
    if (h->aux) (void)hipStreamSynchronize(h->aux);
    for (int s = 0; s < DH_WS_COUNT; ++s)
        if (h->ws[s]) (void)hipFree(h->ws[s]);
    if (h->ev_panel) (void)hipEventDestroy(h->ev_panel);
    if (h->ev_update) (void)hipEventDestroy(h->ev_update);
    if (h->ev_orth) (void)hipEventDestroy(h->ev_orth);
    if (h->aux) (void)hipStreamDestroy(h->aux);
    if (h->host_status) (void)hipHostFree(h->host_status);
    delete static_cast<std::map<std::pair<const void*, int>, int>*>(h->occupancy);
    if (h->own_stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return 0;
}

int* dh_host_status(dh_handle_t h) {
    if (!h->host_status &&
        hipHostMalloc((void**)&h->host_status, 16 * sizeof(int), hipHostMallocDefault) != hipSuccess)
        h->host_status = nullptr;
    return h->host_status;
}

int dh_occupancy(dh_handle_t h, const void* kernel, int threads, int* per_cu) {
    using Table = std::map<std::pair<const void*, int>, int>;
    if (!h->occupancy) h->occupancy = new Table();
    Table& t = *static_cast<Table*>(h->occupancy);
    auto it = t.find({kernel, threads});
    if (it != t.end()) {
        *per_cu = it->second;
        return 0;
    }
    int v = 0;
    DH_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, kernel, threads, 0));
    v = v < 1 ? 1 : v;
    t[{kernel, threads}] = v;
    *per_cu = v;
    return 0;
}

int dh_aux_stream(dh_handle_t h) {
    if (h->aux) return 0;
    // Default priority, like the handle's own stream.
    DH_CHECK_HIP(hipStreamCreateWithFlags(&h->aux, hipStreamNonBlocking));
    DH_CHECK_HIP(hipEventCreateWithFlags(&h->ev_panel, hipEventDisableTiming));
    DH_CHECK_HIP(hipEventCreateWithFlags(&h->ev_update, hipEventDisableTiming));
    return 0;
}

extern "C" int dh_synchronize(dh_handle_t h) {
    if (!h) return dh_fail(-1, "dh_synchronize: NULL handle");
    DH_CHECK_HIP(hipStreamSynchronize(h->stream));
    return 0;
}

extern "C" size_t dh_workspace_bytes(dh_handle_t h) {
    if (!h) return 0;
    size_t s = 0;
    for (int i = 0; i < DH_WS_COUNT; ++i) s += h->ws_size[i];
    return s;
}

// ------------------------------------------------------ device memory helpers
// For callers that include no HIP header (the DArray C++ mirror): device
// allocations on the handle's device and 2-D copies / conversions on its stream.

extern "C" int dh_malloc(dh_handle_t h, size_t bytes, void** ptr) {
    if (!h) return dh_fail(-1, "dh_malloc: NULL handle");
    if (!ptr) return dh_fail(-3, "dh_malloc: NULL output pointer");
    *ptr = nullptr;
    if (bytes == 0) return 0;
    DH_CHECK_HIP(hipSetDevice(h->device));
    if (hipMalloc(ptr, bytes) != hipSuccess) {
        *ptr = nullptr;
        (void)hipGetLastError();
        return dh_fail(DH_ERR_ALLOC, "dh_malloc: %zu bytes", bytes);
    }
    return 0;
}

extern "C" int dh_free(dh_handle_t h, void* ptr) {
    if (!h) return dh_fail(-1, "dh_free: NULL handle");
    if (!ptr) return 0;
    DH_CHECK_HIP(hipSetDevice(h->device));
    DH_CHECK_HIP(hipStreamSynchronize(h->stream));
    DH_CHECK_HIP(hipFree(ptr));
    return 0;
}

extern "C" int dh_copy_2d(dh_handle_t h, int dir, int m, int n, size_t elem, const void* src,
                          int lds, void* dst, int ldd) {
    const char* name = "dh_copy_2d";
    if (!h) return dh_fail(-1, "%s: NULL handle", name);
    if (dir != DH_COPY_H2D && dir != DH_COPY_D2H && dir != DH_COPY_D2D)
        return dh_fail(-2, "%s: bad direction %d", name, dir);
    if (m < 0) return dh_fail(-3, "%s: m < 0", name);
    if (n < 0) return dh_fail(-4, "%s: n < 0", name);
    if (elem == 0) return dh_fail(-5, "%s: elem = 0", name);
    if (lds < max(1, m)) return dh_fail(-7, "%s: lds < m", name);
    if (ldd < max(1, m)) return dh_fail(-9, "%s: ldd < m", name);
    if (m == 0 || n == 0) return 0;
    hipMemcpyKind kind = dir == DH_COPY_H2D   ? hipMemcpyHostToDevice
                         : dir == DH_COPY_D2H ? hipMemcpyDeviceToHost
                                              : hipMemcpyDeviceToDevice;
    DH_CHECK_HIP(hipSetDevice(h->device));
    DH_CHECK_HIP(hipMemcpy2DAsync(dst, elem * (size_t)ldd, src, elem * (size_t)lds,
                                  elem * (size_t)m, n, kind, h->stream));
    DH_CHECK_HIP(hipStreamSynchronize(h->stream));
    return 0;
}

namespace {
template <typename S, typename D>
__global__ __launch_bounds__(256) void convert_kernel(int m, int n, const S* __restrict__ A,
                                                      int lda, D* __restrict__ B, int ldb) {
    long total = (long)m * n;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
        int i = (int)(e % m), j = (int)(e / m);
        B[i + (size_t)j * ldb] = (D)A[i + (size_t)j * lda];
    }
}

template <typename S, typename D>
int convert(dh_handle_t h, const char* name, int m, int n, const S* A, int lda, D* B, int ldb) {
    if (!h) return dh_fail(-1, "%s: NULL handle", name);
    if (m < 0) return dh_fail(-2, "%s: m < 0", name);
    if (n < 0) return dh_fail(-3, "%s: n < 0", name);
    if (lda < max(1, m)) return dh_fail(-5, "%s: lda < m", name);
    if (ldb < max(1, m)) return dh_fail(-7, "%s: ldb < m", name);
    if (m == 0 || n == 0) return 0;
    long total = (long)m * n;
    long want = (total + 255) / 256, cap = 8L * h->num_cu;
    int blocks = (int)(want < cap ? want : cap);
    convert_kernel<S, D><<<blocks, 256, 0, h->stream>>>(m, n, A, lda, B, ldb);
    DH_CHECK_LAUNCH();
    return 0;
}
}  // namespace

extern "C" int dh_convert_f32_f64(dh_handle_t h, int m, int n, const float* A, int lda, double* B,
                                  int ldb) {
    return convert(h, "dh_convert_f32_f64", m, n, A, lda, B, ldb);
}

extern "C" int dh_convert_f64_f32(dh_handle_t h, int m, int n, const double* A, int lda, float* B,
                                  int ldb) {
    return convert(h, "dh_convert_f64_f32", m, n, A, lda, B, ldb);
}

// ------------------------------------------------------------ kernel timing

namespace {
struct Timing {
    std::vector<hipEvent_t> pool;                         // reusable events
    std::vector<std::pair<hipEvent_t, hipEvent_t>> open;  // recorded pairs
    hipEvent_t pending = nullptr;                         // begin without end yet
    hipEvent_t take() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
};
}  // namespace

void dh_timing_begin(dh_handle_t h) {
    Timing* t = static_cast<Timing*>(h->timing);
    if (!t) return;
    t->pending = t->take();
    if (t->pending) (void)hipEventRecord(t->pending, h->stream);
}

void dh_timing_end(dh_handle_t h) {
    Timing* t = static_cast<Timing*>(h->timing);
    if (!t || !t->pending) return;
    hipEvent_t e = t->take();
    if (!e) return;
    (void)hipEventRecord(e, h->stream);
    t->open.emplace_back(t->pending, e);
    t->pending = nullptr;
}

extern "C" int dh_timing_enable(dh_handle_t h, int on) {
    if (!h) return dh_fail(-1, "dh_timing_enable: NULL handle");
    Timing* t = static_cast<Timing*>(h->timing);
    if (on && !t) h->timing = new Timing();
    if (!on && t) {
        (void)hipStreamSynchronize(h->stream);
        for (auto& pr : t->open) {
            (void)hipEventDestroy(pr.first);
            (void)hipEventDestroy(pr.second);
        }
        for (hipEvent_t e : t->pool) (void)hipEventDestroy(e);
        if (t->pending) (void)hipEventDestroy(t->pending);
        delete t;
        h->timing = nullptr;
    }
    return 0;
}

extern "C" int dh_timing_read(dh_handle_t h, double* ms, int* launches) {
    if (!h) return dh_fail(-1, "dh_timing_read: NULL handle");
    if (!ms || !launches) return dh_fail(-2, "dh_timing_read: NULL output");
    *ms = 0.0;
    *launches = 0;
    Timing* t = static_cast<Timing*>(h->timing);
    if (!t) return 0;
    DH_CHECK_HIP(hipStreamSynchronize(h->stream));
    for (auto& pr : t->open) {
        float v = 0.f;
        DH_CHECK_HIP(hipEventElapsedTime(&v, pr.first, pr.second));
        *ms += v;
        ++*launches;
        t->pool.push_back(pr.first);
        t->pool.push_back(pr.second);
    }
    t->open.clear();
    return 0;
}
