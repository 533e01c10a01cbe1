// DArray's local/distributed matrix types for one rank on one MI355X.
//
// Drop-in for the parts of DArray's darray.h that the RBF low-rank path uses
// (Grid darray.h:89-200, LMatrix 257-421, DMatrix 424-1151,
// DMatrixDescriptor 1379-1402), restated for a 1 x 1 process grid: a
// DMatrix's local block is the whole matrix (darray.h:490-498 with
// np = nq = 1), column-major with leading dimension lld.
//
// Difference from the reference: a DMatrix block lives in host memory AND in
// a device mirror, and only one side is current at a time.  Host accessors
// (data, at, operator(), local_view, set_*, fnorm, print, collect ...) bring
// the host copy up to date; the algorithms of dkernel.h, blas_like.h and
// lapack_like.h run on the device copy.  Copies move only when the other
// side changed, so a chain of device algorithms never touches PCIe.  Host
// pointers obtained before a device algorithm ran are stale after it; take
// them again.
#pragma once

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <random>
#include <string>
#include <type_traits>
#include <vector>

#include "../darray_hip.h"

namespace DArray {

enum Layout { RowMajor, ColMajor };

// ----------------------------------------------------------------- runtime

// Status check with the reference's device-error convention (message, then
// exit; dkernel.cu:20-26).
inline void check(int status, const char* where) {
    if (status != 0) {
        std::fprintf(stderr, "DArray: %s: %s (status %d)\n", where, dh_last_error_string(),
                     status);
        std::exit(status < 0 ? 1 : status);
    }
}

// Process-wide handle (device 0, own stream), created on first use.
inline dh_handle_t& handle_slot() {
    static dh_handle_t h = nullptr;
    return h;
}

inline dh_handle_t handle() {
    dh_handle_t& h = handle_slot();
    if (!h) check(dh_create(0, nullptr, &h), "dh_create");
    return h;
}

// Releases the process-wide handle (optional; the OS reclaims it at exit).
inline void release_handle() {
    dh_handle_t& h = handle_slot();
    if (h) dh_destroy(h);
    h = nullptr;
}

// Device buffer owned by the caller's scope.
template <typename T>
class DeviceBuffer {
    T* m_ptr = nullptr;
    size_t m_count = 0;

public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t count) : m_count(count) {
        void* p = nullptr;
        check(dh_malloc(handle(), sizeof(T) * count, &p), "dh_malloc");
        m_ptr = static_cast<T*>(p);
    }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer() {
        if (m_ptr) dh_free(handle(), m_ptr);
    }
    T* get() const { return m_ptr; }
    size_t size() const { return m_count; }
};

// ------------------------------------------------------------------- Grid

// A 1 x 1 process grid.  The constructor accepts DArray's (comm, p, q)
// arguments so callers compile unchanged; any p * q other than 1 is refused
// (the multi-rank layout needs the distributed movement of darray.h:897-1151,
// which this build does not provide).
class Grid {
public:
    Grid() = default;
    template <typename Comm>
    Grid(const Comm&, int p, int q) {
        if (p != 1 || q != 1) {
            std::fprintf(stderr, "DArray: this build runs one rank on one GPU (asked for %dx%d)\n",
                         p, q);
            std::exit(1);
        }
    }
    int rank() const { return 0; }
    int np() const { return 1; }
    std::array<int, 2> dims() const { return {1, 1}; }
    std::array<int, 2> ranks() const { return {0, 0}; }
    Layout layout() const { return ColMajor; }
    int convert_to_linear(int, int) const { return 0; }
    std::array<int, 2> convert_to_pq(int) const { return {0, 0}; }
    void barrier() const {}
    void barrier_row() const {}
    void barrier_col() const {}
    // Same factorisation as darray.h:163-172: the largest k <= round(sqrt(np))
    // dividing np gives p = k, q = np / k.
    static void np_to_pq(int np, int& p, int& q) {
        for (int k = (int)std::lround(std::sqrt((double)np)); k > 0; --k)
            if (np % k == 0) {
                p = k;
                q = np / k;
                return;
            }
    }
    void for_each_process_in_column(const std::function<void(void)>& f) const { f(); }
    void for_each_process_in_row(const std::function<void(void)>& f) const { f(); }
    void for_each_process_in_grid(const std::function<void(void)>& f) const { f(); }
};

// Block-cyclic index helpers (darray.h:65-86, 1118-1124) for one process.
inline int size_of_local_part(int globalsize, int p, int np) { return (globalsize - p + np - 1) / np; }

// ---------------------------------------------------------------- LMatrix

// Host-resident column-major block (darray.h:257-421): owning, or a view of
// somebody else's buffer.
template <typename T>
class LMatrix {
    std::shared_ptr<T[]> m_storage;
    T* m_data = nullptr;
    std::array<int, 2> m_dims{0, 0};
    int m_ld = 0;

public:
    LMatrix() = default;
    LMatrix(int m, int n) : LMatrix(m, n, std::max(m, 1)) {}
    LMatrix(int m, int n, int ld)
        : m_storage(new T[(size_t)std::max(ld, 1) * std::max(n, 1)]), m_dims{m, n}, m_ld(ld) {
        m_data = m_storage.get();
    }
    LMatrix(int m, int n, int ld, T* buffer) : m_data(buffer), m_dims{m, n}, m_ld(ld) {}

    T& operator[](int i) { return m_data[i]; }
    T operator[](int i) const { return m_data[i]; }
    T& operator()(int i, int j) { return m_data[i + (size_t)j * m_ld]; }
    T operator()(int i, int j) const { return m_data[i + (size_t)j * m_ld]; }
    int& ld() { return m_ld; }
    int ld() const { return m_ld; }
    void set_ld(int ld) { m_ld = ld; }
    std::array<int, 2>& dims() { return m_dims; }
    T* data() { return m_data; }
    const T* data() const { return m_data; }

    void set_constant(T a) { set_function([a](int, int) { return a; }); }
    void set_zeros() { set_constant(T(0)); }
    void set_normal(T mean, T stddev, unsigned seed = std::random_device{}()) {
        std::mt19937 gen(seed);
        std::normal_distribution<T> dist(mean, stddev);
        for (int i = 0; i < m_dims[0]; ++i)
            for (int j = 0; j < m_dims[1]; ++j) (*this)(i, j) = dist(gen);
    }
    void set_function(const std::function<T(int, int)>& f) {
        for (int i = 0; i < m_dims[0]; ++i)
            for (int j = 0; j < m_dims[1]; ++j) (*this)(i, j) = f(i, j);
    }
    LMatrix<T> transpose() const {
        LMatrix<T> t(m_dims[1], m_dims[0]);
        for (int j = 0; j < m_dims[1]; ++j)
            for (int i = 0; i < m_dims[0]; ++i) t(j, i) = (*this)(i, j);
        return t;
    }
    void clear_lower() {
        for (int j = 0; j < m_dims[1]; ++j)
            for (int i = j + 1; i < m_dims[0]; ++i) (*this)(i, j) = T(0);
    }
    void clear_upper() {
        for (int j = 0; j < m_dims[1]; ++j)
            for (int i = 0; i < std::min(j, m_dims[0]); ++i) (*this)(i, j) = T(0);
    }
    double fnorm() const {
        double s = 0.0;
        for (int j = 0; j < m_dims[1]; ++j)
            for (int i = 0; i < m_dims[0]; ++i) s += (double)(*this)(i, j) * (*this)(i, j);
        return std::sqrt(s);
    }
    void print(const std::string& title) const {
        std::printf("%s\n", title.c_str());
        for (int i = 0; i < m_dims[0]; ++i) {
            for (int j = 0; j < m_dims[1]; ++j) std::printf("%8.3f ", (double)(*this)(i, j));
            std::printf("\n");
        }
    }
};

// ---------------------------------------------------------------- DMatrix

namespace detail {

// Host block + device mirror with a "which side is current" flag.
template <typename T>
struct Block {
    int rows = 0, cols = 0, ld = 1;
    std::shared_ptr<T[]> host_owned;
    T* host = nullptr;
    T* dev = nullptr;
    bool host_current = true;
    bool dev_current = false;

    Block(int m, int n) : rows(m), cols(n), ld(std::max(m, 1)) {
        host_owned.reset(new T[(size_t)ld * std::max(n, 1)]);
        host = host_owned.get();
    }
    Block(int m, int n, int ldb, T* buffer) : rows(m), cols(n), ld(ldb), host(buffer) {}
    Block(const Block&) = delete;
    Block& operator=(const Block&) = delete;
    ~Block() {
        if (dev) dh_free(handle(), dev);
    }
    void to_host() {
        if (host_current) return;
        check(dh_copy_2d(handle(), DH_COPY_D2H, rows, cols, sizeof(T), dev, ld, host, ld),
              "DMatrix device -> host");
        host_current = true;
    }
    void to_device() {
        if (!dev && (size_t)ld * cols > 0) {
            void* p = nullptr;
            check(dh_malloc(handle(), sizeof(T) * (size_t)ld * cols, &p), "DMatrix device mirror");
            dev = static_cast<T*>(p);
        }
        if (dev_current) return;
        check(dh_copy_2d(handle(), DH_COPY_H2D, rows, cols, sizeof(T), host, ld, dev, ld),
              "DMatrix host -> device");
        dev_current = true;
    }
};

}  // namespace detail

template <typename T>
class DMatrix {
    Grid m_grid;
    std::array<int, 2> m_dims{0, 0};
    std::shared_ptr<detail::Block<T>> m_block;
    T m_unused{};

public:
    DMatrix() = default;
    // Allocating (darray.h:438-452).
    DMatrix(const Grid& grid, int m, int n)
        : m_grid(grid), m_dims{m, n}, m_block(std::make_shared<detail::Block<T>>(m, n)) {}
    // Attached to a caller's host buffer holding A[i1:, j1:] (darray.h:455-467).
    DMatrix(const Grid& grid, int m, int n, T* buf, int ld, int i1, int j1)
        : m_grid(grid), m_dims{m, n},
          m_block(std::make_shared<detail::Block<T>>(m, n, ld, buf - (i1 + (size_t)j1 * ld))) {}

    // Deep copy; the copy's current side is the host.
    DMatrix<T> clone() {
        DMatrix<T> c(m_grid, m_dims[0], m_dims[1]);
        const T* src = host_read();
        for (int j = 0; j < m_dims[1]; ++j)
            std::memcpy(c.m_block->host + (size_t)j * c.lld(), src + (size_t)j * lld(),
                        sizeof(T) * m_dims[0]);
        return c;
    }

    std::array<int, 2> dims() const { return m_dims; }
    std::array<int, 2> ldims() const { return m_dims; }
    Grid& grid() { return m_grid; }
    int lld() const { return m_block ? m_block->ld : 1; }

    // ---- host side
    // Current host block, for reading and writing (the device copy becomes stale).
    T* data() {
        m_block->to_host();
        m_block->dev_current = false;
        return m_block->host;
    }
    // Current host block, read only.
    const T* host_read() {
        m_block->to_host();
        return m_block->host;
    }
    T& at(int i, int j) {
        if (i < 0 || j < 0 || i >= m_dims[0] || j >= m_dims[1]) return m_unused;
        return data()[i + (size_t)j * lld()];
    }
    T operator()(int i, int j) { return host_read()[i + (size_t)j * lld()]; }

    // ---- device side
    // Current device block, for reading and writing (the host copy becomes stale).
    T* device_data() {
        m_block->to_device();
        m_block->host_current = false;
        return m_block->dev;
    }
    // Current device block, read only.
    const T* device_read() {
        m_block->to_device();
        return m_block->dev;
    }
    // Device address of global element (i, j), for reading and writing.
    T* device_at(int i, int j) { return device_data() + i + (size_t)j * lld(); }
    const T* device_read_at(int i, int j) { return device_read() + i + (size_t)j * lld(); }

    // ---- fills (darray.h:513-603), host side.  The unseeded fills draw a
    // seed from std::random_device where the reference uses the clock.
    void set_function(const std::function<T(int, int)>& f) {
        T* a = data();
        for (int i = 0; i < m_dims[0]; ++i)
            for (int j = 0; j < m_dims[1]; ++j) a[i + (size_t)j * lld()] = f(i, j);
    }
    // std::mt19937(seed) + normal_distribution<T>, row-outer / column-inner
    // fill order (darray.h:535-544).
    void set_normal_seed(T mean, T stddev, int seed) {
        std::mt19937 gen((unsigned)seed);
        std::normal_distribution<T> dist(mean, stddev);
        T* a = data();
        for (int i = 0; i < m_dims[0]; ++i)
            for (int j = 0; j < m_dims[1]; ++j) a[i + (size_t)j * lld()] = dist(gen);
    }
    void set_normal(T mean, T stddev) { set_normal_seed(mean, stddev, (int)std::random_device{}()); }
    // std::mt19937(seed) + uniform_real_distribution<T>(a, b), same order.
    void set_uniform_seed(T lo, T hi, int seed) {
        std::mt19937 gen((unsigned)seed);
        std::uniform_real_distribution<T> dist(lo, hi);
        T* a = data();
        for (int i = 0; i < m_dims[0]; ++i)
            for (int j = 0; j < m_dims[1]; ++j) a[i + (size_t)j * lld()] = dist(gen);
    }
    void set_uniform(T lo, T hi) { set_uniform_seed(lo, hi, (int)std::random_device{}()); }
    void set_constant(T v) { set_function([v](int, int) { return v; }); }
    void set_identity() { set_function([](int i, int j) { return i == j ? T(1) : T(0); }); }
    void set_value_from_matrix(const T* x, int ldx) {
        set_function([x, ldx](int i, int j) { return x[i + (size_t)j * ldx]; });
    }

    // Frobenius norm accumulated in double (darray.h:607-617).
    double fnorm() {
        const T* a = host_read();
        double s = 0.0;
        for (int j = 0; j < m_dims[1]; ++j)
            for (int i = 0; i < m_dims[0]; ++i) {
                double v = a[i + (size_t)j * lld()];
                s += v * v;
            }
        return std::sqrt(s);
    }

    // ---- views and copies (darray.h:720-1151 at one rank)
    // Host view of A[i1:i2, j1:j2] (writable: the device copy becomes stale).
    LMatrix<T> local_view(int i1, int j1, int i2, int j2) {
        T* a = data();
        return LMatrix<T>(i2 - i1, j2 - j1, lld(), a + i1 + (size_t)j1 * lld());
    }
    // Host copy of A[i1:i2, j1:j2].
    LMatrix<T> replicate_in_all(int i1, int j1, int i2, int j2) {
        const T* a = host_read();
        LMatrix<T> r(i2 - i1, j2 - j1);
        for (int j = j1; j < j2; ++j)
            for (int i = i1; i < i2; ++i) r(i - i1, j - j1) = a[i + (size_t)j * lld()];
        return r;
    }
    LMatrix<T> replicate_in_rows(int i1, int j1, int i2, int j2) {
        return replicate_in_all(i1, j1, i2, j2);
    }
    LMatrix<T> replicate_in_columns(int i1, int j1, int i2, int j2) {
        return replicate_in_all(i1, j1, i2, j2);
    }
    // Host copy of A[i1:i2, j1:j2]^T.
    LMatrix<T> transpose_and_replicate_in_rows(int i1, int j1, int i2, int j2) {
        return replicate_in_all(i1, j1, i2, j2).transpose();
    }
    // Write a host block back into A[i1:i2, j1:j2].
    void dereplicate_in_all(LMatrix<T>& src, int i1, int j1, int i2, int j2) {
        T* a = data();
        for (int j = j1; j < j2; ++j)
            for (int i = i1; i < i2; ++i) a[i + (size_t)j * lld()] = src(i - i1, j - j1);
    }
    void dereplicate_in_rows(LMatrix<T>& src, int i1, int j1, int i2, int j2) {
        dereplicate_in_all(src, i1, j1, i2, j2);
    }
    void dereplicate_in_columns(LMatrix<T>& src, int i1, int j1, int i2, int j2) {
        dereplicate_in_all(src, i1, j1, i2, j2);
    }
    // The whole matrix on "root" (a copy; darray.h:640-690 at one rank).
    LMatrix<T> collect(int root) {
        (void)root;
        return replicate_in_all(0, 0, m_dims[0], m_dims[1]);
    }
    void collect_and_print(const std::string& title, int root = 0) { collect(root).print(title); }
    void print_by_process(const std::string& title) { collect_and_print(title); }

    // (start, end) of global range ij in the local index space of process p of
    // np (darray.h:1118-1124); the identity at one rank.
    static std::array<int, 2> global_to_local_index(std::array<int, 2> ij, int p, int np, int align) {
        int off = (p - align + np) % np;
        int lo = (ij[0] - off + np - 1) / np;
        if (ij[0] - off < 0) lo = 0;
        int hi = (ij[1] - off + np - 1) / np;
        if (ij[1] - off < 0) hi = 0;
        return {lo, hi};
    }
};

// A sub-block A[i1:i2, j1:j2] of a DMatrix (darray.h:1379-1402).
template <typename T>
struct DMatrixDescriptor {
    DMatrix<T> matrix;
    int i1, j1, i2, j2;
    int m() const { return i2 - i1; }
    int n() const { return j2 - j1; }
    LMatrix<T> replicate_in_rows() { return matrix.replicate_in_rows(i1, j1, i2, j2); }
    LMatrix<T> replicate_in_columns() { return matrix.replicate_in_columns(i1, j1, i2, j2); }
    LMatrix<T> replicate_in_all() { return matrix.replicate_in_all(i1, j1, i2, j2); }
    T* device_data() { return matrix.device_at(i1, j1); }
    const T* device_read() { return matrix.device_read_at(i1, j1); }
    int ld() const { return matrix.lld(); }
};

// ---------------------------------------------------------------- timing

// Wall-clock milliseconds (the reference's ElapsedTimer, utils.h:436-456,
// reads CLOCK_MONOTONIC_COARSE; steady_clock is finer).
class ElapsedTimer {
    std::chrono::steady_clock::time_point m_t0 = std::chrono::steady_clock::now();

public:
    void start() { m_t0 = std::chrono::steady_clock::now(); }
    int elapsed() const {
        return (int)std::chrono::duration_cast<std::chrono::milliseconds>(
                   std::chrono::steady_clock::now() - m_t0)
            .count();
    }
    double elapsed_s() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - m_t0).count();
    }
};

template <typename T>
inline T relative_difference(T a, T b) {
    T d = std::fabs(a - b), s = std::max(std::fabs(a), std::fabs(b));
    return s > T(0) ? d / s : d;
}

}  // namespace DArray
