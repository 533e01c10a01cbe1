// Barrier-method SVM dual solver on the Nystrom factor (testsvm.cpp:191-321,
// dbarrier.h:13-139) with everything of size n resident on the device.
//
// Problem: min f(a) = a^T Z Z^T a / 2 - sum(a) over 0 < a < C, Z (n x k) the
// Nystrom factor of diag(y) K diag(y) (the reference drops the equality
// constraint y^T a = 0 and estimates the bias afterwards).  Barrier
// g(a) = f(a) - sum(log a_i + log(C - a_i)) / t, Newton steps from
// Jacobi-preconditioned CG on (Z Z^T + diag(D)) s = -grad g, backtracking line
// search, t *= 20 until the gap 2n/t is below 1e-5 |f|.
//
// Per CG iteration Z is read twice (Z^T p by row chunks and column groups,
// then Z v by rows): 2 n k 8 bytes, HBM-bound.  All reductions are
// per-workgroup partial sums combined in a fixed order, so results are
// deterministic.  Scalars of the CG recurrence stay on the device, and so
// does the stopping test: the last kernel of an iteration clears an active
// flag that the kernels of later iterations check first, so the host queues
// iterations in batches and reads the flag once per batch (the iterate is the
// one an iteration-by-iteration loop would stop at).  The line-search values
// are read once per Newton step.
//
// Deviations from the reference (bug fixes), as the oracle states them: the
// CG right-hand side is -grad (the reference solves with b = e_1,
// testsvm.cpp:234-235), no n I is added to the operator (dbarrier.h:99,113),
// the preconditioner is diag(Z Z^T) + D (the reference uses diag(Z Z^T)),
// the Newton CG starts from 0 (the reference from a, dbarrier.h:108), and
// every loop has an iteration bound.
#include <algorithm>
#include <cmath>
#include <vector>

#include "dh_common.h"

namespace {

constexpr int kRows = 256;   // rows per workgroup of the row-parallel kernels
constexpr int kColGrp = 32;  // columns per workgroup of Z^T x
constexpr int kTrials = 16;  // line-search step sizes evaluated per launch

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Sum of the 256 threads' values, valid in thread 0 (fixed order).
__device__ __forceinline__ double block_sum(double v, double* sh) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) s = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    __syncthreads();
    return s;
}

// Sum of count values part[0..count) in a fixed order by one workgroup of
// 256 threads; the result is returned to every thread.
__device__ double fixed_sum(int count, const double* __restrict__ part, double* sh) {
    double v = 0.0;
    for (int i = threadIdx.x; i < count; i += 256) v += part[i];
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    const double total = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    __syncthreads();  // sh is reused by the caller
    return total;
}

// part[c * nchunks + chunk] = sum over the chunk's rows of Z[r, c] * x[r].
// Kernels of the CG iteration take `ctl` (the loop's control words, ctl[0] =
// active flag) and do nothing once the loop has stopped; NULL = always run.
__device__ __forceinline__ bool cg_stopped(const double* ctl) { return ctl && ctl[0] == 0.0; }

__global__ __launch_bounds__(256) void ztx_partial_kernel(int n, int k, const double* __restrict__ Z,
                                                          int ldz, const double* __restrict__ x,
                                                          double* __restrict__ part,
                                                          const double* __restrict__ ctl) {
    if (cg_stopped(ctl)) return;
    const int chunk = blockIdx.x, nchunks = gridDim.x;
    const int r0 = chunk * kRows;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double xv[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int r = r0 + lane + 64 * i;
        xv[i] = r < n ? x[r] : 0.0;
    }
    const int c0 = blockIdx.y * kColGrp;
    for (int cc = wave; cc < kColGrp; cc += 4) {
        const int c = c0 + cc;
        if (c >= k) break;
        const double* z = Z + (size_t)c * ldz;
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int r = r0 + lane + 64 * i;
            if (r < n) s += z[r] * xv[i];
        }
        s = wave_sum(s);
        if (lane == 0) part[(size_t)c * nchunks + chunk] = s;
    }
}

// out[c] = sum over chunks of part[c * nchunks + chunk]; one wave per column.
__global__ __launch_bounds__(256) void reduce_cols_kernel(int nchunks, int k,
                                                          const double* __restrict__ part,
                                                          double* __restrict__ out,
                                                          const double* __restrict__ ctl) {
    if (cg_stopped(ctl)) return;
    const int c = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (c >= k) return;
    const double* p = part + (size_t)c * nchunks;
    double s = 0.0;
    for (int i = lane; i < nchunks; i += 64) s += p[i];
    s = wave_sum(s);
    if (lane == 0) out[c] = s;
}

// (Z u)_i for the thread's row i, u staged in LDS (k <= kMaxK).
constexpr int kMaxK = 2048;

__device__ __forceinline__ double z_row_dot(int i, int n, int k, const double* __restrict__ Z,
                                            int ldz, const double* su) {
    double s = 0.0;
    if (i < n)
        for (int c = 0; c < k; ++c) s += Z[i + (size_t)c * ldz] * su[c];
    return s;
}

// Newton setup at a: grad = Z u - 1 + (-1/a + 1/(C-a))/t, D = (1/a^2 + 1/(C-a)^2)/t,
// Pc = M + D, and the CG start s = 0, r = -grad, z = r / Pc, p = z; partials:
// [0] r.z, [1] r.r, [2] max D, [3] min D per workgroup.
__global__ __launch_bounds__(256) void newton_setup_kernel(
    int n, int k, const double* __restrict__ Z, int ldz, const double* __restrict__ u,
    const double* __restrict__ a, const double* __restrict__ M, double C, double inv_t,
    double* __restrict__ grad, double* __restrict__ D, double* __restrict__ Pc,
    double* __restrict__ s, double* __restrict__ r, double* __restrict__ z,
    double* __restrict__ p, double* __restrict__ part) {
    __shared__ double su[kMaxK];
    __shared__ double sh[4];
    __shared__ double smx[4], smn[4];
    for (int c = threadIdx.x; c < k; c += 256) su[c] = u[c];
    __syncthreads();
    const int i = blockIdx.x * kRows + threadIdx.x;
    const double zu = z_row_dot(i, n, k, Z, ldz, su);
    double rz = 0.0, rr = 0.0, dmax = 0.0, dmin = INFINITY;
    if (i < n) {
        const double ai = a[i], ci = C - ai;
        const double g = zu - 1.0 + (-1.0 / ai + 1.0 / ci) * inv_t;
        const double d = (1.0 / (ai * ai) + 1.0 / (ci * ci)) * inv_t;
        const double pc = M[i] + d;
        const double ri = -g, zi = ri / pc;
        grad[i] = g;
        D[i] = d;
        Pc[i] = pc;
        s[i] = 0.0;
        r[i] = ri;
        z[i] = zi;
        p[i] = zi;
        rz = ri * zi;
        rr = ri * ri;
        dmax = d;
        dmin = d;
    }
    const double srz = block_sum(rz, sh);
    const double srr = block_sum(rr, sh);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        dmax = fmax(dmax, __shfl_xor(dmax, o));
        dmin = fmin(dmin, __shfl_xor(dmin, o));
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        smx[wave] = dmax;
        smn[wave] = dmin;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double* P = part + (size_t)blockIdx.x * 4;
        P[0] = srz;
        P[1] = srr;
        P[2] = fmax(fmax(smx[0], smx[1]), fmax(smx[2], smx[3]));
        P[3] = fmin(fmin(smn[0], smn[1]), fmin(smn[2], smn[3]));
    }
}

// CG state on the device: sc[0] = r.z, sc[1] = r.r, sc[2] = max D, sc[3] = min D.
__global__ __launch_bounds__(256) void setup_reduce_kernel(int nblk, const double* __restrict__ part,
                                                           double* __restrict__ sc) {
    __shared__ double sh[4];
    double v0 = 0.0, v1 = 0.0, mx = 0.0, mn = INFINITY;
    for (int b = threadIdx.x; b < nblk; b += 256) {
        v0 += part[4 * b];
        v1 += part[4 * b + 1];
        mx = fmax(mx, part[4 * b + 2]);
        mn = fmin(mn, part[4 * b + 3]);
    }
    const double s0 = block_sum(v0, sh);
    const double s1 = block_sum(v1, sh);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        mx = fmax(mx, __shfl_xor(mx, o));
        mn = fmin(mn, __shfl_xor(mn, o));
    }
    __shared__ double smx[4], smn[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        smx[wave] = mx;
        smn[wave] = mn;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        sc[0] = s0;
        sc[1] = s1;
        sc[2] = fmax(fmax(smx[0], smx[1]), fmax(smx[2], smx[3]));
        sc[3] = fmin(fmin(smn[0], smn[1]), fmin(smn[2], smn[3]));
    }
}

// q = Z v + D o p; part[b] = p.q over the workgroup's rows.
__global__ __launch_bounds__(256) void cg_apply_kernel(int n, int k, const double* __restrict__ Z,
                                                       int ldz, const double* __restrict__ v,
                                                       const double* __restrict__ D,
                                                       const double* __restrict__ p,
                                                       double* __restrict__ q,
                                                       double* __restrict__ part,
                                                       const double* __restrict__ ctl) {
    if (cg_stopped(ctl)) return;
    __shared__ double su[kMaxK];
    __shared__ double sh[4];
    for (int c = threadIdx.x; c < k; c += 256) su[c] = v[c];
    __syncthreads();
    const int i = blockIdx.x * kRows + threadIdx.x;
    const double zv = z_row_dot(i, n, k, Z, ldz, su);
    double pq = 0.0;
    if (i < n) {
        const double qi = zv + D[i] * p[i];
        q[i] = qi;
        pq = p[i] * qi;
    }
    const double s = block_sum(pq, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// alpha = r.z / p.q (p.q summed from the partials in a fixed order by every
// workgroup); s += alpha p, r -= alpha q, z = r / Pc; part2[2b..] = (r.z, r.r).
__global__ __launch_bounds__(256) void cg_update_kernel(int n, int nblk,
                                                        const double* __restrict__ part_pq,
                                                        const double* __restrict__ sc,
                                                        const double* __restrict__ p,
                                                        const double* __restrict__ q,
                                                        const double* __restrict__ Pc,
                                                        double* __restrict__ s,
                                                        double* __restrict__ r,
                                                        double* __restrict__ z,
                                                        double* __restrict__ part2,
                                                        const double* __restrict__ ctl) {
    if (cg_stopped(ctl)) return;
    __shared__ double sh[4];
    const double pq = fixed_sum(nblk, part_pq, sh);
    const double alpha = sc[0] / pq;
    const int i = blockIdx.x * kRows + threadIdx.x;
    double rz = 0.0, rr = 0.0;
    if (i < n) {
        s[i] += alpha * p[i];
        const double ri = r[i] - alpha * q[i];
        const double zi = ri / Pc[i];
        r[i] = ri;
        z[i] = zi;
        rz = ri * zi;
        rr = ri * ri;
    }
    const double a0 = block_sum(rz, sh);
    const double a1 = block_sum(rr, sh);
    if (threadIdx.x == 0) {
        part2[2 * blockIdx.x] = a0;
        part2[2 * blockIdx.x + 1] = a1;
    }
}

// beta = (r.z)_new / (r.z); p = z + beta p.  Every workgroup reads the old
// r.z from sc, so the new r.z and r.r go to sc_next and a separate launch
// (cg_commit_kernel) moves them into sc.
__global__ __launch_bounds__(256) void cg_direction_kernel(int n, int nblk,
                                                           const double* __restrict__ part2,
                                                           const double* __restrict__ sc,
                                                           const double* __restrict__ z,
                                                           double* __restrict__ p,
                                                           double* __restrict__ sc_next,
                                                           const double* __restrict__ ctl) {
    if (cg_stopped(ctl)) return;
    __shared__ double sh[4];
    double v0 = 0.0, v1 = 0.0;
    for (int b = threadIdx.x; b < nblk; b += 256) {
        v0 += part2[2 * b];
        v1 += part2[2 * b + 1];
    }
    v0 = wave_sum(v0);
    v1 = wave_sum(v1);
    __shared__ double sh1[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        sh[wave] = v0;
        sh1[wave] = v1;
    }
    __syncthreads();
    const double rz_new = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    const double rr_new = (sh1[0] + sh1[1]) + (sh1[2] + sh1[3]);
    const double beta = rz_new / sc[0];
    const int i = blockIdx.x * kRows + threadIdx.x;
    if (i < n) p[i] = z[i] + beta * p[i];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        sc_next[0] = rz_new;
        sc_next[1] = rr_new;
    }
}

// End of an iteration: sc[0..1] = sc_next[0..1] (the next iteration's r.z and
// the residual norm), one more iteration counted in ctl[1], and the stopping
// test of the loop, sqrt(r.r) <= tol or max_iter iterations, into ctl[0];
// ctl[2] = r.r for the host.
__global__ void cg_commit_kernel(double* __restrict__ sc, const double* __restrict__ sc_next,
                                 double* __restrict__ ctl, double tol, int max_iter) {
    if (threadIdx.x == 0 && ctl[0] != 0.0) {
        sc[0] = sc_next[0];
        sc[1] = sc_next[1];
        ctl[1] += 1.0;
        ctl[2] = sc[1];
        ctl[0] = (sqrt(sc[1]) > tol && ctl[1] < (double)max_iter) ? 1.0 : 0.0;
    }
}

// Loop start: active, no iterations yet.
__global__ void cg_start_kernel(double* __restrict__ ctl) {
    if (threadIdx.x == 0) {
        ctl[0] = 1.0;
        ctl[1] = 0.0;
    }
}

// Line-search sums per workgroup: part[b * W + j] for j = 0: grad.s,
// 1: sum a, 2: sum s, 3: log barrier at a, 4 + t: log barrier at
// a + tt_t s for the kTrials step sizes tt_t = tt0 * 2^-t (-inf outside the box).
constexpr int kLsW = 4 + kTrials;

__global__ __launch_bounds__(256) void ls_partial_kernel(int n, const double* __restrict__ a,
                                                         const double* __restrict__ s,
                                                         const double* __restrict__ grad,
                                                         double C, double tt0,
                                                         double* __restrict__ part) {
    __shared__ double sh[4];
    const int i = blockIdx.x * kRows + threadIdx.x;
    double v[kLsW];
#pragma unroll
    for (int j = 0; j < kLsW; ++j) v[j] = 0.0;
    if (i < n) {
        const double ai = a[i], si = s[i];
        v[0] = grad[i] * si;
        v[1] = ai;
        v[2] = si;
        v[3] = (ai > 0.0 && C - ai > 0.0) ? log(ai) + log(C - ai) : -INFINITY;
        double tt = tt0;
#pragma unroll
        for (int t = 0; t < kTrials; ++t) {
            const double x = ai + tt * si;
            v[4 + t] = (x > 0.0 && C - x > 0.0) ? log(x) + log(C - x) : -INFINITY;
            tt *= 0.5;
        }
    }
#pragma unroll
    for (int j = 0; j < kLsW; ++j) {
        const double t = block_sum(v[j], sh);
        if (threadIdx.x == 0) part[(size_t)blockIdx.x * kLsW + j] = t;
    }
}

// out[j] = sum over workgroups of part[b * W + j] (fixed order), j < W;
// then out[W..W+2] = u.u, u.v, v.v for the k-vectors u, v.
__global__ __launch_bounds__(256) void ls_reduce_kernel(int nblk, int W,
                                                        const double* __restrict__ part, int k,
                                                        const double* __restrict__ u,
                                                        const double* __restrict__ v,
                                                        double* __restrict__ out) {
    __shared__ double sh[4];
    for (int j = 0; j < W; ++j) {
        double x = 0.0;
        for (int b = threadIdx.x; b < nblk; b += 256) x += part[(size_t)b * W + j];
        const double t = block_sum(x, sh);
        if (threadIdx.x == 0) out[j] = t;
    }
    double uu = 0.0, uv = 0.0, vv = 0.0;
    for (int c = threadIdx.x; c < k; c += 256) {
        const double uc = u[c], vc = v ? v[c] : 0.0;
        uu += uc * uc;
        uv += uc * vc;
        vv += vc * vc;
    }
    const double a0 = block_sum(uu, sh), a1 = block_sum(uv, sh), a2 = block_sum(vv, sh);
    if (threadIdx.x == 0) {
        out[W] = a0;
        out[W + 1] = a1;
        out[W + 2] = a2;
    }
}

// a += tt s.
__global__ __launch_bounds__(256) void axpy_kernel(int n, double tt, const double* __restrict__ s,
                                                   double* __restrict__ a) {
    const int i = blockIdx.x * kRows + threadIdx.x;
    if (i < n) a[i] += tt * s[i];
}

// Initial point (testsvm.cpp:204-219) and M = diag(Z Z^T) (testsvm.cpp:236-243).
__global__ __launch_bounds__(256) void svm_init_kernel(int n, int k, const double* __restrict__ Z,
                                                       int ldz, const double* __restrict__ y,
                                                       double pval, double mval,
                                                       double* __restrict__ a,
                                                       double* __restrict__ M) {
    const int i = blockIdx.x * kRows + threadIdx.x;
    if (i >= n) return;
    a[i] = y[i] > 0.0 ? pval : mval;
    double s = 0.0;
    for (int c = 0; c < k; ++c) {
        const double z = Z[i + (size_t)c * ldz];
        s += z * z;
    }
    M[i] = s;
}

struct Work {
    int n, k, nblk, nchunks, ngrp;
    double *a, *M, *grad, *D, *Pc, *s, *r, *z, *p, *q;
    double *u, *v;              // k-vectors
    double *part, *part2;       // per-workgroup partials
    double *ztpart;             // nchunks x k partials of Z^T x
    double *sc, *sc_next, *ls;  // device scalars
    double* ctl;                // CG loop control: active flag, iterations, r.r
};

int grab(dh_handle_t h, int n, int k, Work& w) {
    w.n = n;
    w.k = k;
    w.nblk = dh_cdiv(n, kRows);
    w.nchunks = w.nblk;
    w.ngrp = dh_cdiv(k, kColGrp);
    const size_t nn = (size_t)std::max(n, 1);
    const size_t doubles = 10 * nn + 2 * (size_t)k + (size_t)w.nblk * (kLsW + 4) +
                           2 * (size_t)w.nblk + (size_t)w.nchunks * k + 80 + kLsW + 8;
    double* base = (double*)dh_ws(h, DH_WS_SVM, sizeof(double) * doubles);
    if (!base) return DH_ERR_ALLOC;
    double* x = base;
    auto take = [&](size_t cnt) {
        double* r = x;
        x += cnt;
        return r;
    };
    w.a = take(nn);
    w.M = take(nn);
    w.grad = take(nn);
    w.D = take(nn);
    w.Pc = take(nn);
    w.s = take(nn);
    w.r = take(nn);
    w.z = take(nn);
    w.p = take(nn);
    w.q = take(nn);
    w.u = take(k);
    w.v = take(k);
    w.part = take((size_t)w.nblk * (kLsW + 4));
    w.part2 = take(2 * (size_t)w.nblk);
    w.ztpart = take((size_t)w.nchunks * k);
    w.sc = take(16);
    w.sc_next = take(16);
    w.ls = take(kLsW + 8 + 32);
    w.ctl = take(16);
    return 0;
}

// out (k) = Z^T x (skipped once the CG loop controlled by ctl has stopped;
// ctl = NULL: always).
int zt(dh_handle_t h, const Work& w, const double* Z, int ldz, const double* x, double* out,
       const double* ctl = nullptr) {
    ztx_partial_kernel<<<dim3(w.nchunks, w.ngrp), 256, 0, h->stream>>>(w.n, w.k, Z, ldz, x,
                                                                       w.ztpart, ctl);
    DH_CHECK_LAUNCH();
    reduce_cols_kernel<<<dh_cdiv(w.k, 4), 256, 0, h->stream>>>(w.nchunks, w.k, w.ztpart, out,
                                                               ctl);
    DH_CHECK_LAUNCH();
    return 0;
}

// Host copies of device scalars (pinned words are 16 ints; a small vector
// of doubles is copied synchronously instead).
int read_scalars(dh_handle_t h, const double* dev, double* host, int count) {
    DH_CHECK_HIP(hipMemcpyAsync(host, dev, sizeof(double) * count, hipMemcpyDeviceToHost,
                                h->stream));
    DH_CHECK_HIP(hipStreamSynchronize(h->stream));
    return 0;
}

// Iterations queued between two reads of the CG loop's control words.
constexpr int kCgBatch = 8;

// Preconditioned CG from the state left by newton_setup_kernel (or
// pcg_setup): iterates until sqrt(r.r) <= tol or max_iter.  The test runs on
// the device at the end of every iteration (cg_commit_kernel); iterations
// queued after it fails do nothing.
int cg_loop(dh_handle_t h, const Work& w, const double* Z, int ldz, double tol, int max_iter,
            int* iters, double* rnorm) {
    double rr = 0.0;
    DH_CHECK_STATUS(read_scalars(h, w.sc + 1, &rr, 1));
    int it = 0;
    if (std::sqrt(rr) > tol && it < max_iter) {
        cg_start_kernel<<<1, 64, 0, h->stream>>>(w.ctl);
        DH_CHECK_LAUNCH();
        for (;;) {
            const int batch = std::min(kCgBatch, max_iter - it);
            for (int b = 0; b < batch; ++b) {
                DH_CHECK_STATUS(zt(h, w, Z, ldz, w.p, w.v, w.ctl));
                cg_apply_kernel<<<w.nblk, 256, 0, h->stream>>>(w.n, w.k, Z, ldz, w.v, w.D, w.p,
                                                                w.q, w.part, w.ctl);
                DH_CHECK_LAUNCH();
                cg_update_kernel<<<w.nblk, 256, 0, h->stream>>>(w.n, w.nblk, w.part, w.sc, w.p,
                                                                 w.q, w.Pc, w.s, w.r, w.z,
                                                                 w.part2, w.ctl);
                DH_CHECK_LAUNCH();
                cg_direction_kernel<<<w.nblk, 256, 0, h->stream>>>(w.n, w.nblk, w.part2, w.sc,
                                                                    w.z, w.p, w.sc_next, w.ctl);
                DH_CHECK_LAUNCH();
                cg_commit_kernel<<<1, 64, 0, h->stream>>>(w.sc, w.sc_next, w.ctl, tol, max_iter);
                DH_CHECK_LAUNCH();
            }
            double c[3];
            DH_CHECK_STATUS(read_scalars(h, w.ctl, c, 3));
            it = (int)c[1];
            rr = c[2];
            if (c[0] == 0.0) break;
        }
    }
    if (iters) *iters = it;
    if (rnorm) *rnorm = std::sqrt(rr);
    return 0;
}

int check_args(const char* name, dh_handle_t h, int n, int k, const void* Z, int ldz) {
    if (!h) return dh_fail(-1, "%s: NULL handle", name);
    if (n < 0) return dh_fail(-2, "%s: n < 0", name);
    if (k < 0 || k > kMaxK) return dh_fail(-3, "%s: need 0 <= k <= %d", name, kMaxK);
    if (n > 0 && k > 0 && !Z) return dh_fail(-4, "%s: Z is NULL", name);
    if (ldz < std::max(1, n)) return dh_fail(-5, "%s: ldz < n", name);
    return 0;
}

// Preconditioner and CG start for dh_svm_pcg: Pc = diag(Z Z^T) + D,
// r = b - (Z Z^T + D) x, z = r / Pc, p = z; partials (r.z, r.r, 0, inf).
__global__ __launch_bounds__(256) void pcg_setup_kernel(
    int n, int k, const double* __restrict__ Z, int ldz, const double* __restrict__ u,
    const double* __restrict__ Dv, const double* __restrict__ b, const double* __restrict__ x,
    double* __restrict__ Pc, double* __restrict__ r, double* __restrict__ z,
    double* __restrict__ p, double* __restrict__ part) {
    __shared__ double su[kMaxK];
    __shared__ double sh[4];
    for (int c = threadIdx.x; c < k; c += 256) su[c] = u[c];
    __syncthreads();
    const int i = blockIdx.x * kRows + threadIdx.x;
    double rz = 0.0, rr = 0.0;
    if (i < n) {
        double zu = 0.0, m = 0.0;
        for (int c = 0; c < k; ++c) {
            const double zc = Z[i + (size_t)c * ldz];
            zu += zc * su[c];
            m += zc * zc;
        }
        const double pc = m + Dv[i];
        const double ri = b[i] - (zu + Dv[i] * x[i]);
        const double zi = ri / pc;
        Pc[i] = pc;
        r[i] = ri;
        z[i] = zi;
        p[i] = zi;
        rz = ri * zi;
        rr = ri * ri;
    }
    const double a0 = block_sum(rz, sh);
    const double a1 = block_sum(rr, sh);
    if (threadIdx.x == 0) {
        double* P = part + (size_t)blockIdx.x * 4;
        P[0] = a0;
        P[1] = a1;
        P[2] = 0.0;
        P[3] = INFINITY;
    }
}

__global__ __launch_bounds__(256) void copy_vec_kernel(int n, const double* __restrict__ src,
                                                       double* __restrict__ dst) {
    const int i = blockIdx.x * kRows + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

}  // namespace

extern "C" int dh_svm_pcg(dh_handle_t h, int n, int k, const double* Z, int ldz, const double* D,
                          const double* b, double* x, double tol, int max_iter, int* iters,
                          double* rnorm) {
    const char* name = "dh_svm_pcg";
    DH_CHECK_STATUS(check_args(name, h, n, k, Z, ldz));
    if (n > 0 && (!D || !b || !x)) return dh_fail(-6, "%s: NULL vector", name);
    if (max_iter < 0) return dh_fail(-10, "%s: max_iter < 0", name);
    if (iters) *iters = 0;
    if (rnorm) *rnorm = 0.0;
    if (n == 0) return 0;
    DH_CHECK_HIP(hipSetDevice(h->device));
    Work w;
    DH_CHECK_STATUS(grab(h, n, k, w));
    // The iterate lives in w.s (the CG loop accumulates into s); D in w.D.
    copy_vec_kernel<<<w.nblk, 256, 0, h->stream>>>(n, x, w.s);
    DH_CHECK_LAUNCH();
    copy_vec_kernel<<<w.nblk, 256, 0, h->stream>>>(n, D, w.D);
    DH_CHECK_LAUNCH();
    DH_CHECK_STATUS(zt(h, w, Z, ldz, w.s, w.u));
    pcg_setup_kernel<<<w.nblk, 256, 0, h->stream>>>(n, k, Z, ldz, w.u, w.D, b, w.s, w.Pc, w.r,
                                                     w.z, w.p, w.part);
    DH_CHECK_LAUNCH();
    setup_reduce_kernel<<<1, 256, 0, h->stream>>>(w.nblk, w.part, w.sc);
    DH_CHECK_LAUNCH();
    DH_CHECK_STATUS(cg_loop(h, w, Z, ldz, tol, max_iter, iters, rnorm));
    copy_vec_kernel<<<w.nblk, 256, 0, h->stream>>>(n, w.s, x);
    DH_CHECK_LAUNCH();
    DH_CHECK_HIP(hipStreamSynchronize(h->stream));
    return 0;
}

extern "C" int dh_svm_barrier(dh_handle_t h, int n, int k, const double* Z, int ldz,
                              const double* y, double C, double cg_tol, int max_outer,
                              int max_newton, int max_cg, double* a, int* stats, double* fval) {
    const char* name = "dh_svm_barrier";
    DH_CHECK_STATUS(check_args(name, h, n, k, Z, ldz));
    if (n > 0 && (!y || !a)) return dh_fail(-6, "%s: NULL vector", name);
    if (!(C > 0.0)) return dh_fail(-7, "%s: C must be > 0", name);
    if (max_outer < 1 || max_newton < 1 || max_cg < 0)
        return dh_fail(-9, "%s: iteration bounds must be positive", name);
    if (stats) stats[0] = stats[1] = stats[2] = 0;
    if (fval) *fval = 0.0;
    if (n == 0) return 0;
    DH_CHECK_HIP(hipSetDevice(h->device));
    // Class sizes from the labels (host copy of y).
    std::vector<double> hy(n);
    DH_CHECK_HIP(hipMemcpyAsync(hy.data(), y, sizeof(double) * n, hipMemcpyDeviceToHost,
                                h->stream));
    DH_CHECK_HIP(hipStreamSynchronize(h->stream));
    int plus = 0, minus = 0;
    for (int i = 0; i < n; ++i) {
        if (hy[i] > 0) ++plus;
        else ++minus;
    }
    if (plus == 0 || minus == 0)
        return dh_fail(-6, "%s: both classes must be present (%d +1, %d -1)", name, plus, minus);
    double pval, mval;  // testsvm.cpp:204-211
    if (plus > minus) {
        mval = 0.9 * C;
        pval = mval * minus / plus;
    } else {
        pval = 0.9 * C;
        mval = pval * plus / minus;
    }
    Work w;
    DH_CHECK_STATUS(grab(h, n, k, w));
    svm_init_kernel<<<w.nblk, 256, 0, h->stream>>>(n, k, Z, ldz, y, pval, mval, w.a, w.M);
    DH_CHECK_LAUNCH();
    const double ls_alpha = 0.1, mu = 20.0;
    double t = 1.0, fcur = 0.0;
    int outer = 0, newton = 0, cgit = 0;
    std::vector<double> hs(kLsW + 3);
    for (outer = 1; outer <= max_outer; ++outer) {
        for (int ni = 1; ni <= max_newton; ++ni) {
            DH_CHECK_STATUS(zt(h, w, Z, ldz, w.a, w.u));
            newton_setup_kernel<<<w.nblk, 256, 0, h->stream>>>(n, k, Z, ldz, w.u, w.a, w.M, C,
                                                                1.0 / t, w.grad, w.D, w.Pc, w.s,
                                                                w.r, w.z, w.p, w.part);
            DH_CHECK_LAUNCH();
            setup_reduce_kernel<<<1, 256, 0, h->stream>>>(w.nblk, w.part, w.sc);
            DH_CHECK_LAUNCH();
            double dd[4];
            DH_CHECK_STATUS(read_scalars(h, w.sc, dd, 4));
            if (dd[2] / dd[3] > 1e18) break;  // D too ill-conditioned (testsvm.cpp:252-256)
            int its = 0;
            DH_CHECK_STATUS(cg_loop(h, w, Z, ldz, cg_tol, max_cg, &its, nullptr));
            cgit += its;
            ++newton;
            // Line search: f(a + tt s) from u = Z^T a and v = Z^T s, barrier sums per trial.
            DH_CHECK_STATUS(zt(h, w, Z, ldz, w.s, w.v));
            double tt0 = 1.0, g0 = 0.0, gs = 0.0, tt = 0.0, gnew = 0.0;
            bool accepted = false;
            for (int batch = 0; batch < 4 && !accepted; ++batch) {
                ls_partial_kernel<<<w.nblk, 256, 0, h->stream>>>(n, w.a, w.s, w.grad, C, tt0,
                                                                  w.part);
                DH_CHECK_LAUNCH();
                ls_reduce_kernel<<<1, 256, 0, h->stream>>>(w.nblk, kLsW, w.part, k, w.u, w.v,
                                                           w.ls);
                DH_CHECK_LAUNCH();
                DH_CHECK_STATUS(read_scalars(h, w.ls, hs.data(), kLsW + 3));
                gs = hs[0];
                const double sum_a = hs[1], sum_s = hs[2], lb0 = hs[3];
                const double uu = hs[kLsW], uv = hs[kLsW + 1], vv = hs[kLsW + 2];
                g0 = 0.5 * uu - sum_a - lb0 / t;
                double trial = tt0;
                for (int j = 0; j < kTrials; ++j, trial *= 0.5) {
                    const double fn = 0.5 * (uu + 2.0 * trial * uv + trial * trial * vv) -
                                      (sum_a + trial * sum_s);
                    const double g = fn - hs[4 + j] / t;
                    if (g <= g0 + ls_alpha * trial * gs) {
                        tt = trial;
                        gnew = g;
                        accepted = true;
                        break;
                    }
                }
                tt0 = trial;  // next batch continues the halving
            }
            if (!accepted) {
                tt = 0.0;
                gnew = g0;
            }
            if (tt > 0.0) {
                axpy_kernel<<<w.nblk, 256, 0, h->stream>>>(n, tt, w.s, w.a);
                DH_CHECK_LAUNCH();
            }
            const double nd2 = -gs;
            if (nd2 / 2 < 1e-6) break;                                  // testsvm.cpp:287-290
            if (ni > 1 && (g0 - gnew) < 1e-9 * std::fabs(gnew)) break;  // testsvm.cpp:293-299
        }
        // f(a) for the duality-gap test (testsvm.cpp:307-311).
        DH_CHECK_STATUS(zt(h, w, Z, ldz, w.a, w.u));
        ls_partial_kernel<<<w.nblk, 256, 0, h->stream>>>(n, w.a, w.a, w.a, C, 0.0, w.part);
        DH_CHECK_LAUNCH();
        ls_reduce_kernel<<<1, 256, 0, h->stream>>>(w.nblk, kLsW, w.part, k, w.u, nullptr, w.ls);
        DH_CHECK_LAUNCH();
        DH_CHECK_STATUS(read_scalars(h, w.ls, hs.data(), kLsW + 3));
        fcur = 0.5 * hs[kLsW] - hs[1];
        if (2.0 * n / t < std::fabs(fcur) * 1e-5) break;
        t *= mu;
    }
    if (outer > max_outer) outer = max_outer;
    DH_CHECK_HIP(hipMemcpyAsync(a, w.a, sizeof(double) * n, hipMemcpyDeviceToDevice, h->stream));
    DH_CHECK_HIP(hipStreamSynchronize(h->stream));
    if (stats) {
        stats[0] = outer;
        stats[1] = newton;
        stats[2] = cgit;
    }
    if (fval) *fval = fcur;
    return 0;
}
