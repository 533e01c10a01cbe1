"""MI355X-native RBF randomised low-rank approximation (DArray's kernel-SVM hot path).

The compute path is libdarray_hip.so (hand-written gfx950 HIP kernels behind
the C ABI in include/darray_hip.h).  This module is a thin host-side mirror
for Python callers: device buffers come from torch, and matrices follow
DArray's column-major convention -- a logical m x n matrix lives in a torch
tensor of shape (n, m) (row j of the tensor is column j of the matrix), so a
point set X (d x n, one point per column) is simply an (n, d) tensor.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi
from ._capi import DH_PREC_BF16, DH_PREC_F16, DH_PREC_F64, DH_PREC_FP8, DHError, call, lib

__all__ = ["Handle", "DHError", "DH_PREC_F64", "DH_PREC_BF16", "DH_PREC_F16", "DH_PREC_FP8", "lib", "fill_uniform",
           "fill_normal", "fill_mixture", "mt19937_raw", "to_dev", "from_dev", "read_libsvm",
           "write_libsvm_model"]


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _ld(t):
    """Leading dimension of a column-major matrix held as a (cols, rows) tensor."""
    assert t.dim() == 2, "expected a 2-D (cols, rows) tensor"
    if t.numel() == 0:
        return max(1, t.shape[1])
    assert t.stride(1) == 1 or t.shape[1] == 1, "rows must be contiguous"
    return max(1, t.stride(0), t.shape[1])


def to_dev(a: np.ndarray, dtype=None, device="cuda"):
    """Logical column-major matrix (m x n numpy array) -> (n, m) device tensor."""
    import torch
    a = np.asarray(a)
    if a.ndim == 1:
        a = a[:, None]
    t = torch.from_numpy(np.ascontiguousarray(a.T))
    if dtype is not None:
        t = t.to(dtype)
    return t.to(device)


def from_dev(t) -> np.ndarray:
    """(n, m) device tensor -> logical m x n numpy array."""
    return t.detach().cpu().numpy().T


class Handle:
    """One device + one stream + workspaces (dh_create / dh_destroy)."""

    def __init__(self, device: int = 0, stream=None):
        import torch
        if stream is None:
            stream = torch.cuda.current_stream(device)
        self.stream = stream
        h = C.c_void_p()
        call("dh_create", device, C.c_void_p(stream.cuda_stream), C.byref(h))
        self.h = h

    def close(self):
        if self.h:
            lib().dh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        call("dh_synchronize", self.h)

    def timing(self, on: bool):
        """Record HIP events around the fused projection kernels (dh_timing_enable)."""
        call("dh_timing_enable", self.h, 1 if on else 0)

    def timing_read(self):
        """(summed kernel milliseconds, launches) since the previous read."""
        ms, n = C.c_double(), C.c_int()
        call("dh_timing_read", self.h, C.byref(ms), C.byref(n))
        return ms.value, n.value

    def workspace_bytes(self) -> int:
        return lib().dh_workspace_bytes(self.h)

    # ---------------------------------------------------------------- RBF
    def sqnorms_f64(self, X, out):
        n, d = X.shape
        call("dh_sqnorms_f64", self.h, n, d, _ptr(X), _ld(X), _ptr(out))
        return out

    def rbf_tile(self, Xi, Xj, gamma, K, yi=None, yj=None):
        """K (ldk = K rows) = RBF tile of the point sets Xi (m, d), Xj (n, d)."""
        import torch
        m, d = Xi.shape
        n = Xj.shape[0]
        fn = "dh_rbf_tile_f64" if Xi.dtype == torch.float64 else "dh_rbf_tile_f32"
        call(fn, self.h, m, n, d, _ptr(Xi), _ld(Xi), _ptr(Xj), _ld(Xj), _ptr(yi), _ptr(yj),
             float(gamma), _ptr(K), _ld(K))
        return K

    def rbf_project_f64(self, Xi, Xj, O, gamma, Y, yi=None, yj=None, beta=0.0):
        """Y (m x l) = K(Xi, Xj) O + beta Y; O is an (l, n) tensor, Y (l, m)."""
        m, d = Xi.shape
        n = Xj.shape[0]
        l = O.shape[0]
        call("dh_rbf_project_f64", self.h, m, n, d, l, _ptr(Xi), _ld(Xi), _ptr(Xj), _ld(Xj),
             _ptr(yi), _ptr(yj), float(gamma), _ptr(O), _ld(O), float(beta), _ptr(Y), _ld(Y))
        return Y

    def rbf_project_bf16(self, Xi, Xj, O, gamma, Y, yi=None, yj=None):
        """Y (m x l, FP64) = K(Xi, Xj) O through the BF16 MFMA path (FP64 in/out)."""
        m, d = Xi.shape
        n = Xj.shape[0]
        l = O.shape[0]
        call("dh_rbf_project_bf16", self.h, m, n, d, l, _ptr(Xi), _ld(Xi), _ptr(Xj), _ld(Xj),
             _ptr(yi), _ptr(yj), float(gamma), _ptr(O), _ld(O), _ptr(Y), _ld(Y))
        return Y

    def rbf_project_fp8(self, Xi, Xj, O, gamma, Y, yi=None, yj=None):
        """FP8 e4m3 distance, BF16 projection (FP64 in/out)."""
        m, d = Xi.shape
        n = Xj.shape[0]
        l = O.shape[0]
        call("dh_rbf_project_fp8", self.h, m, n, d, l, _ptr(Xi), _ld(Xi), _ptr(Xj), _ld(Xj),
             _ptr(yi), _ptr(yj), float(gamma), _ptr(O), _ld(O), _ptr(Y), _ld(Y))
        return Y

    def rbf_project_f16(self, Xi, Xj, O, gamma, Y, yi=None, yj=None):
        """As rbf_project_bf16 with FP16 operands (FP64 in/out)."""
        m, d = Xi.shape
        n = Xj.shape[0]
        l = O.shape[0]
        call("dh_rbf_project_f16", self.h, m, n, d, l, _ptr(Xi), _ld(Xi), _ptr(Xj), _ld(Xj),
             _ptr(yi), _ptr(yj), float(gamma), _ptr(O), _ld(O), _ptr(Y), _ld(Y))
        return Y

    # -------------------------------------------------------------- dense
    def gemm(self, ta, tb, A, B, C_, alpha=1.0, beta=0.0):
        import torch
        m = C_.shape[1]
        n = C_.shape[0]
        k = A.shape[1] if ta.upper() == "T" else A.shape[0]
        fn = "dh_gemm_f64" if C_.dtype == torch.float64 else "dh_gemm_f32"
        call(fn, self.h, ta.encode(), tb.encode(), m, n, k, alpha, _ptr(A), _ld(A), _ptr(B),
             _ld(B), beta, _ptr(C_), _ld(C_))
        return C_

    def syrk_ut(self, A, C_, alpha=1.0, beta=0.0):
        import torch
        k, n = A.shape[1], A.shape[0]
        fn = "dh_syrk_ut_f64" if C_.dtype == torch.float64 else "dh_syrk_ut_f32"
        call(fn, self.h, n, k, alpha, _ptr(A), _ld(A), beta, _ptr(C_), _ld(C_))
        return C_

    def trsm(self, side, trans, diag, A, B, alpha=1.0, uplo="U"):
        import torch
        m, n = B.shape[1], B.shape[0]
        fn = "dh_trsm_f64" if B.dtype == torch.float64 else "dh_trsm_f32"
        call(fn, self.h, side.encode(), uplo.encode(), trans.encode(), diag.encode(), m, n,
             alpha, _ptr(A), _ld(A), _ptr(B), _ld(B))
        return B

    def potrf(self, A, bs=256, uplo="U") -> int:
        """Upper Cholesky in place; returns LAPACK info (0 = success)."""
        import torch
        n = A.shape[0]
        fn = "dh_potrf_f64" if A.dtype == torch.float64 else "dh_potrf_f32"
        return call(fn, self.h, uplo.encode(), n, _ptr(A), _ld(A), bs)

    def posv(self, A, B, bs=256, uplo="U") -> int:
        """FP64 Cholesky solve in place: A <- U, B <- A^{-1} B; returns info."""
        n, nrhs = A.shape[0], B.shape[0]
        return call("dh_posv_f64", self.h, uplo.encode(), n, nrhs, _ptr(A), _ld(A), _ptr(B),
                    _ld(B), bs)

    def posv_mixed(self, A, B, X, bs=256, tol=1e-10, max_iter=50, uplo="U"):
        """FP32 Cholesky + FP64 refinement: X = A^{-1} B; returns (info, iters, berr)."""
        n, nrhs = A.shape[0], B.shape[0]
        it = C.c_int(0)
        be = C.c_double(0.0)
        info = call("dh_posv_mixed", self.h, uplo.encode(), n, nrhs, _ptr(A), _ld(A), _ptr(B),
                    _ld(B), _ptr(X), _ld(X), bs, tol, max_iter, C.byref(it), C.byref(be))
        return info, it.value, be.value

    def cholqr2(self, Y, R=None) -> int:
        m, n = Y.shape[1], Y.shape[0]
        return call("dh_cholqr2_f64", self.h, m, n, _ptr(Y), _ld(Y), _ptr(R),
                    _ld(R) if R is not None else 1)

    # ---------------------------------------------------------------- LRA
    def lra(self, X, Omega, gamma, q, Q, W, y=None, prec=DH_PREC_F64) -> int:
        n, d = X.shape
        l = Omega.shape[0]
        return call("dh_lra", self.h, prec, n, d, l, q, _ptr(X), _ld(X), _ptr(y), float(gamma),
                    _ptr(Omega), _ld(Omega), _ptr(Q), _ld(Q), _ptr(W), _ld(W))

    def lra_batched(self, problems, prec=DH_PREC_BF16):
        """dh_lra_batched over a list of dicts with keys X, Omega, gamma, q, Q, W
        and optionally y (tensors as for lra).  Returns (status, per-problem statuses)."""
        arr = (_capi.LraProblem * len(problems))()
        for i, p in enumerate(problems):
            X, Om, Q, W = p["X"], p["Omega"], p["Q"], p["W"]
            y = p.get("y")
            arr[i] = _capi.LraProblem(X.shape[0], X.shape[1], Om.shape[0], p["q"], X.data_ptr(),
                                      _ld(X), None if y is None else y.data_ptr(),
                                      float(p["gamma"]), Om.data_ptr(), _ld(Om), Q.data_ptr(),
                                      _ld(Q), W.data_ptr(), _ld(W))
        st = (C.c_int * max(1, len(problems)))()
        r = call("dh_lra_batched", self.h, prec, len(problems), arr, st)
        return r, list(st)[:len(problems)]

    def gesvj(self, A, sva, V):
        """One-sided Jacobi SVD in place: A (n, m) tensor <- U diag(sva); returns
        (status, sweeps)."""
        m, n = A.shape[1], A.shape[0]
        sw = C.c_int(0)
        st = call("dh_gesvj_f64", self.h, m, n, _ptr(A), _ld(A), _ptr(sva), _ptr(V), _ld(V),
                  C.byref(sw))
        return st, sw.value

    def nystrom(self, Q, W, k, Z, lam=None) -> int:
        """Z (n x k) with Z Z^T the rank-k truncation of Q Q^T K Q Q^T, from Q and
        W = K Q ((l, n) tensors); lam (k) receives the eigenvalues of Q^T K Q."""
        n, l = Q.shape[1], Q.shape[0]
        return call("dh_nystrom_f64", self.h, n, l, k, _ptr(Q), _ld(Q), _ptr(W), _ld(W), _ptr(Z),
                    _ld(Z), _ptr(lam))

    def svm_pcg(self, Z, D, b, x, tol=1e-2, max_iter=1000):
        """(Z Z^T + diag(D)) x = b by preconditioned CG; Z a (k, n) tensor (n x k
        column-major), x the initial guess (updated).  Returns (iters, rnorm)."""
        n, k = Z.shape[1], Z.shape[0]
        it, rn = C.c_int(0), C.c_double(0.0)
        call("dh_svm_pcg", self.h, n, k, _ptr(Z), _ld(Z), _ptr(D), _ptr(b), _ptr(x), float(tol),
             int(max_iter), C.byref(it), C.byref(rn))
        return it.value, rn.value

    def svm_barrier(self, Z, y, C_, a, cg_tol=1e-2, max_outer=50, max_newton=100, max_cg=None):
        """Barrier-method SVM dual on the Nystrom factor Z ((k, n) tensor); a (n)
        receives the solution.  Returns (stats [outer, newton, cg], f(a))."""
        n, k = Z.shape[1], Z.shape[0]
        st = (C.c_int * 3)()
        fv = C.c_double(0.0)
        call("dh_svm_barrier", self.h, n, k, _ptr(Z), _ld(Z), _ptr(y), float(C_), float(cg_tol),
             int(max_outer), int(max_newton), int(max_cg if max_cg is not None else 10 * n),
             _ptr(a), st, C.byref(fv))
        return list(st), fv.value

    def lra_error(self, X, gamma, Q, W, y=None):
        """(||K - Q W^T||_F^2, ||K||_F^2) in FP64."""
        n, d = X.shape
        l = Q.shape[0]
        out = (C.c_double * 2)()
        call("dh_lra_error", self.h, n, d, l, _ptr(X), _ld(X), _ptr(y), float(gamma), _ptr(Q),
             _ld(Q), _ptr(W), _ld(W), out)
        return out[0], out[1]


# ------------------------------------------------------ host generators

def mt19937_raw(seed: int, count: int) -> np.ndarray:
    out = np.empty(count, dtype=np.uint32)
    call("dh_mt19937_raw", seed, count, out.ctypes.data_as(C.POINTER(C.c_uint32)))
    return out


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def fill_uniform(seed: int, m: int, n: int, a=0.0, b=1.0) -> np.ndarray:
    """DMatrix::set_uniform with a fixed seed; returns the logical m x n matrix."""
    buf = np.empty((n, m), dtype=np.float64)  # column-major storage
    call("dh_fill_uniform", seed, a, b, m, n, _dptr(buf), max(1, m))
    return buf.T


def fill_normal(seed: int, m: int, n: int, mean=0.0, std=1.0) -> np.ndarray:
    """DMatrix::set_normal_seed; returns the logical m x n matrix."""
    buf = np.empty((n, m), dtype=np.float64)
    call("dh_fill_normal", seed, mean, std, m, n, _dptr(buf), max(1, m))
    return buf.T


def fill_mixture(seed: int, d: int, n: int, ncentres: int, sigma: float, labels=False):
    """Gaussian-mixture point set as an (n, d) array (point per row); optional labels."""
    X = np.empty((n, d), dtype=np.float64)
    y = np.empty(n, dtype=np.float64) if labels else None
    call("dh_fill_mixture", seed, d, n, ncentres, sigma, _dptr(X), d,
         _dptr(y) if labels else None, None)
    return (X, y) if labels else X


# ---------------------------------------------------------- LIBSVM files

def read_libsvm(path, n: int, d: int):
    """(X (n, d) points as rows, y (n), points read) from a LIBSVM data file."""
    X = np.empty((n, d), dtype=np.float64)  # = d x n column-major, ld d
    y = np.empty(n, dtype=np.float64)
    got = call("dh_read_libsvm", str(path).encode(), n, d, _dptr(X), max(1, d), _dptr(y))
    return X, y, got


def write_libsvm_model(path, a, C_, Z, y, X, gamma):
    """Write the LIBSVM model of the dual solution a; Z is (n, k) (Nystrom
    factor rows), X (n, d) points as rows.  Returns (support vectors, bias)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    Zf = np.asfortranarray(Z, dtype=np.float64)
    Xp = np.ascontiguousarray(X, dtype=np.float64)
    n, k = Zf.shape
    d = Xp.shape[1]
    b = C.c_double(0.0)
    nsv = call("dh_write_libsvm_model", str(path).encode(), n, d, k, _dptr(a), float(C_), _dptr(Zf),
               max(1, n), _dptr(y), _dptr(Xp), max(1, d), float(gamma), C.byref(b))
    return nsv, b.value
