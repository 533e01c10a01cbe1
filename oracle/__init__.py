"""ctypes binding of the FP64 CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.

Matrices are numpy arrays in column-major (Fortran) order with ld = rows;
point sets are (n, d) C-ordered arrays (one point per row = DArray's d x n
column-major training matrix).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().with_name("liboracle.so")
_lib = None

_i, _c, _d, _p = C.c_int, C.c_char, C.c_double, C.c_void_p
_SIG = {
    "orc_sqnorms": (None, [_i, _i, _p, _i, _p]),
    "orc_rbf_tile": (None, [_i, _i, _i, _p, _i, _p, _i, _p, _p, _d, _i, _p, _i]),
    "orc_rbf_project": (None, [_i, _i, _i, _i, _p, _i, _p, _i, _p, _p, _d, _p, _i, _d, _p, _i,
                               _i]),
    "orc_gemm": (None, [_c, _c, _i, _i, _i, _d, _p, _i, _p, _i, _d, _p, _i]),
    "orc_matrix_multiply_tn": (None, [_i, _i, _i, _d, _p, _i, _p, _i, _d, _p, _i, _i]),
    "orc_syrk_ut": (None, [_i, _i, _d, _p, _i, _d, _p, _i]),
    "orc_potf2_upper": (_i, [_i, _p, _i]),
    "orc_potrf_upper": (_i, [_i, _p, _i, _i]),
    "orc_potrf_upper_recursive": (_i, [_i, _p, _i, _i]),
    "orc_trsm_lut": (None, [_c, _i, _i, _d, _p, _i, _p, _i, _i]),
    "orc_trsm_lun": (None, [_c, _i, _i, _d, _p, _i, _p, _i, _i]),
    "orc_trsm_run": (None, [_c, _i, _i, _d, _p, _i, _p, _i]),
    "orc_householder_qr": (None, [_i, _i, _p, _i, _p, _i]),
    "orc_cholqr2": (_i, [_i, _i, _p, _i, _p, _i]),
    "orc_lra": (_i, [_i, _i, _i, _i, _p, _i, _p, _d, _p, _i, _i, _i, _p, _i, _p, _i]),
    "orc_lra_error": (None, [_i, _i, _i, _p, _i, _p, _d, _p, _i, _p, _i, _i, _p]),
    "orc_gesvj": (_i, [_i, _i, _p, _i, _p, _p, _i, _d, _i]),
    "orc_nystrom": (_i, [_i, _i, _i, _p, _i, _p, _i, _p, _i, _p]),
    "orc_svm_barrier": (_i, [_i, _i, _p, _i, _p, _d, _d, _i, _i, _i, _p, _p, _p]),
    "orc_set_threads": (None, [_i]),
    "orc_get_threads": (_i, []),
}


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        h = C.CDLL(str(LIB_PATH))
        for k, (r, a) in _SIG.items():
            f = getattr(h, k)
            f.restype = r
            f.argtypes = a
        _lib = h
    return _lib


def _p_(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _f(a):
    """Fortran-ordered float64 copy (or the array itself if already so)."""
    return np.asfortranarray(a, dtype=np.float64)


def _pts(X):
    return np.ascontiguousarray(X, dtype=np.float64)


def set_threads(n: int):
    lib().orc_set_threads(int(n))


def sqnorms(X):
    X = _pts(X)
    n, d = X.shape
    out = np.empty(n)
    lib().orc_sqnorms(n, d, _p_(X), d, _p_(out))
    return out


def rbf_tile(Xi, Xj, gamma, yi=None, yj=None, clamp=False):
    Xi, Xj = _pts(Xi), _pts(Xj)
    m, d = Xi.shape
    n = Xj.shape[0]
    K = np.zeros((m, n), order="F")
    yi = None if yi is None else np.ascontiguousarray(yi, dtype=np.float64)
    yj = None if yj is None else np.ascontiguousarray(yj, dtype=np.float64)
    lib().orc_rbf_tile(m, n, d, _p_(Xi), d, _p_(Xj), d, _p_(yi), _p_(yj), gamma, int(clamp),
                       _p_(K), max(1, m))
    return K


def rbf_project(Xi, Xj, O, gamma, yi=None, yj=None, beta=0.0, Y=None, chunk=0):
    Xi, Xj, O = _pts(Xi), _pts(Xj), _f(O)
    m, d = Xi.shape
    n, l = O.shape
    Y = np.zeros((m, l), order="F") if Y is None else _f(Y).copy(order="F")
    yi = None if yi is None else np.ascontiguousarray(yi, dtype=np.float64)
    yj = None if yj is None else np.ascontiguousarray(yj, dtype=np.float64)
    lib().orc_rbf_project(m, n, d, l, _p_(Xi), d, _p_(Xj), d, _p_(yi), _p_(yj), gamma, _p_(O),
                          max(1, n), beta, _p_(Y), max(1, m), chunk)
    return Y


def gemm(ta, tb, A, B, alpha=1.0, beta=0.0, Cm=None):
    A, B = _f(A), _f(B)
    m = A.shape[1] if ta == "T" else A.shape[0]
    k = A.shape[0] if ta == "T" else A.shape[1]
    n = B.shape[0] if tb == "T" else B.shape[1]
    Cm = np.zeros((m, n), order="F") if Cm is None else _f(Cm).copy(order="F")
    lib().orc_gemm(ta.encode(), tb.encode(), m, n, k, alpha, _p_(A), max(1, A.shape[0]), _p_(B),
                   max(1, B.shape[0]), beta, _p_(Cm), max(1, m))
    return Cm


def matrix_multiply_tn(A, B, alpha=1.0, beta=0.0, Cm=None, bs=128):
    A, B = _f(A), _f(B)
    k, m = A.shape
    n = B.shape[1]
    Cm = np.zeros((m, n), order="F") if Cm is None else _f(Cm).copy(order="F")
    lib().orc_matrix_multiply_tn(m, n, k, alpha, _p_(A), max(1, k), _p_(B), max(1, k), beta,
                                 _p_(Cm), max(1, m), bs)
    return Cm


def syrk_ut(A, Cm, alpha=1.0, beta=0.0):
    A = _f(A)
    k, n = A.shape
    Cm = _f(Cm).copy(order="F")
    lib().orc_syrk_ut(n, k, alpha, _p_(A), max(1, k), beta, _p_(Cm), max(1, n))
    return Cm


def potf2_upper(A):
    A = _f(A).copy(order="F")
    info = lib().orc_potf2_upper(A.shape[0], _p_(A), max(1, A.shape[0]))
    return A, info


def potrf_upper(A, bs=256):
    A = _f(A).copy(order="F")
    info = lib().orc_potrf_upper(A.shape[0], _p_(A), max(1, A.shape[0]), bs)
    return A, info


def potrf_upper_recursive(A, bs=32):
    A = _f(A).copy(order="F")
    info = lib().orc_potrf_upper_recursive(A.shape[0], _p_(A), max(1, A.shape[0]), bs)
    return A, info


def trsm(side, trans, A, B, alpha=1.0, diag="N", bs=64):
    """side L: B := alpha op(U)^{-1} B; side R (trans N): B := alpha B U^{-1}."""
    A = _f(A)
    B = _f(B).copy(order="F")
    m, n = B.shape
    ldb = max(1, m)
    lda = max(1, A.shape[0])
    if side == "L" and trans == "T":
        lib().orc_trsm_lut(diag.encode(), m, n, alpha, _p_(A), lda, _p_(B), ldb, bs)
    elif side == "L":
        lib().orc_trsm_lun(diag.encode(), m, n, alpha, _p_(A), lda, _p_(B), ldb, bs)
    else:
        lib().orc_trsm_run(diag.encode(), m, n, alpha, _p_(A), lda, _p_(B), ldb)
    return B


def householder_qr(A):
    A = _f(A).copy(order="F")
    m, n = A.shape
    R = np.zeros((n, n), order="F")
    lib().orc_householder_qr(m, n, _p_(A), max(1, m), _p_(R), max(1, n))
    return A, R


def cholqr2(Y):
    Y = _f(Y).copy(order="F")
    m, n = Y.shape
    R = np.zeros((n, n), order="F")
    info = lib().orc_cholqr2(m, n, _p_(Y), max(1, m), _p_(R), max(1, n))
    return Y, R, info


def lra(X, Omega, gamma, q, y=None, orth="cholqr2", chunk=0):
    """Returns (Q, W, info); B = W^T."""
    X, Omega = _pts(X), _f(Omega)
    n, d = X.shape
    l = Omega.shape[1]
    Q = np.zeros((n, l), order="F")
    W = np.zeros((n, l), order="F")
    y = None if y is None else np.ascontiguousarray(y, dtype=np.float64)
    info = lib().orc_lra(n, d, l, q, _p_(X), d, _p_(y), gamma, _p_(Omega), max(1, n),
                         0 if orth == "householder" else 1, chunk, _p_(Q), max(1, n), _p_(W),
                         max(1, n))
    return Q, W, info


def lra_error(X, gamma, Q, W, y=None):
    """(||K - Q W^T||_F^2, ||K||_F^2)."""
    X, Q, W = _pts(X), _f(Q), _f(W)
    n, d = X.shape
    l = Q.shape[1]
    out = np.zeros(2)
    y = None if y is None else np.ascontiguousarray(y, dtype=np.float64)
    lib().orc_lra_error(n, d, l, _p_(X), d, _p_(y), gamma, _p_(Q), max(1, n), _p_(W), max(1, n),
                        0, _p_(out))
    return out[0], out[1]


def gesvj(A, tol=1e-14, max_sweeps=60):
    """One-sided Jacobi SVD: returns (A V = U diag(s), s descending, V, sweeps)."""
    A = _f(A).copy(order="F")
    m, n = A.shape
    s = np.zeros(n)
    V = np.zeros((n, n), order="F")
    sweeps = lib().orc_gesvj(m, n, _p_(A), max(1, m), _p_(s), _p_(V), max(1, n), tol, max_sweeps)
    return A, s, V, sweeps


def nystrom(Q, W, k):
    """Nystrom factor Z (n x k) of the range finder's Q, W = K Q and the k
    largest eigenvalues of Q^T K Q; returns (Z, lam, status)."""
    Q, W = _f(Q), _f(W)
    n, l = Q.shape
    Z = np.zeros((n, k), order="F")
    lam = np.zeros(k)
    st = lib().orc_nystrom(n, l, k, _p_(Q), max(1, n), _p_(W), max(1, n), _p_(Z), max(1, n),
                           _p_(lam))
    return Z, lam, st


def svm_barrier(Z, y, Cb, cg_tol=1e-2, max_outer=50, max_newton=100, max_cg=None):
    """Barrier-method SVM dual on the Nystrom factor Z (n x k); returns
    (a, [outer, newton, cg], f(a), status)."""
    Z = _f(Z)
    n, k = Z.shape
    y = np.ascontiguousarray(y, dtype=np.float64)
    a = np.zeros(n)
    st = np.zeros(3, dtype=np.int32)
    fv = C.c_double(0.0)
    r = lib().orc_svm_barrier(n, k, _p_(Z), max(1, n), _p_(y), float(Cb), float(cg_tol),
                              int(max_outer), int(max_newton),
                              int(max_cg if max_cg is not None else 10 * n), _p_(a), _p_(st),
                              C.byref(fv))
    return a, [int(v) for v in st], fv.value, r
