"""Pins the FP64 oracle: golden vectors (numpy/scipy), the reference's own
known-answer Cholesky tests, closed forms and invariants.  CPU only."""
from pathlib import Path

import numpy as np
import pytest
import scipy.linalg as sla

import oracle as orc

G = np.load(Path(__file__).resolve().parent / "golden" / "golden.npz")


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300)


# ---------------------------------------------------------------- RBF tile

def test_rbf_tile_golden():
    K = orc.rbf_tile(G["tile_Xi"], G["tile_Xj"], float(G["tile_gamma"]), G["tile_yi"],
                     G["tile_yj"])
    assert rel(K, G["tile_K"]) < 1e-14


def test_rbf_tile_known_distances():
    # Points on a line at integer positions: |xi - xj|^2 = (i - j)^2 exactly.
    X = np.zeros((6, 3))
    X[:, 1] = np.arange(6)
    K = orc.rbf_tile(X, X, 0.25)
    i, j = np.meshgrid(np.arange(6), np.arange(6), indexing="ij")
    np.testing.assert_allclose(K, np.exp(-0.25 * (i - j) ** 2), rtol=1e-15, atol=0)


def test_rbf_tile_invariants():
    rng = np.random.default_rng(1)
    X = rng.uniform(0, 1, (40, 9))
    y = rng.choice([-1.0, 1.0], 40)
    K = orc.rbf_tile(X, X, 0.3, y, y, clamp=True)
    # Self-tile diagonal: d2 = 2|x|^2 - 2 x.x is rounding noise (clamped at 0 from below).
    np.testing.assert_allclose(np.diag(K), y * y, rtol=0, atol=1e-14)
    np.testing.assert_allclose(K, K.T, rtol=0, atol=1e-15)
    K1 = orc.rbf_tile(X, X, 0.3)
    assert np.all(K1 > 0) and np.all(K1 <= 1 + 1e-15)


def test_sqnorms_exact():
    X = np.arange(12, dtype=np.float64).reshape(3, 4)
    np.testing.assert_array_equal(orc.sqnorms(X), (X**2).sum(1))


# --------------------------------------------------------------- projection

def test_rbf_project_golden():
    Y = orc.rbf_project(G["proj_X"], G["proj_X"], G["proj_O"], float(G["proj_gamma"]),
                        beta=float(G["proj_beta"]), Y=G["proj_Y0"])
    assert rel(Y, G["proj_Y"]) < 1e-14


@pytest.mark.parametrize("chunk", [0, 7, 32])
def test_rbf_project_chunking_matches_tile_times_omega(chunk):
    # DArray::LRA's NN-chunk loop (dkernel.h:20-42) must equal K @ O for any chunk.
    rng = np.random.default_rng(2)
    Xi, Xj = rng.uniform(0, 1, (45, 5)), rng.uniform(0, 1, (38, 5))
    yi, yj = rng.choice([-1.0, 1.0], 45), rng.choice([-1.0, 1.0], 38)
    O = rng.standard_normal((38, 6))
    Y = orc.rbf_project(Xi, Xj, O, 0.7, yi, yj, chunk=chunk)
    K = orc.rbf_tile(Xi, Xj, 0.7, yi, yj)
    assert rel(Y, K @ O) < 1e-14


# -------------------------------------------------------------------- GEMM

@pytest.mark.parametrize("ta", ["N", "T"])
@pytest.mark.parametrize("tb", ["N", "T"])
def test_gemm_exact_small_integers(ta, tb):
    rng = np.random.default_rng(3)
    A = rng.integers(-4, 5, (9, 7)).astype(np.float64)
    B = rng.integers(-4, 5, (7, 5) if tb == "N" else (5, 7)).astype(np.float64)
    if ta == "T":
        A = A.T.copy()
    C0 = rng.integers(-4, 5, (9, 5)).astype(np.float64)
    opA = A.T if ta == "T" else A
    opB = B.T if tb == "T" else B
    out = orc.gemm(ta, tb, A, B, alpha=2.0, beta=-1.0, Cm=C0)
    np.testing.assert_array_equal(out, 2 * opA @ opB - C0)


def test_matrix_multiply_tn_blocking():
    rng = np.random.default_rng(4)
    A, B, C0 = rng.standard_normal((300, 20)), rng.standard_normal((300, 30)), rng.standard_normal((20, 30))
    out = orc.matrix_multiply_tn(A, B, alpha=-1.0, beta=1.0, Cm=C0, bs=128)
    assert rel(out, C0 - A.T @ B) < 1e-14


def test_syrk_upper_only():
    rng = np.random.default_rng(5)
    A = rng.standard_normal((17, 11))
    C0 = rng.standard_normal((11, 11))
    out = orc.syrk_ut(A, C0, alpha=-1.0, beta=1.0)
    full = C0 - A.T @ A
    iu = np.triu_indices(11)
    il = np.tril_indices(11, -1)
    assert rel(out[iu], full[iu]) < 1e-14
    np.testing.assert_array_equal(out[il], C0[il])


# ---------------------------------------------------------------- Cholesky

@pytest.mark.parametrize("n,bs", [(1000, 256), (300, 64), (257, 256)])
def test_potrf_min_ij_known_answer(n, bs):
    # perftest_cholesky.cpp:46-53,76-84: A(i,j) = min(i+1,j+1) = C C^T with C
    # lower-triangular ones, so U is all ones; the reference probes A(n-1,n-1)=1.
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    A = np.minimum(i + 1, j + 1).astype(np.float64)
    U, info = orc.potrf_upper(A, bs)
    assert info == 0
    assert abs(U[n - 1, n - 1] - 1) <= 1e-5  # the reference's own probe and tolerance
    np.testing.assert_allclose(np.triu(U), np.triu(np.ones((n, n))), rtol=0, atol=1e-9)
    np.testing.assert_array_equal(np.tril(U, -1), np.tril(A, -1))  # lower never written


def test_potrf_test3_known_answer():
    # test3.cpp:23-46: A = m m^T with m(i,j) = 1+i+j for i >= j (lower), so A = U^T U
    # with U = m^T (positive diagonal => the unique Cholesky factor).
    n = 10
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    m = np.where(i >= j, 1.0 + i + j, 0.0)
    A = m @ m.T
    for fn in (lambda a: orc.potrf_upper(a, 4), lambda a: orc.potrf_upper_recursive(a, 4)):
        U, info = fn(A)
        assert info == 0
        np.testing.assert_allclose(np.triu(U), m.T, rtol=1e-12, atol=1e-12)


def test_potrf_golden_and_variants():
    A = G["chol_A"]
    for fn in (lambda a: orc.potrf_upper(a, 16), lambda a: orc.potrf_upper(a, 256),
               lambda a: orc.potf2_upper(a), lambda a: orc.potrf_upper_recursive(a, 8)):
        U, info = fn(A)
        assert info == 0
        assert rel(np.triu(U), G["chol_U"]) < 1e-13


def test_potrf_not_positive_definite_reports_global_index():
    rng = np.random.default_rng(6)
    M = rng.standard_normal((40, 40))
    A = M.T @ M + 40 * np.eye(40)
    A[23, 23] = -5.0  # leading minor of order 24 is not positive definite
    _, info = orc.potrf_upper(A, 8)
    assert info == 24
    _, info2 = orc.potf2_upper(A)
    assert info2 == 24


# -------------------------------------------------------------------- TRSM

def test_trsm_golden():
    U, B, Br = G["trsm_U"], G["trsm_B"], G["trsm_Br"]
    assert rel(orc.trsm("L", "T", U, B, bs=8), G["trsm_LUT"]) < 1e-13
    assert rel(orc.trsm("L", "N", U, B, bs=8), G["trsm_LUN"]) < 1e-13
    assert rel(orc.trsm("R", "N", U, Br), G["trsm_RUN"]) < 1e-13


def test_trsm_exact_unit_triangular():
    # Integer unit-upper U and integer X: B = U^T X (or U X, X U) is exact,
    # and the solve must return X exactly (no rounding on small integers).
    rng = np.random.default_rng(7)
    U = np.triu(rng.integers(-2, 3, (12, 12)).astype(np.float64), 1) + np.eye(12)
    X = rng.integers(-3, 4, (12, 4)).astype(np.float64)
    np.testing.assert_array_equal(orc.trsm("L", "T", U, U.T @ X, diag="U", bs=4), X)
    np.testing.assert_array_equal(orc.trsm("L", "N", U, U @ X, diag="U", bs=4), X)
    Xr = X.T.copy()
    np.testing.assert_array_equal(orc.trsm("R", "N", U, Xr @ U, diag="U"), Xr)


def test_trsm_alpha():
    U, B = G["trsm_U"], G["trsm_B"]
    assert rel(orc.trsm("L", "T", U, B, alpha=-2.5), -2.5 * G["trsm_LUT"]) < 1e-13


# ------------------------------------------------------- orthogonalisation

def test_householder_qr():
    Y = G["qr_Y"]
    Q, R = orc.householder_qr(Y)
    assert np.linalg.norm(Q.T @ Q - np.eye(Y.shape[1])) < 1e-13
    assert rel(Q @ R, Y) < 1e-14
    np.testing.assert_array_equal(R, np.triu(R))
    s = np.sign(np.diag(R))
    assert rel(Q * s, G["qr_Q"]) < 1e-12
    assert rel(s[:, None] * R, G["qr_R"]) < 1e-12


def test_cholqr2():
    Y = G["qr_Y"]
    Q, R, info = orc.cholqr2(Y)
    assert info == 0
    assert np.all(np.diag(R) > 0)
    assert np.linalg.norm(Q.T @ Q - np.eye(Y.shape[1])) < 1e-14
    assert rel(Q @ R, Y) < 1e-14
    assert rel(Q, G["qr_Q"]) < 1e-12 and rel(R, G["qr_R"]) < 1e-12


def test_cholqr2_breakdown_rank_deficient():
    rng = np.random.default_rng(8)
    Y = rng.standard_normal((50, 6))
    Y[:, 4] = Y[:, 1]  # exactly dependent columns: Gram matrix singular
    _, _, info = orc.cholqr2(Y)
    assert info != 0


# --------------------------------------------------------------- pipeline

@pytest.mark.parametrize("orth", ["householder", "cholqr2"])
def test_lra_golden(orth):
    Q, W, info = orc.lra(G["lra_X"], G["lra_Omega"], float(G["lra_gamma"]), 1, orth=orth)
    assert info == 0
    assert rel(Q @ W.T, G["lra_QWt"]) < 1e-10
    e2, k2 = orc.lra_error(G["lra_X"], float(G["lra_gamma"]), Q, W)
    np.testing.assert_allclose([e2, k2], G["lra_err2"], rtol=1e-8, atol=1e-22)


def test_lra_exact_low_rank():
    # r distinct points => rank(K) <= r: with l >= r the range finder is exact.
    rng = np.random.default_rng(9)
    base = rng.uniform(0, 1, (5, 4))
    X = base[rng.integers(0, 5, 120)]
    Om = rng.standard_normal((120, 8))
    Q, W, info = orc.lra(X, Om, 0.5, 1, orth="householder")
    e2, k2 = orc.lra_error(X, 0.5, Q, W)
    assert np.sqrt(e2 / k2) < 1e-12


def test_lra_error_identity():
    # Q orthonormal => ||K - Q Q^T K||^2 = ||K||^2 - ||Q^T K||^2.
    rng = np.random.default_rng(10)
    X = rng.uniform(0, 1, (90, 3))
    Q, W, info = orc.lra(X, rng.standard_normal((90, 7)), 2.0, 0, orth="cholqr2")
    e2, k2 = orc.lra_error(X, 2.0, Q, W)
    np.testing.assert_allclose(e2, k2 - np.sum(W**2), rtol=1e-8)
    K = orc.rbf_tile(X, X, 2.0)
    np.testing.assert_allclose(k2, np.sum(K**2), rtol=1e-14)


def test_lra_householder_and_cholqr2_same_subspace():
    rng = np.random.default_rng(11)
    X = rng.uniform(0, 1, (200, 5))
    Om = rng.standard_normal((200, 12))
    Q1, W1, _ = orc.lra(X, Om, 0.4, 2, orth="householder")
    Q2, W2, _ = orc.lra(X, Om, 0.4, 2, orth="cholqr2")
    assert rel(Q1 @ Q1.T, Q2 @ Q2.T) < 1e-9
    assert rel(Q1 @ W1.T, Q2 @ W2.T) < 1e-9


# ---------------------------------------------------------------- Nystrom factor

@pytest.mark.parametrize("m,n", [(50, 50), (300, 272), (273, 273), (100, 3), (7, 1)])
def test_gesvj_vs_numpy_svd(m, n):
    rng = np.random.default_rng(m * 1000 + n)
    A = rng.standard_normal((m, n))
    AV, s, V, sweeps = orc.gesvj(A)
    assert sweeps > 0 or n == 1
    ref = np.linalg.svd(A, compute_uv=False)
    assert np.abs(s - ref).max() <= 1e-13 * ref[0]
    assert np.all(np.diff(s) <= 0)
    assert np.linalg.norm(V.T @ V - np.eye(n)) < 1e-12
    # A V = U diag(s) with orthonormal U: the columns of A V are orthogonal, of norm s
    G = AV.T @ AV
    assert np.abs(G - np.diag(s ** 2)).max() <= 1e-12 * s[0] ** 2
    assert rel(AV, A @ V) < 1e-13


def test_gesvj_closed_forms():
    # A diagonal matrix needs no rotation: exact |d| in descending order, V a permutation.
    d = np.array([3.0, -7.0, 0.5, 0.0, 2.0])
    AV, s, V, sweeps = orc.gesvj(np.diag(d))
    assert sweeps == 1 and list(s) == [7.0, 3.0, 2.0, 0.5, 0.0]
    assert np.array_equal(np.abs(V), np.abs(V).round())
    # [[a, b], [b, a]]: eigenvalues a + b, a - b with vectors (1, 1)/sqrt2, (1, -1)/sqrt2.
    AV, s, V, _ = orc.gesvj(np.array([[2.0, 1.0], [1.0, 2.0]]))
    np.testing.assert_allclose(s, [3.0, 1.0], rtol=1e-15)
    np.testing.assert_allclose(np.abs(V), np.full((2, 2), np.sqrt(0.5)), rtol=1e-15)


def _nystrom_numpy(Q, W, k):
    C = Q.T @ W
    C = (C + C.T) / 2
    w, U = np.linalg.eigh(C)
    idx = np.argsort(w)[::-1][:k]
    return Q @ (U[:, idx] * np.sqrt(np.maximum(w[idx], 0))), w[idx]


@pytest.mark.parametrize("labels", [False, True])
def test_nystrom_vs_numpy_eigh(labels):
    rng = np.random.default_rng(40)
    n, d, l, k = 600, 8, 50, 40
    X = rng.uniform(-1, 1, (n, d))
    y = rng.choice([-1.0, 1.0], n) if labels else None
    Om = rng.standard_normal((n, l))
    Q, W, info = orc.lra(X, Om, 0.5, 1, y)
    assert info == 0
    Z, lam, st = orc.nystrom(Q, W, k)
    assert st == 0
    Zr, lr = _nystrom_numpy(Q, W, k)
    assert np.abs(lam - lr).max() <= 1e-12 * lr[0]
    assert rel(Z @ Z.T, Zr @ Zr.T) < 1e-12
    # K - Z Z^T against the rank-k truncation of Q Q^T K Q Q^T: no better than the range finder
    e2, k2 = orc.lra_error(X, 0.5, Z, Z, y)
    q2, _ = orc.lra_error(X, 0.5, Q, W, y)
    assert q2 <= e2 * (1 + 1e-12)


def test_nystrom_exact_for_low_rank_kernel():
    # r distinct points: rank(K) <= r <= k, so Z Z^T reproduces K to rounding.
    rng = np.random.default_rng(41)
    r, n, d = 12, 900, 5
    P = rng.uniform(-1, 1, (r, d))
    X = P[np.arange(n) % r]
    Om = rng.standard_normal((n, 20))
    # l = 20 > rank: Householder orthogonalisation (CholeskyQR breaks down here)
    Q, W, info = orc.lra(X, Om, 0.7, 1, orth="householder")
    assert info == 0
    Z, lam, st = orc.nystrom(Q, W, 16)
    e2, k2 = orc.lra_error(X, 0.7, Z, Z)
    assert st == 0 and np.sqrt(e2 / k2) < 1e-12
    assert np.all(lam[r:] < 1e-10 * lam[0])  # eigenvalues past the rank vanish


# ---------------------------------------------------------------- barrier SVM solver

def _svm_problem(n, d, k, seed):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-1, 1, (n, d))
    y = np.where(X[:, 0] + 0.3 * rng.standard_normal(n) > 0, 1.0, -1.0)
    K = np.exp(-0.5 * ((X[:, None, :] - X[None, :, :]) ** 2).sum(-1))
    w, U = np.linalg.eigh((y[:, None] * y[None, :]) * K)
    idx = np.argsort(w)[::-1][:k]
    return np.asfortranarray(U[:, idx] * np.sqrt(np.maximum(w[idx], 0))), y


@pytest.mark.parametrize("n,d,k,Cb", [(300, 4, 40, 1.0), (800, 8, 60, 10.0)])
def test_svm_barrier_reaches_the_box_qp_optimum(n, d, k, Cb):
    """The barrier method stops when the gap bound 2n/t is below 1e-5 |f|, so
    f(a) lies within 1e-5 |f*| of the box-constrained optimum, computed here
    independently by L-BFGS-B; a stays strictly inside (0, C)."""
    from scipy.optimize import minimize
    Z, y = _svm_problem(n, d, k, n)
    a, stats, fval, st = orc.svm_barrier(Z, y, Cb)
    assert st == 0 and stats[0] >= 2 and stats[2] > 0
    G = Z @ Z.T
    assert abs(fval - (0.5 * a @ G @ a - a.sum())) <= 1e-10 * abs(fval)
    res = minimize(lambda x: 0.5 * x @ G @ x - x.sum(), np.full(n, Cb / 2), jac=lambda x: G @ x - 1,
                   method="L-BFGS-B", bounds=[(0, Cb)] * n,
                   options=dict(ftol=1e-15, gtol=1e-12, maxiter=100000))
    assert res.fun <= fval <= res.fun + 1e-5 * abs(res.fun)
    assert np.all(a > 0) and np.all(a < Cb)


def test_svm_barrier_start_point_and_empty_class():
    # One Newton step with no CG iteration (s = 0): a stays at the start point,
    # y^T a = 0 with the smaller class at 0.9 C (testsvm.cpp:204-219).
    Z, y = _svm_problem(200, 3, 20, 3)
    a, stats, _, st = orc.svm_barrier(Z, y, 2.0, max_outer=1, max_newton=1, max_cg=0)
    assert st == 0 and stats == [1, 1, 0]
    small = -1.0 if (y > 0).sum() > (y < 0).sum() else 1.0
    assert np.all(a[y == small] == 0.9 * 2.0) and abs(y @ a) < 1e-9
    _, _, _, st = orc.svm_barrier(Z, np.ones(200), 2.0)
    assert st == -1
