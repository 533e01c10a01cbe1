"""HIP path vs the FP64 oracle and the golden vectors, through the C ABI.

Tolerances are relative Frobenius errors written per test: FP64 paths must
agree with the oracle to <= 1e-10 (the path's stated bar; most are far
tighter), FP32 paths to FP32 rounding."""
from pathlib import Path

import numpy as np
import pytest

import darray_optimizer_amd as dh
import oracle as orc

pytestmark = pytest.mark.gpu
G = np.load(Path(__file__).resolve().parent / "golden" / "golden.npz")


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def dev_pts(X, dtype=None):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(X))
    return (t.to(dtype) if dtype else t).cuda()


def dev_vec(v, dtype=None):
    import torch
    if v is None:
        return None
    t = torch.from_numpy(np.ascontiguousarray(v))
    return (t.to(dtype) if dtype else t).cuda()


def dev_mat(A, dtype=None):
    return dh.to_dev(np.asarray(A, dtype=np.float64), dtype)


def empty_mat(m, n, dtype):
    import torch
    return torch.zeros((n, m), dtype=dtype, device="cuda")


# ---------------------------------------------------------------- RBF tile

@pytest.mark.parametrize("m,n,d,labels", [(1, 1, 1, False), (64, 64, 16, True),
                                          (100, 77, 13, True), (130, 257, 130, False)])
def test_rbf_tile_f64(dev, m, n, d, labels):
    import torch
    rng = np.random.default_rng(m * 1000 + n)
    Xi, Xj = rng.uniform(0, 1, (m, d)), rng.uniform(0, 1, (n, d))
    yi = rng.choice([-1.0, 1.0], m) if labels else None
    yj = rng.choice([-1.0, 1.0], n) if labels else None
    K = empty_mat(m, n, torch.float64)
    dev.rbf_tile(dev_pts(Xi), dev_pts(Xj), 1.0 / d, K, dev_vec(yi), dev_vec(yj))
    ref = orc.rbf_tile(Xi, Xj, 1.0 / d, yi, yj, clamp=True)
    assert rel(dh.from_dev(K), ref) < 1e-13


def test_rbf_tile_f32(dev):
    import torch
    rng = np.random.default_rng(5)
    Xi, Xj = rng.uniform(0, 1, (300, 64)), rng.uniform(0, 1, (200, 64))
    K = empty_mat(300, 200, torch.float32)
    dev.rbf_tile(dev_pts(Xi, torch.float32), dev_pts(Xj, torch.float32), 1 / 64.0, K)
    ref = orc.rbf_tile(Xi.astype(np.float32), Xj.astype(np.float32), 1 / 64.0, clamp=True)
    assert rel(dh.from_dev(K), ref) < 2e-6


def test_rbf_tile_golden(dev):
    import torch
    m, n = G["tile_Xi"].shape[0], G["tile_Xj"].shape[0]
    K = empty_mat(m, n, torch.float64)
    dev.rbf_tile(dev_pts(G["tile_Xi"]), dev_pts(G["tile_Xj"]), float(G["tile_gamma"]), K,
                 dev_vec(G["tile_yi"]), dev_vec(G["tile_yj"]))
    assert rel(dh.from_dev(K), G["tile_K"]) < 1e-14


# --------------------------------------------------------- fused projection

PROJ_SHAPES = [(1, 1, 1, 1), (64, 64, 64, 16), (100, 77, 13, 9), (129, 65, 33, 40),
               (300, 257, 130, 138), (200, 190, 200, 300), (96, 500, 128, 272),
               (150, 300, 257, 40), (70, 130, 600, 17)]  # d > 256: materialised K blocks


@pytest.mark.parametrize("m,n,d,l", PROJ_SHAPES)
def test_rbf_project_f64(dev, m, n, d, l):
    import torch
    rng = np.random.default_rng(m + 7 * n + 13 * d + l)
    Xi, Xj = rng.uniform(0, 1, (m, d)), rng.uniform(0, 1, (n, d))
    yi, yj = rng.choice([-1.0, 1.0], m), rng.choice([-1.0, 1.0], n)
    O = rng.standard_normal((n, l))
    Y0 = rng.standard_normal((m, l))
    Y = dev_mat(Y0)
    dev.rbf_project_f64(dev_pts(Xi), dev_pts(Xj), dev_mat(O), 1.0 / d, Y, dev_vec(yi),
                        dev_vec(yj), beta=0.5)
    ref = orc.rbf_project(Xi, Xj, O, 1.0 / d, yi, yj, beta=0.5, Y=Y0)
    assert rel(dh.from_dev(Y), ref) < 1e-12


def test_rbf_project_golden(dev):
    X = dev_pts(G["proj_X"])
    Y = dev_mat(G["proj_Y0"])
    dev.rbf_project_f64(X, X, dev_mat(G["proj_O"]), float(G["proj_gamma"]), Y,
                        beta=float(G["proj_beta"]))
    assert rel(dh.from_dev(Y), G["proj_Y"]) < 1e-14


# -------------------------------------------------------------------- GEMM

GEMM_SHAPES = [(1, 1, 1), (40, 40, 3000), (257, 129, 65), (2100, 2050, 40), (17, 300, 0)]


@pytest.mark.parametrize("ta", ["N", "T"])
@pytest.mark.parametrize("tb", ["N", "T"])
@pytest.mark.parametrize("m,n,k", GEMM_SHAPES)
def test_gemm_f64(dev, ta, tb, m, n, k):
    import torch
    rng = np.random.default_rng(m + n + k)
    A = rng.standard_normal((k, m) if ta == "T" else (m, k))
    B = rng.standard_normal((n, k) if tb == "T" else (k, n))
    C0 = rng.standard_normal((m, n))
    Cd = dev_mat(C0)
    dev.gemm(ta, tb, dev_mat(A), dev_mat(B), Cd, alpha=-1.5, beta=0.25)
    opA = A.T if ta == "T" else A
    opB = B.T if tb == "T" else B
    assert rel(dh.from_dev(Cd), -1.5 * opA @ opB + 0.25 * C0) < 1e-13


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T")])
def test_gemm_f32(dev, ta, tb):
    import torch
    rng = np.random.default_rng(11)
    m, n, k = 300, 200, 1500
    A = rng.standard_normal((k, m) if ta == "T" else (m, k)).astype(np.float32)
    B = rng.standard_normal((n, k) if tb == "T" else (k, n)).astype(np.float32)
    Cd = empty_mat(m, n, torch.float32)
    dev.gemm(ta, tb, dh.to_dev(A), dh.to_dev(B), Cd)
    opA = (A.T if ta == "T" else A).astype(np.float64)
    opB = (B.T if tb == "T" else B).astype(np.float64)
    assert rel(dh.from_dev(Cd), opA @ opB) < 1e-5


def test_syrk_upper_only(dev):
    rng = np.random.default_rng(12)
    A = rng.standard_normal((300, 150))
    C0 = rng.standard_normal((150, 150))
    Cd = dev_mat(C0)
    dev.syrk_ut(dev_mat(A), Cd, alpha=-1.0, beta=1.0)
    out = dh.from_dev(Cd)
    iu, il = np.triu_indices(150), np.tril_indices(150, -1)
    assert rel(out[iu], (C0 - A.T @ A)[iu]) < 1e-13
    np.testing.assert_array_equal(out[il], C0[il])


# -------------------------------------------------------------------- TRSM

@pytest.mark.parametrize("m", [45, 130, 300])
@pytest.mark.parametrize("side,trans", [("L", "T"), ("L", "N"), ("R", "N")])
def test_trsm_f64(dev, m, side, trans):
    rng = np.random.default_rng(m)
    U = np.triu(rng.standard_normal((m, m))) + 8 * np.eye(m)
    U_full = U + np.tril(rng.standard_normal((m, m)), -1)  # lower garbage must be ignored
    B = rng.standard_normal((m, 37)) if side == "L" else rng.standard_normal((53, m))
    Bd = dev_mat(B)
    dev.trsm(side, trans, "N", dev_mat(U_full), Bd, alpha=-0.5)
    ref = orc.trsm(side, trans, U, B, alpha=-0.5)
    assert rel(dh.from_dev(Bd), ref) < 1e-12


@pytest.mark.parametrize("side,trans,m,nrhs", [("L", "T", 700, 64), ("L", "N", 700, 64),
                                                ("L", "T", 300, 600), ("L", "N", 130, 513)])
def test_trsm_f64_block_sweep_and_recursion(dev, side, trans, m, nrhs):
    """Side-L solves with up to 512 right-hand sides run as a sweep over 64-row
    blocks; wider ones take the recursive halving.  Both against the oracle."""
    rng = np.random.default_rng(m + nrhs)
    U = np.triu(rng.standard_normal((m, m))) + 8 * np.eye(m)
    B = rng.standard_normal((m, nrhs))
    Bd = dev_mat(B)
    dev.trsm(side, trans, "N", dev_mat(U), Bd)
    assert rel(dh.from_dev(Bd), orc.trsm(side, trans, U, B)) < 1e-12


def test_trsm_golden(dev):
    U = G["trsm_U"]
    for side, trans, Bk, ref in (("L", "T", "trsm_B", "trsm_LUT"), ("L", "N", "trsm_B", "trsm_LUN"),
                                 ("R", "N", "trsm_Br", "trsm_RUN")):
        Bd = dev_mat(G[Bk])
        dev.trsm(side, trans, "N", dev_mat(U), Bd)
        assert rel(dh.from_dev(Bd), G[ref]) < 1e-13


def test_trsm_f32(dev):
    import torch
    rng = np.random.default_rng(13)
    U = (np.triu(rng.standard_normal((200, 200))) + 8 * np.eye(200)).astype(np.float32)
    B = rng.standard_normal((200, 64)).astype(np.float32)
    Bd = dh.to_dev(B)
    dev.trsm("L", "T", "N", dh.to_dev(U), Bd)
    ref = orc.trsm("L", "T", U.astype(np.float64), B.astype(np.float64))
    assert rel(dh.from_dev(Bd), ref) < 1e-5


# ---------------------------------------------------------------- Cholesky

@pytest.mark.parametrize("n,bs", [(1000, 256), (300, 64), (257, 256)])
def test_potrf_f64_min_ij_known_answer(dev, n, bs):
    # perftest_cholesky.cpp:46-53,76-84: U must be all ones; probe A(n-1,n-1) = 1.
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    A = np.minimum(i + 1, j + 1).astype(np.float64)
    Ad = dev_mat(A)
    assert dev.potrf(Ad, bs) == 0
    U = dh.from_dev(Ad)
    assert abs(U[n - 1, n - 1] - 1) <= 1e-5
    np.testing.assert_allclose(np.triu(U), np.triu(np.ones((n, n))), rtol=0, atol=1e-9)
    np.testing.assert_array_equal(np.tril(U, -1), np.tril(A, -1))


def test_potrf_f32_min_ij_probe(dev):
    import torch
    n = 512
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    A = np.minimum(i + 1, j + 1).astype(np.float32)
    Ad = dh.to_dev(A)
    assert dev.potrf(Ad, 256) == 0
    U = dh.from_dev(Ad)
    assert abs(U[n - 1, n - 1] - 1) <= 1e-5  # the reference's FP32 probe tolerance


def test_potrf_f64_vs_oracle_and_golden(dev):
    Ad = dev_mat(G["chol_A"])
    assert dev.potrf(Ad, 32) == 0
    assert rel(np.triu(dh.from_dev(Ad)), G["chol_U"]) < 1e-13
    rng = np.random.default_rng(14)
    M = rng.standard_normal((700, 700))
    A = M.T @ M / 700 + np.eye(700)
    Ad = dev_mat(A)
    assert dev.potrf(Ad, 256) == 0
    Uo, info = orc.potrf_upper(A, 256)
    assert info == 0
    assert rel(np.triu(dh.from_dev(Ad)), np.triu(Uo)) < 1e-12


def test_potrf_test3_known_answer(dev):
    n = 10
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    m = np.where(i >= j, 1.0 + i + j, 0.0)
    Ad = dev_mat(m @ m.T)
    assert dev.potrf(Ad, 4) == 0
    np.testing.assert_allclose(np.triu(dh.from_dev(Ad)), m.T, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("bs", [64, 256])
@pytest.mark.parametrize("bad", [0, 23, 200, 299])
def test_potrf_reports_first_bad_minor(dev, bad, bs):
    # bs = 256: the FP64 diagonal block of order 256 is the register
    # factorisation; the last 44 columns and every bs = 64 block are not.
    rng = np.random.default_rng(15)
    M = rng.standard_normal((300, 300))
    A = M.T @ M + 300 * np.eye(300)
    A[bad, bad] = -1.0
    Ad = dev_mat(A)
    assert dev.potrf(Ad, bs) == bad + 1


def test_potrf_lookahead_orders_with_the_handle_stream(dev):
    """The look-ahead runs part of each trailing update on a second stream:
    it must start after work queued before the call on the handle's stream
    (here the copy that fills A) and finish before the call returns."""
    import torch
    rng = np.random.default_rng(16)
    n = 1500
    M = rng.standard_normal((n, n))
    A = M.T @ M / n + np.eye(n)
    src = dev_mat(A)
    side = torch.cuda.Stream()
    h = dh.Handle(0, stream=side)
    try:
        with torch.cuda.stream(side):
            Ad = torch.full_like(src, float("nan"))
            torch.cuda._sleep(2_000_000)  # keep the copy below queued behind other work
            Ad.copy_(src)
        assert h.potrf(Ad, 256) == 0
        side.synchronize()
        Uo, info = orc.potrf_upper(A, 256)
        assert info == 0
        assert rel(np.triu(dh.from_dev(Ad)), np.triu(Uo)) < 1e-12
    finally:
        h.close()


# ------------------------------------------------------------- CholeskyQR2

def test_cholqr2_golden(dev):
    Yd = dev_mat(G["qr_Y"])
    R = empty_mat(12, 12, __import__("torch").float64)
    assert dev.cholqr2(Yd, R) == 0
    assert rel(dh.from_dev(Yd), G["qr_Q"]) < 1e-12
    assert rel(dh.from_dev(R), G["qr_R"]) < 1e-12


@pytest.mark.parametrize("m,n", [(20000, 138), (3000, 1), (3000, 257), (5000, 272), (6000, 544),
                                 (5000, 300), (2503, 417), (4000, 600), (131, 130), (4099, 200),
                                 (70001, 129)])
def test_cholqr2_tall(dev, m, n):
    # n <= 272: register-resident Cholesky (257: one real column in the last
    # 16-block); n <= 544: one-workgroup Cholesky and the row solve as two
    # register solves around a GEMM (300, 417: ragged second block); n = 600:
    # blocked fallback.
    # 128 < n <= 272 also takes the register-blocked Gram: (131, 130) one slab
    # per 16 rows with a ragged last chunk, (70001, 129) ragged row slabs.
    import torch
    rng = np.random.default_rng(16)
    Y = rng.standard_normal((m, n)) @ np.diag(np.logspace(0, -4, n))
    Yd = dev_mat(Y)
    R = empty_mat(n, n, torch.float64)
    assert dev.cholqr2(Yd, R) == 0
    Q, Rh = dh.from_dev(Yd), dh.from_dev(R)
    assert np.linalg.norm(Q.T @ Q - np.eye(n)) < 1e-13
    assert rel(Q @ Rh, Y) < 1e-13
    np.testing.assert_array_equal(np.tril(Rh, -1), 0)
    _, Ro, info = orc.cholqr2(Y)
    assert info == 0 and rel(Rh, Ro) < 1e-10


@pytest.mark.parametrize("n", [10, 250, 400, 600])
def test_cholqr2_breakdown(dev, n):
    # Exactly dependent columns: every CholeskyQR variant must report breakdown
    # (for n = 250 and 400 in a late panel).
    rng = np.random.default_rng(17)
    Y = rng.standard_normal((2000, n))
    Y[:, (3 * n) // 4] = Y[:, 2]
    assert dev.cholqr2(dev_mat(Y)) > 0


@pytest.mark.parametrize("n", [30, 400])
def test_cholqr_shifted_fallback_ill_conditioned(dev, n):
    # kappa(Y) = 1e11 is beyond CholeskyQR2 (~u^{-1/2}); the shifted
    # CholeskyQR3 fallback must still return an orthonormal Q with QR = Y
    # (n = 400: the shifted Cholesky as two order-272 register blocks).
    import torch
    rng = np.random.default_rng(18)
    U, _ = np.linalg.qr(rng.standard_normal((4000, n)))
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    Y = U @ np.diag(np.logspace(0, -11, n)) @ V.T
    Yd = dev_mat(Y)
    R = empty_mat(n, n, torch.float64)
    assert dev.cholqr2(Yd, R) == 0
    Q, Rh = dh.from_dev(Yd), dh.from_dev(R)
    assert np.linalg.norm(Q.T @ Q - np.eye(n)) < 1e-12
    assert rel(Q @ Rh, Y) < 1e-13


# ------------------------------------------------------------------ LRA

@pytest.mark.parametrize("q", [0, 1, 2])
@pytest.mark.parametrize("labels", [False, True])
def test_lra_f64_vs_oracle(dev, labels, q):
    # q = 2: the power iterations alternate the basis between Q and W and end in Q.
    import torch
    n, d, l = 1536, 64, 40
    X = dh.fill_uniform(0, d, n).T.copy()  # DArray set_uniform layout, points as rows
    Om = dh.fill_normal(1, n, l)
    y = np.where(np.arange(n) % 3 == 0, -1.0, 1.0) if labels else None
    Q = empty_mat(n, l, torch.float64)
    W = empty_mat(n, l, torch.float64)
    assert dev.lra(dev_pts(X), dev_mat(Om), 1.0 / d, q, Q, W, dev_vec(y)) == 0
    Qo, Wo, info = orc.lra(X, Om, 1.0 / d, q, y, orth="householder")
    assert info == 0
    Qg, Wg = dh.from_dev(Q), dh.from_dev(W)
    assert rel(Qg @ Wg.T, Qo @ Wo.T) < 1e-10  # QB with B = W^T, sign invariant
    assert rel(Qg @ Qg.T, Qo @ Qo.T) < 1e-10
    e2, k2 = dev.lra_error(dev_pts(X), 1.0 / d, Q, W, dev_vec(y))
    eo, ko = orc.lra_error(X, 1.0 / d, Qo, Wo, y)
    assert abs(k2 - ko) / ko < 1e-12
    assert abs(np.sqrt(e2 / k2) - np.sqrt(eo / ko)) <= 1e-10 + 1e-6 * np.sqrt(eo / ko)


def test_lra_above_256_features(dev):
    """Above 256 features the passes materialise K in row blocks (FP64); every
    precision runs them, so the low-precision requests return the FP64 result
    bit for bit."""
    import torch
    n, d, l = 700, 300, 24
    X = dh.fill_uniform(3, d, n).T.copy()
    Om = dh.fill_normal(4, n, l)
    y = np.where(np.arange(n) % 4 == 0, -1.0, 1.0)
    out = {}
    for prec in (dh.DH_PREC_F64, dh.DH_PREC_BF16, dh.DH_PREC_F16, dh.DH_PREC_FP8):
        Q = empty_mat(n, l, torch.float64)
        W = empty_mat(n, l, torch.float64)
        assert dev.lra(dev_pts(X), dev_mat(Om), 1.0 / d, 1, Q, W, dev_vec(y), prec=prec) == 0
        out[prec] = (dh.from_dev(Q), dh.from_dev(W))
    Qo, Wo, info = orc.lra(X, Om, 1.0 / d, 1, y, orth="householder")
    assert info == 0
    Qg, Wg = out[dh.DH_PREC_F64]
    assert rel(Qg @ Wg.T, Qo @ Wo.T) < 1e-10
    for prec in (dh.DH_PREC_BF16, dh.DH_PREC_F16, dh.DH_PREC_FP8):
        assert np.array_equal(out[prec][0], Qg) and np.array_equal(out[prec][1], Wg)


def test_lra_golden(dev):
    import torch
    X = G["lra_X"]
    n, l = X.shape[0], G["lra_Omega"].shape[1]
    Q = empty_mat(n, l, torch.float64)
    W = empty_mat(n, l, torch.float64)
    assert dev.lra(dev_pts(X), dev_mat(G["lra_Omega"]), float(G["lra_gamma"]), 1, Q, W) == 0
    Qg, Wg = dh.from_dev(Q), dh.from_dev(W)
    assert rel(Qg @ Wg.T, G["lra_QWt"]) < 1e-10


def test_lra_error_golden(dev):
    import torch
    X = G["lra_X"]
    Qo, Wo, _ = orc.lra(X, G["lra_Omega"], float(G["lra_gamma"]), 1)
    e2, k2 = dev.lra_error(dev_pts(X), float(G["lra_gamma"]), dev_mat(Qo), dev_mat(Wo))
    np.testing.assert_allclose([e2, k2], G["lra_err2"], rtol=1e-8, atol=1e-22)


def test_lra_f64_full_size_properties(dev):
    """n = 65536, d = 128, l = 272, q = 2 (the FP64 GPU configuration): the oracle
    cannot run here, so check size-independent properties."""
    import torch
    n, d, k, p, q = 65536, 128, 256, 16, 2
    l = k + p
    X = dh.fill_mixture(0, d, n, 16, 0.25)
    Om = dh.fill_normal(1, n, l)
    Xd = dev_pts(X)
    Q = empty_mat(n, l, torch.float64)
    W = empty_mat(n, l, torch.float64)
    assert dev.lra(Xd, dev_mat(Om), 0.5 / d, q, Q, W) == 0
    QtQ = (Q @ Q.T).cpu().numpy()
    assert np.linalg.norm(QtQ - np.eye(l)) < 1e-12
    e2, k2 = dev.lra_error(Xd, 0.5 / d, Q, W)
    w2 = float((W * W).sum())
    # Q orthonormal: ||K - QQ^T K||^2 = ||K||^2 - ||Q^T K||^2
    assert abs(e2 - (k2 - w2)) <= 1e-9 * k2
    assert 0 <= np.sqrt(e2 / k2) < 0.5


# ------------------------------------------------------------- core solve

def _rbf_core(n, d, seed=21, shift=1e-3):
    X = dh.fill_mixture(seed, d, n, 16, 0.25)
    return orc.rbf_tile(X, X, 0.5 / d) + shift * np.eye(n)


@pytest.mark.parametrize("n,nrhs,bs", [(700, 13, 256), (1000, 64, 128)])
def test_posv_f64_vs_oracle(dev, n, nrhs, bs):
    rng = np.random.default_rng(n)
    M = rng.standard_normal((n, n))
    A = M.T @ M / n + np.eye(n)
    B = rng.standard_normal((n, nrhs))
    Ad, Bd = dev_mat(A), dev_mat(B)
    assert dev.posv(Ad, Bd, bs) == 0
    U, info = orc.potrf_upper(A, bs)
    Xo = orc.trsm("L", "N", np.triu(U), orc.trsm("L", "T", np.triu(U), B))
    assert info == 0 and rel(dh.from_dev(Bd), Xo) < 1e-10
    assert rel(np.triu(dh.from_dev(Ad)), np.triu(U)) < 1e-12


def test_posv_f64_reports_failure(dev):
    A = np.eye(50)
    A[30, 30] = -1.0
    assert dev.posv(dev_mat(A), dev_mat(np.ones((50, 2))), 16) == 31


def test_posv_mixed_rbf_core(dev):
    """FP32 Cholesky + FP64 refinement on an RBF core + 1e-3 I (kappa ~ 1e6)."""
    import torch
    n, nrhs = 2048, 16
    A = _rbf_core(n, 32)
    rng = np.random.default_rng(22)
    B = rng.standard_normal((n, nrhs))
    X = torch.zeros((nrhs, n), dtype=torch.float64, device="cuda")
    info, iters, berr = dev.posv_mixed(dev_mat(np.triu(A)), dev_mat(B), X, 256, 1e-10, 100)
    assert info == 0 and berr <= 1e-10, (iters, berr)
    Xh = dh.from_dev(X)
    R = B - A @ Xh
    anorm = np.linalg.norm(A)
    be = max(np.linalg.norm(R[:, j]) / (anorm * np.linalg.norm(Xh[:, j]) + np.linalg.norm(B[:, j]))
             for j in range(nrhs))
    assert be <= 1e-10
    Xref = np.linalg.solve(A, B)
    assert rel(Xh, Xref) <= 1e-10 * np.linalg.cond(A)


def test_posv_mixed_falls_back_to_fp64(dev):
    # Eigenvalues down to 1e-10 relative: FP32 Cholesky breaks down, FP64 does not.
    import torch
    rng = np.random.default_rng(23)
    n = 300
    Qm, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = (Qm * np.logspace(0, -10, n)) @ Qm.T
    A = (A + A.T) / 2
    B = rng.standard_normal((n, 3))
    X = torch.zeros((3, n), dtype=torch.float64, device="cuda")
    info, iters, berr = dev.posv_mixed(dev_mat(A), dev_mat(B), X, 64, 1e-10, 30)
    assert info == 0 and iters == -1
    assert rel(A @ dh.from_dev(X), B) < 1e-5


@pytest.mark.parametrize("fn", ["potrf", "posv"])
def test_lookahead_cholesky_with_split_k_updates(dev, fn):
    """bs = 1024 at n = 4096: the trailing updates have k = 1024 and few output
    tiles, so they split k on both the handle's stream and the look-ahead
    stream, which run concurrently with separate partials buffers."""
    import os
    n, bs = 4096, 1024
    rng = np.random.default_rng(31)
    M = rng.standard_normal((n, n))
    A = M.T @ M / n + np.eye(n)
    orc.set_threads(min(16, os.cpu_count() or 1))
    try:
        U, info = orc.potrf_upper(A, bs)
    finally:
        orc.set_threads(1)
    assert info == 0
    Ad = dev_mat(A)
    if fn == "potrf":
        assert dev.potrf(Ad, bs) == 0
    else:
        B = rng.standard_normal((n, 8))
        Bd = dev_mat(B)
        assert dev.posv(Ad, Bd, bs) == 0
        assert rel(A @ dh.from_dev(Bd), B) < 1e-10
    e = rel(np.triu(dh.from_dev(Ad)), np.triu(U))
    print(f"{fn} n={n} bs={bs}: factor vs oracle {e:.2e}", flush=True)
    assert e < 1e-12


def test_posv_mixed_max_iter_reached_falls_back(dev):
    """A tolerance the FP32 refinement cannot reach within one correction: the
    solve is redone in FP64 (iters = -1) and the returned X is the FP64 one."""
    import torch
    n = 400
    A = _rbf_core(n, 16)
    B = np.random.default_rng(24).standard_normal((n, 4))
    X = torch.zeros((4, n), dtype=torch.float64, device="cuda")
    info, iters, berr = dev.posv_mixed(dev_mat(np.triu(A)), dev_mat(B), X, 64, 1e-15, 1)
    assert info == 0 and iters == -1
    assert berr <= 1e3 * n * 2.0**-53
    assert rel(dh.from_dev(X), np.linalg.solve(A, B)) < 1e-6


def test_posv_mixed_diverging_refinement_falls_back(dev):
    """kappa = 1e9: u32 * kappa > 1, so FP32 refinement cannot converge (or the
    FP32 factorisation breaks down); either way the FP64 path answers."""
    import torch
    rng = np.random.default_rng(25)
    n = 300
    Qm, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = (Qm * np.logspace(0, -9, n)) @ Qm.T
    A = (A + A.T) / 2
    B = rng.standard_normal((n, 2))
    X = torch.zeros((2, n), dtype=torch.float64, device="cuda")
    info, iters, berr = dev.posv_mixed(dev_mat(A), dev_mat(B), X, 64, 1e-10, 30)
    print(f"kappa 1e9: info {info} iters {iters} berr {berr:.2e}", flush=True)
    assert info == 0 and iters == -1 and berr <= 1e-10


def test_posv_mixed_nan_rhs_is_reported(dev):
    """A NaN in B gives a non-finite backward error: never reported as converged."""
    import torch
    n = 128
    A = _rbf_core(n, 8)
    B = np.ones((n, 2))
    B[5, 1] = np.nan
    X = torch.zeros((2, n), dtype=torch.float64, device="cuda")
    with pytest.raises(dh.DHError) as ei:
        dev.posv_mixed(dev_mat(np.triu(A)), dev_mat(B), X, 64, 1e-10, 10)
    assert ei.value.status == -1003
