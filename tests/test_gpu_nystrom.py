"""Nystrom factor (dh_nystrom_f64) and the one-sided Jacobi SVD under it
(dh_gesvj_f64) against the oracle (orc.gesvj / orc.nystrom, pinned to numpy's
SVD and eigh) and through size-independent properties at the C1 size.

Bars: singular values and eigenvalues to 1e-12 of the largest; Z Z^T (sign-
and rotation-invariant) to 1e-11 relative Frobenius; Z^T Z = diag(lam) to
1e-12 of lam_1 (Q and V_k orthonormal)."""
import os

import numpy as np
import pytest

import darray_optimizer_amd as dh
import oracle as orc

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("m,n", [(300, 272), (273, 273), (100, 3), (7, 1), (64, 64), (600, 544)])
def test_gesvj_vs_oracle(dev, m, n):
    import torch
    rng = np.random.default_rng(m + 7 * n)
    A = rng.standard_normal((m, n))
    Ad = dh.to_dev(A)
    s = torch.zeros(n, dtype=torch.float64, device="cuda")
    V = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    st, sweeps = dev.gesvj(Ad, s, V)
    AV, so, Vo, swo = orc.gesvj(A)
    sg, Vg, AVg = s.cpu().numpy(), dh.from_dev(V), dh.from_dev(Ad)
    print(f"gesvj {m}x{n}: sweeps device {sweeps} oracle {swo}; "
          f"|s - s_oracle| / s_1 = {np.abs(sg - so).max() / so[0]:.2e}", flush=True)
    assert st == 0 and sweeps > 0 or n == 1
    assert np.abs(sg - so).max() <= 1e-12 * so[0]
    assert np.abs(sg - np.linalg.svd(A, compute_uv=False)).max() <= 1e-12 * so[0]
    assert np.linalg.norm(Vg.T @ Vg - np.eye(n)) < 1e-11
    assert rel(AVg, A @ Vg) < 1e-12


def test_gesvj_rank_deficient_and_zero_columns(dev):
    import torch
    rng = np.random.default_rng(3)
    B = rng.standard_normal((200, 5))
    A = np.hstack([B @ rng.standard_normal((5, 30)), np.zeros((200, 2))])  # rank 5
    n = A.shape[1]
    Ad = dh.to_dev(A)
    s = torch.zeros(n, dtype=torch.float64, device="cuda")
    V = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    st, _ = dev.gesvj(Ad, s, V)
    sg = s.cpu().numpy()
    ref = np.linalg.svd(A, compute_uv=False)
    assert st == 0
    assert np.abs(sg - ref).max() <= 1e-12 * ref[0]
    assert np.all(sg[5:] <= 1e-12 * sg[0])


@pytest.mark.parametrize("labels", [False, True])
def test_nystrom_vs_oracle(dev, labels):
    import torch
    orc.set_threads(min(16, os.cpu_count() or 1))
    try:
        n, d, l, k = 3000, 16, 74, 64
        X, y = dh.fill_mixture(51, d, n, 4, 0.3, labels=True)
        y = y if labels else None
        Om = dh.fill_normal(52, n, l)
        Qo, Wo, info = orc.lra(X, Om, 1.0 / d, 1, y)
        assert info == 0
        Zo, lo, sto = orc.nystrom(Qo, Wo, k)
    finally:
        orc.set_threads(1)
    assert sto == 0
    Z = torch.zeros((k, n), dtype=torch.float64, device="cuda")
    lam = torch.zeros(k, dtype=torch.float64, device="cuda")
    assert dev.nystrom(dh.to_dev(Qo), dh.to_dev(Wo), k, Z, lam) == 0
    Zg, lg = dh.from_dev(Z), lam.cpu().numpy()
    e_zz = rel(Zg @ Zg.T, Zo @ Zo.T)
    print(f"nystrom n={n} l={l} k={k}: ZZ^T vs oracle {e_zz:.2e}, "
          f"lam {np.abs(lg - lo).max() / lo[0]:.2e}", flush=True)
    assert np.abs(lg - lo).max() <= 1e-12 * lo[0]
    assert e_zz < 1e-11


def test_nystrom_after_mixed_range_finder_full_size(dev):
    """C1/C2 size: FP16 range finder (n = 65536, d = 128, l = 272, q = 2), then
    Z for k = 256; Z^T Z = diag(lam), and ||K - Z Z^T|| is no better than the
    range finder's ||K - Q W^T|| but close to it."""
    import time

    import torch
    n, d, k, p, q = 65536, 128, 256, 16, 2
    l, gamma = k + p, 0.5 / d
    Xd = torch.from_numpy(dh.fill_mixture(0, d, n, 16, 0.25)).cuda()
    Omd = dh.to_dev(dh.fill_normal(1, n, l))
    Q = torch.empty((l, n), dtype=torch.float64, device="cuda")
    W = torch.empty_like(Q)
    assert dev.lra(Xd, Omd, gamma, q, Q, W, prec=dh.DH_PREC_F16) == 0
    Z = torch.zeros((k, n), dtype=torch.float64, device="cuda")
    lam = torch.zeros(k, dtype=torch.float64, device="cuda")
    dev.nystrom(Q, W, k, Z, lam)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    assert dev.nystrom(Q, W, k, Z, lam) == 0
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3
    lg = lam.cpu().numpy()
    G = (Z @ Z.T).cpu().numpy()
    e_g = np.abs(G - np.diag(np.maximum(lg, 0))).max() / lg[0]
    e2, k2 = dev.lra_error(Xd, gamma, Z, Z)
    q2, _ = dev.lra_error(Xd, gamma, Q, W)
    ez, eq = np.sqrt(max(e2, 0) / k2), np.sqrt(max(q2, 0) / k2)
    print(f"nystrom full size: {ms:.2f} ms; |Z^T Z - diag(lam)|/lam_1 {e_g:.1e}; "
          f"||K-ZZ^T||/||K|| {ez:.4e} vs ||K-QW^T||/||K|| {eq:.4e}; lam_1 {lg[0]:.4e} "
          f"lam_k {lg[-1]:.4e}", flush=True)
    assert e_g < 1e-12
    assert np.all(np.diff(lg) <= 0)
    assert eq <= ez * (1 + 1e-9) and ez < 3 * eq


def test_nystrom_exact_low_rank(dev):
    """r = 9 distinct points: rank(K) <= 9, the range finder completes its
    rank-deficient basis, and Z (k = 12) reproduces K to rounding."""
    import torch
    r, n, d, l, k = 9, 5000, 7, 20, 12
    P = np.random.default_rng(8).uniform(-1, 1, (r, d))
    X = P[np.arange(n) % r]
    Xd = torch.from_numpy(X.copy()).cuda()
    Omd = dh.to_dev(dh.fill_normal(9, n, l))
    Q = torch.empty((l, n), dtype=torch.float64, device="cuda")
    W = torch.empty_like(Q)
    dev.lra(Xd, Omd, 0.5, 1, Q, W)
    Z = torch.zeros((k, n), dtype=torch.float64, device="cuda")
    assert dev.nystrom(Q, W, k, Z) == 0
    e2, k2 = dev.lra_error(Xd, 0.5, Z, Z)
    assert np.sqrt(max(e2, 0) / k2) < 1e-11


def test_nystrom_argument_errors(dev):
    import torch
    Q = torch.zeros((10, 100), dtype=torch.float64, device="cuda")
    Z = torch.zeros((11, 100), dtype=torch.float64, device="cuda")
    with pytest.raises(dh.DHError):
        dev.nystrom(Q, Q, 11, Z)  # k > l
