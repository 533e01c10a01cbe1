"""Barrier SVM solver (dh_svm_barrier) and its preconditioned CG (dh_svm_pcg)
against the oracle's restatement and against the mathematics: the CG solve
against numpy's dense solve; the barrier solution against the box-constrained
QP optimum (L-BFGS-B), which the method's stopping rule (2n/t <= 1e-5 |f|)
bounds to 1e-5 relative."""
import os

import numpy as np
import pytest

import darray_optimizer_amd as dh
import oracle as orc

pytestmark = pytest.mark.gpu


def svm_problem(n, d, k, seed):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-1, 1, (n, d))
    y = np.where(X[:, 0] + 0.3 * rng.standard_normal(n) > 0, 1.0, -1.0)
    K = np.exp(-0.5 * ((X[:, None, :] - X[None, :, :]) ** 2).sum(-1))
    w, U = np.linalg.eigh((y[:, None] * y[None, :]) * K)
    idx = np.argsort(w)[::-1][:k]
    return np.asfortranarray(U[:, idx] * np.sqrt(np.maximum(w[idx], 0))), y


def dvec(v):
    import torch
    return torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64)).cuda()


@pytest.mark.parametrize("n,k", [(3000, 50), (777, 13), (1, 1)])
def test_pcg_vs_dense_solve(dev, n, k):
    rng = np.random.default_rng(n + k)
    Z = rng.standard_normal((n, k))
    D = rng.uniform(0.5, 2.0, n)
    b = rng.standard_normal(n)
    x = dvec(np.zeros(n))
    it, rn = dev.svm_pcg(dh.to_dev(Z), dvec(D), dvec(b), x, tol=1e-11, max_iter=5 * n)
    A = Z @ Z.T + np.diag(D)
    xr = np.linalg.solve(A, b)
    xg = x.cpu().numpy()
    print(f"pcg n={n} k={k}: {it} iterations, |r| {rn:.2e}, rel err {np.linalg.norm(xg - xr) / np.linalg.norm(xr):.2e}",
          flush=True)
    assert rn <= 1e-11
    assert np.linalg.norm(A @ xg - b) <= 1e-10 * np.linalg.norm(b)
    # From the exact solution: no iteration needed.
    it2, _ = dev.svm_pcg(dh.to_dev(Z), dvec(D), dvec(b), dvec(xr), tol=1e-8)
    assert it2 == 0


def test_pcg_reference_operator_with_n_added(dev):
    """pre_conjgrad's operator Z Z^T + D + n I (dbarrier.h:99) is dh_svm_pcg with D + n."""
    rng = np.random.default_rng(9)
    n, k = 500, 20
    Z, D, b = rng.standard_normal((n, k)), rng.uniform(0, 1, n), np.zeros(n)
    b[0] = 1.0  # the reference's right-hand side (set_Identity of an n x 1 matrix)
    x0 = rng.uniform(0.1, 0.5, n)
    x = dvec(x0)
    dev.svm_pcg(dh.to_dev(Z), dvec(D + n), dvec(b), x, tol=1e-12)
    xr = np.linalg.solve(Z @ Z.T + np.diag(D + n), b)
    assert np.linalg.norm(x.cpu().numpy() - xr) <= 1e-10 * np.linalg.norm(xr)


@pytest.mark.parametrize("n,d,k,Cb", [(300, 4, 40, 1.0), (800, 8, 60, 10.0), (1500, 16, 80, 100.0)])
def test_barrier_vs_oracle_and_qp_optimum(dev, n, d, k, Cb):
    from scipy.optimize import minimize
    Z, y = svm_problem(n, d, k, n)
    orc.set_threads(min(16, os.cpu_count() or 1))
    try:
        ao, so, fo, sto = orc.svm_barrier(Z, y, Cb)
    finally:
        orc.set_threads(1)
    a = dvec(np.zeros(n))
    stats, fg = dev.svm_barrier(dh.to_dev(Z), dvec(y), Cb, a)
    ag = a.cpu().numpy()
    G = Z @ Z.T
    res = minimize(lambda x: 0.5 * x @ G @ x - x.sum(), np.full(n, Cb / 2), jac=lambda x: G @ x - 1,
                   method="L-BFGS-B", bounds=[(0, Cb)] * n,
                   options=dict(ftol=1e-15, gtol=1e-12, maxiter=100000))
    print(f"barrier n={n} k={k} C={Cb}: device {stats} f={fg:.10e}; oracle {so} f={fo:.10e}; "
          f"L-BFGS-B f*={res.fun:.10e}; max|a - a_oracle|/C {np.abs(ag - ao).max() / Cb:.2e}",
          flush=True)
    assert sto == 0
    assert np.all(ag > 0) and np.all(ag < Cb)
    assert abs(fg - (0.5 * ag @ G @ ag - ag.sum())) <= 1e-9 * abs(fg)
    assert res.fun <= fg <= res.fun + 1e-5 * abs(res.fun)
    assert abs(fg - fo) <= 1e-5 * abs(fo)
    assert stats[0] == so[0]


def test_barrier_empty_class_is_an_error(dev):
    Z, _ = svm_problem(100, 3, 10, 1)
    with pytest.raises(dh.DHError):
        dev.svm_barrier(dh.to_dev(Z), dvec(np.ones(100)), 1.0, dvec(np.zeros(100)))


def test_svm_on_range_finder_nystrom_factor(dev):
    """n = 16384, d = 16, two-blob labels: labelled FP16 range finder (l = 144),
    Nystrom Z (k = 128), barrier solver; the decision values
    y_j (Z Z^T a)_j + b classify the training set."""
    import time

    import torch
    n, d, k, p = 16384, 16, 128, 16
    rng = np.random.default_rng(12)
    y = np.where(rng.random(n) < 0.4, 1.0, -1.0)
    X = rng.normal(0, 0.7, (n, d))
    X[:, :2] += 1.1 * y[:, None]
    gamma, Cb = 0.1, 10.0
    Xd, yd = dvec(X.reshape(n, d)), dvec(y)
    Q = torch.empty((k + p, n), dtype=torch.float64, device="cuda")
    W = torch.empty_like(Q)
    assert dev.lra(Xd, dh.to_dev(rng.standard_normal((n, k + p))), gamma, 1, Q, W, yd,
                   prec=dh.DH_PREC_F16) == 0
    Z = torch.zeros((k, n), dtype=torch.float64, device="cuda")
    assert dev.nystrom(Q, W, k, Z) == 0
    a = dvec(np.zeros(n))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats, f = dev.svm_barrier(Z, yd, Cb, a)
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    ag, Zh = a.cpu().numpy(), Z.cpu().numpy().T
    s = Zh @ (Zh.T @ ag)
    free = np.flatnonzero((ag > 1e-6) & (ag < Cb - 1e-6))[:50]
    b = np.mean(y[free] * (1 - s[free]))
    acc = np.mean(np.sign(y * s + b) == y)
    print(f"svm n={n}: {stats} in {secs:.2f} s, f={f:.6e}, training accuracy {acc:.4f}", flush=True)
    assert np.all(ag > 0) and np.all(ag < Cb)
    assert acc >= 0.95
