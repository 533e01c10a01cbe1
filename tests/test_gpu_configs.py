"""The benchmark's five workloads through the C ABI, each against the FP64
oracle at a size the oracle finishes in seconds and through size-independent
properties at full size.

  C0  FP64 range finder at its exact shape: n = 8192 points uniform in [0,1)^64
      (DArray's set_uniform, seed 0), Omega from set_normal_seed(1), l = 138,
      q = 1, gamma = 1/64 (testsvm.cpp:86-118) -- oracle bar 1e-10.
  C2  FP16 and BF16 range finders on the 16-centre Gaussian mixture (sigma 0.25,
      gamma = 0.5/d), d = 128, l = 272, q = 2: at n = 8192 against the oracle,
      at n = 65536 through orthonormality, the error identity and the error
      ratio to the FP64 path.  The FP16 mode carries the 1.1x error gate.
  C3  batched problems of uneven n, d, l with labels: each against the oracle's
      FP64 error, and the full 96-problem batch for orthonormality and the
      error identity.
  C4  FP8-distance passes at l = 544 (two column groups) against the oracle
      and an e4m3 emulation; the full n = 262144, d = 256 solve through its
      properties.

Errors are relative Frobenius norms; ||K - Q W^T|| / ||K|| comes from the
FP64 error pass (K regenerated on the device) or the oracle's."""
import os

import numpy as np
import pytest

import darray_optimizer_amd as dh
import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def oracle_threads():
    orc.set_threads(min(16, os.cpu_count() or 1))
    yield
    orc.set_threads(1)


def dev_pts(X):
    import torch
    return torch.from_numpy(np.ascontiguousarray(X)).cuda()


def empty_panel(n, l):
    import torch
    return torch.empty((l, n), dtype=torch.float64, device="cuda")


def lra(dev, Xd, Omd, gamma, q, prec, y=None):
    n, l = Xd.shape[0], Omd.shape[0]
    Q, W = empty_panel(n, l), empty_panel(n, l)
    st = dev.lra(Xd, Omd, gamma, q, Q, W, y, prec=prec)
    assert st == 0, st
    return Q, W


def rel_err(dev, Xd, gamma, Q, W, y=None):
    e2, k2 = dev.lra_error(Xd, gamma, Q, W, y)
    return float(np.sqrt(max(e2, 0.0) / k2)), e2, k2


def product_rel(Qa, Wa, Qb, Wb, block=2048):
    """||Qa Wa^T - Qb Wb^T||_F / ||Qb Wb^T||_F for (l, n) device panels (n x l
    column-major), formed in row blocks on the device in FP64 -- a
    sign-invariant comparison of two range finders."""
    n = Qa.shape[1]
    num = den = 0.0
    for i0 in range(0, n, block):
        i1 = min(n, i0 + block)
        A = Qa[:, i0:i1].T.contiguous() @ Wa
        B = Qb[:, i0:i1].T.contiguous() @ Wb
        num += float(((A - B) ** 2).sum())
        den += float((B ** 2).sum())
    return float(np.sqrt(num / den))


def orthonormality(Q):
    import torch
    l = Q.shape[0]
    return float((Q @ Q.T - torch.eye(l, dtype=torch.float64, device="cuda")).norm())


def c1_mixture(n, d=128):
    return dh.fill_mixture(0, d, n, 16, 0.25)


# ---------------------------------------------------------------- C0

def test_c0_fp64_range_finder_at_exact_shape(dev):
    n, d, k, p, q = 8192, 64, 128, 10, 1
    l, gamma = k + p, 1.0 / d
    X = dh.fill_uniform(0, d, n).T.copy()
    Om = dh.fill_normal(1, n, l)
    print(f"C0: n={n} d={d} l={l} q={q}", flush=True)
    Xd = dev_pts(X)
    Q, W = lra(dev, Xd, dh.to_dev(Om), gamma, q, dh.DH_PREC_F64)
    Qo, Wo, info = orc.lra(X, Om, gamma, q, orth="householder")
    assert info == 0
    Qod, Wod = dh.to_dev(Qo), dh.to_dev(Wo)
    e_qw = product_rel(Q, W, Qod, Wod)
    e_qq = product_rel(Q, Q, Qod, Qod)
    err, _, _ = rel_err(dev, Xd, gamma, Q, W)
    eo, ko = orc.lra_error(X, gamma, Qo, Wo)
    err_o = float(np.sqrt(eo / ko))
    print(f"C0: QW^T vs oracle {e_qw:.2e}, QQ^T {e_qq:.2e}; ||K-QB||/||K|| device {err:.6e} "
          f"oracle {err_o:.6e}", flush=True)
    assert e_qw < 1e-10 and e_qq < 1e-10
    assert abs(err - err_o) <= 1e-10 * err_o + 1e-14
    assert orthonormality(Q) < 1e-12


# ---------------------------------------------------------------- C2

@pytest.fixture(scope="module")
def c2_small():
    """n = 8192 of the C1/C2 mixture with the oracle's FP64 range finder."""
    n, d, l, q = 8192, 128, 272, 2
    gamma = 0.5 / d
    X = c1_mixture(n, d)
    Om = dh.fill_normal(1, n, l)
    Qo, Wo, info = orc.lra(X, Om, gamma, q)
    assert info == 0
    eo, ko = orc.lra_error(X, gamma, Qo, Wo)
    return {"X": X, "Om": Om, "gamma": gamma, "q": q, "Qo": Qo, "Wo": Wo,
            "err": float(np.sqrt(eo / ko))}


# Mixed-precision bars at n = 8192 (FP64 error 5.3e-4): QW^T of the mixed
# range finder against the oracle's (its rounding moves the basis within the
# approximation error), and the error ratio.  FP16 meets the 1.1x gate
# (measured 1.009); the BF16 mode (3 fewer mantissa bits on K and the
# operand: 2^-9 relative per entry of K against an approximation error of
# 5e-4) is the faster mode and does not (measured 1.50 here, 1.135 at
# n = 65536): it is held to 1.6x.
C2_BARS = {dh.DH_PREC_F16: (2e-3, 1.1), dh.DH_PREC_BF16: (5e-3, 1.6)}

# ||K - Q W^T||^2 = ||K||^2 - ||W||^2 holds exactly for W = K Q; a mixed pass
# returns W = K Q + E, which moves the identity by 2 <K Q, E> + 2 ||E||^2, of
# the order of the pass's rounding relative to ||K||^2.
# Measured: FP64 5e-17, FP16 1.3e-6, BF16 1.0e-4 to 1.4e-4, FP8 8e-5.
IDENTITY_TOL = {dh.DH_PREC_F64: 1e-9, dh.DH_PREC_F16: 1e-5, dh.DH_PREC_BF16: 5e-4,
                dh.DH_PREC_FP8: 1e-3}


@pytest.mark.parametrize("prec", [dh.DH_PREC_F16, dh.DH_PREC_BF16], ids=["f16", "bf16"])
def test_c2_mixed_range_finder_vs_oracle(dev, c2_small, prec):
    c = c2_small
    Xd = dev_pts(c["X"])
    Q, W = lra(dev, Xd, dh.to_dev(c["Om"]), c["gamma"], c["q"], prec)
    e_qw = product_rel(Q, W, dh.to_dev(c["Qo"]), dh.to_dev(c["Wo"]))
    err, e2, k2 = rel_err(dev, Xd, c["gamma"], Q, W)
    orth = orthonormality(Q)
    ratio = err / c["err"]
    print(f"C2 n=8192 prec={prec}: QW^T vs oracle {e_qw:.2e}; error {err:.6e} vs oracle FP64 "
          f"{c['err']:.6e} (ratio {ratio:.4f}); |QtQ-I| {orth:.1e}", flush=True)
    bar_qw, bar_ratio = C2_BARS[prec]
    assert orth < 1e-12
    assert e_qw < bar_qw
    assert ratio <= bar_ratio


def test_c2_full_size_properties_and_error_gate(dev):
    """n = 65536, d = 128, l = 272, q = 2: FP16 and BF16 solves against the FP64
    HIP path (the oracle is validated against it at n = 8192 and C0)."""
    import torch
    n, d, l, q = 65536, 128, 272, 2
    gamma = 0.5 / d
    Xd = dev_pts(c1_mixture(n, d))
    Omd = dh.to_dev(dh.fill_normal(1, n, l))
    errs = {}
    for prec in (dh.DH_PREC_F64, dh.DH_PREC_F16, dh.DH_PREC_BF16):
        Q, W = lra(dev, Xd, Omd, gamma, q, prec)
        err, e2, k2 = rel_err(dev, Xd, gamma, Q, W)
        w2 = float((W * W).sum())
        orth = orthonormality(Q)
        print(f"C2 n=65536 prec={prec}: error {err:.6e}, |QtQ-I| {orth:.1e}, "
              f"identity residual {abs(e2 - (k2 - w2)) / k2:.1e}", flush=True)
        assert orth < 1e-12
        # Q orthonormal: ||K - Q Q^T K||^2 = ||K||^2 - ||Q^T K||^2 with W = K Q.
        assert abs(e2 - (k2 - w2)) <= IDENTITY_TOL[prec] * k2
        errs[prec] = err
        del Q, W
        torch.cuda.empty_cache()
    r16 = errs[dh.DH_PREC_F16] / errs[dh.DH_PREC_F64]
    rbf = errs[dh.DH_PREC_BF16] / errs[dh.DH_PREC_F64]
    print(f"C2 n=65536 error ratios vs FP64: f16 {r16:.4f}, bf16 {rbf:.4f}", flush=True)
    assert r16 <= 1.1
    assert rbf <= C2_BARS[dh.DH_PREC_BF16][1]


# ---------------------------------------------------------------- C3

def c3_problem(seed, n, d, k, p=10):
    X, y = dh.fill_mixture(seed, d, n, 4, 0.3, labels=True)
    rng = np.random.default_rng(seed)
    return X, y, rng.standard_normal((n, k + p)), 1.0 / d


# (n, d, k): every d class and every k of the batched workload at reduced n.
C3_SHAPES = [(2048, 16, 32), (2600, 64, 64), (3000, 256, 128), (2222, 16, 256),
             (4096, 64, 256), (3333, 256, 32), (2900, 64, 128), (2048, 256, 64)]
# Error ratio to the oracle's FP64 error, per problem.  Where the FP64 error
# is far above the rounding of K (most problems) both modes stay within 1.01x;
# at d = 16, l = 266 the FP64 error (2.8e-4) is of the order of K's 16-bit
# rounding, 2^-12 (FP16) or 2^-9 (BF16) per entry, which sets the floor:
# measured 1.12 (FP16) and 4.4 (BF16) there.
C3_BARS = {dh.DH_PREC_F16: 1.2, dh.DH_PREC_BF16: 5.0}


@pytest.mark.parametrize("prec", [dh.DH_PREC_F16, dh.DH_PREC_BF16], ids=["f16", "bf16"])
def test_c3_batch_vs_oracle(dev, prec):
    import torch
    probs, ref = [], []
    for i, (n, d, k) in enumerate(C3_SHAPES):
        X, y, Om, gamma = c3_problem(300 + i, n, d, k)
        Qo, Wo, info = orc.lra(X, Om, gamma, 1, y)
        assert info == 0
        eo, ko = orc.lra_error(X, gamma, Qo, Wo, y)
        ref.append(float(np.sqrt(eo / ko)))
        l = k + 10
        probs.append({"X": dev_pts(X), "y": dev_pts(y), "Omega": dh.to_dev(Om), "gamma": gamma,
                      "q": 1, "Q": empty_panel(n, l), "W": empty_panel(n, l)})
    r, st = dev.lra_batched(probs, prec=prec)
    assert r == 0 and st == [0] * len(probs), st
    worst = 0.0
    for i, p in enumerate(probs):
        err, _, _ = rel_err(dev, p["X"], p["gamma"], p["Q"], p["W"], p["y"])
        orth = float((p["Q"] @ p["Q"].T - torch.eye(p["Q"].shape[0], dtype=torch.float64,
                                                      device="cuda")).norm())
        ratio = err / ref[i]
        worst = max(worst, ratio)
        print(f"C3 prec={prec} problem {i} {C3_SHAPES[i]}: error {err:.4e} oracle {ref[i]:.4e} "
              f"ratio {ratio:.4f} |QtQ-I| {orth:.1e}", flush=True)
        assert orth < 1e-12
    assert worst <= C3_BARS[prec], worst


def test_c3_full_batch_of_96(dev):
    """The batched workload itself: 96 problems, n ~ U{2048..32768}, d in
    {16, 64, 256}, k in {32, 64, 128, 256}, p = 10, q = 1, labels; every
    problem returns an orthonormal basis and W = K Q satisfies the error
    identity (checked on every 12th problem)."""
    import torch
    rng = np.random.default_rng(3)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(3)
    probs = []
    for i in range(96):
        n = int(rng.integers(2048, 32769))
        d = int(rng.choice([16, 64, 256]))
        k = int(rng.choice([32, 64, 128, 256]))
        l = k + 10
        X, y = dh.fill_mixture(3000 + i, d, n, 4, 0.3, labels=True)
        probs.append({"X": dev_pts(X), "y": dev_pts(y),
                      "Omega": torch.randn((l, n), dtype=torch.float64, device="cuda",
                                           generator=gen),
                      "gamma": 1.0 / d, "q": 1, "Q": empty_panel(n, l), "W": empty_panel(n, l)})
    for prec in (dh.DH_PREC_F16, dh.DH_PREC_BF16):
        r, st = dev.lra_batched(probs, prec=prec)
        assert r == 0 and st == [0] * 96, st
        worst = max(orthonormality(p["Q"]) for p in probs)
        print(f"C3 96 problems prec={prec}: max |QtQ-I| {worst:.1e}", flush=True)
        assert worst < 1e-12
        for p in probs[::12]:
            e2, k2 = dev.lra_error(p["X"], p["gamma"], p["Q"], p["W"], p["y"])
            w2 = float((p["W"] * p["W"]).sum())
            assert abs(e2 - (k2 - w2)) <= IDENTITY_TOL[prec] * k2
    del probs
    torch.cuda.empty_cache()


# ---------------------------------------------------------------- C4

@pytest.mark.parametrize("m,n", [(2048, 2048), (1000, 3000)])
def test_c4_fp8_pass_at_l544(dev, m, n):
    """l = 544 runs as two column groups of 17 blocks; C4's data (32-centre
    mixture in [-1,1]^256, sigma 0.2, gamma = 0.25/d)."""
    import torch
    from test_gpu_mixed import emulate_fp8, rel
    d, l = 256, 544
    X = dh.fill_mixture(4, d, max(m, n), 32, 0.2)
    Xi, Xj = X[:m], X[-n:]
    O = np.random.default_rng(m + n).standard_normal((n, l))
    gamma = 0.25 / d
    Y = torch.empty((l, m), dtype=torch.float64, device="cuda")
    dev.rbf_project_fp8(dev_pts(Xi), dev_pts(Xj), dh.to_dev(O), gamma, Y)
    Yg = dh.from_dev(Y)
    e_emu = rel(Yg, emulate_fp8(Xi, Xj, O, gamma))
    e_orc = rel(Yg, orc.rbf_project(Xi, Xj, O, gamma))
    print(f"C4 fp8 pass m={m} n={n} l={l}: vs e4m3 emulation {e_emu:.2e}, vs oracle {e_orc:.2e}",
          flush=True)
    assert e_emu < 5e-4
    assert e_orc < 3e-2


def test_c4_fp8_range_finder_vs_oracle(dev):
    """FP8 range finder at l = 544 on C4 data at n = 8192 against the oracle."""
    n, d, l, q = 8192, 256, 544, 1
    gamma = 0.25 / d
    X = dh.fill_mixture(4, d, n, 32, 0.2)
    Om = dh.fill_normal(5, n, l)
    Qo, Wo, info = orc.lra(X, Om, gamma, q)
    assert info == 0
    eo, ko = orc.lra_error(X, gamma, Qo, Wo)
    err_o = float(np.sqrt(eo / ko))
    Xd = dev_pts(X)
    Q, W = lra(dev, Xd, dh.to_dev(Om), gamma, q, dh.DH_PREC_FP8)
    err, _, _ = rel_err(dev, Xd, gamma, Q, W)
    orth = orthonormality(Q)
    print(f"C4 n=8192 fp8: error {err:.4e}, oracle FP64 {err_o:.4e}, |QtQ-I| {orth:.1e}",
          flush=True)
    assert orth < 1e-12
    # e4m3 coordinates (3-bit mantissa) set the floor: a 2e-3 absolute bar.
    assert err < 2e-3


def test_c4_full_size_properties(dev):
    """n = 262144, d = 256, l = 544, q = 1, FP8 distance + BF16 projection."""
    import torch
    n, d, k, p, q = 262144, 256, 512, 32, 1
    l, gamma = k + p, 0.25 / d
    Xd = dev_pts(dh.fill_mixture(4, d, n, 32, 0.2))
    gen = torch.Generator(device="cuda")
    gen.manual_seed(5)
    Omd = torch.randn((l, n), dtype=torch.float64, device="cuda", generator=gen)
    print(f"C4: n={n} d={d} l={l} q={q}", flush=True)
    Q, W = lra(dev, Xd, Omd, gamma, q, dh.DH_PREC_FP8)
    del Omd
    orth = orthonormality(Q)
    err, e2, k2 = rel_err(dev, Xd, gamma, Q, W)
    w2 = float((W * W).sum())
    print(f"C4 full: error {err:.4e}, |QtQ-I| {orth:.1e}, identity residual "
          f"{abs(e2 - (k2 - w2)) / k2:.1e}", flush=True)
    assert orth < 1e-12
    assert abs(e2 - (k2 - w2)) <= IDENTITY_TOL[dh.DH_PREC_FP8] * k2
    assert err < 2e-3
    del Q, W, Xd
    torch.cuda.empty_cache()
