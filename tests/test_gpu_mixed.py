"""Mixed-precision (BF16 MFMA) K-passes and LRA through the C ABI.

Two bars per kernel: (1) against a numpy emulation of the same rounding
scheme -- centred BF16 points, BF16 K, BF16 Omega, wide accumulation -- to
1e-4 relative Frobenius (only FP32 accumulation order and exp2 rounding
differ); (2) against the FP64 oracle to BF16 accuracy (5e-3)."""
import numpy as np
import pytest

import darray_optimizer_amd as dh
import oracle as orc

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def bf16(x):
    """Round to the nearest bfloat16 (ties to even) through float32, as float64."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def f16(x):
    """Round to the nearest IEEE half (ties to even), as float64."""
    return np.asarray(x, dtype=np.float64).astype(np.float16).astype(np.float64)


def emulate(Xi, Xj, O, gamma, yi=None, yj=None, rnd=bf16, scale=1.0):
    mu = Xj.mean(axis=0)
    a = rnd((Xi - mu) * scale)
    b = rnd((Xj - mu) * scale)
    g = gamma / scale**2
    na = (a * a).sum(1)
    nb = (b * b).sum(1)
    d2 = np.maximum(na[:, None] + nb[None, :] - 2 * a @ b.T, 0)
    K = rnd(np.exp(-g * d2))
    Ow = O * (yj[:, None] if yj is not None else 1.0)
    Y = K @ rnd(Ow)
    if yi is not None:
        Y = Y * yi[:, None]
    return Y


def dev_pts(X):
    import torch
    return torch.from_numpy(np.ascontiguousarray(X)).cuda()


def dev_vec(v):
    return None if v is None else dev_pts(v)


SHAPES = [(1, 1, 1, 1), (64, 64, 64, 16), (100, 77, 13, 9), (300, 257, 130, 138),
          (513, 1000, 128, 272), (200, 190, 200, 300), (96, 500, 256, 40),
          # l > 272 with d <= 128: wave pairs over two column groups (and, for n = 2100,
          # a j range split in four)
          (130, 257, 64, 300), (300, 2100, 100, 544)]


@pytest.mark.parametrize("m,n,d,l", SHAPES)
def test_rbf_project_bf16(dev, m, n, d, l):
    import torch
    rng = np.random.default_rng(3 * m + n + d + l)
    Xi, Xj = rng.uniform(-1, 1, (m, d)), rng.uniform(-1, 1, (n, d))
    yi, yj = rng.choice([-1.0, 1.0], m), rng.choice([-1.0, 1.0], n)
    O = rng.standard_normal((n, l))
    gamma = 1.0 / d
    Y = torch.empty((l, m), dtype=torch.float64, device="cuda")
    dev.rbf_project_bf16(dev_pts(Xi), dev_pts(Xj), dh.to_dev(O), gamma, Y, dev_vec(yi),
                         dev_vec(yj))
    Yg = dh.from_dev(Y)
    assert rel(Yg, emulate(Xi, Xj, O, gamma, yi, yj)) < 1e-4
    assert rel(Yg, orc.rbf_project(Xi, Xj, O, gamma, yi, yj)) < 5e-3


def test_rbf_project_bf16_self_tile(dev):
    # Xi is Xj (one conversion shared), mixture data at the C1 scale of gamma.
    import torch
    n, d, l = 2048, 128, 272
    X = dh.fill_mixture(3, d, n, 16, 0.25)
    O = np.asfortranarray(dh.fill_normal(4, n, l))
    Xd = dev_pts(X)
    Y = torch.empty((l, n), dtype=torch.float64, device="cuda")
    dev.rbf_project_bf16(Xd, Xd, dh.to_dev(O), 0.5 / d, Y)
    Yg = dh.from_dev(Y)
    assert rel(Yg, emulate(X, X, O, 0.5 / d)) < 1e-4
    assert rel(Yg, orc.rbf_project(X, X, O, 0.5 / d)) < 5e-3


def f16_scale(Xi, Xj):
    """Power of two 2^-e with max |x - mean| 2^-e in [1/2, 1), as the FP16 path picks it."""
    mu = Xj.mean(axis=0)
    mx = max(np.abs(Xi - mu).max(), np.abs(Xj - mu).max())
    return 1.0 if mx == 0 else 2.0 ** -np.frexp(np.float32(mx))[1]


@pytest.mark.parametrize("m,n,d,l", SHAPES)
def test_rbf_project_f16(dev, m, n, d, l):
    """FP16 operands: emulation to 1e-4, oracle to FP16 accuracy (1e-3, tighter than BF16)."""
    import torch
    rng = np.random.default_rng(7 * m + n + d + l)
    Xi, Xj = rng.uniform(-1, 1, (m, d)) * 300.0, rng.uniform(-1, 1, (n, d)) * 300.0
    yi, yj = rng.choice([-1.0, 1.0], m), rng.choice([-1.0, 1.0], n)
    O = rng.standard_normal((n, l))
    gamma = 1.0 / (d * 300.0**2)
    Y = torch.empty((l, m), dtype=torch.float64, device="cuda")
    dev.rbf_project_f16(dev_pts(Xi), dev_pts(Xj), dh.to_dev(O), gamma, Y, dev_vec(yi), dev_vec(yj))
    Yg = dh.from_dev(Y)
    sc = f16_scale(Xi, Xj)
    assert rel(Yg, emulate(Xi, Xj, O, gamma, yi, yj, rnd=f16, scale=sc)) < 1e-4
    assert rel(Yg, orc.rbf_project(Xi, Xj, O, gamma, yi, yj)) < 1e-3


def test_rbf_project_f16_beats_bf16(dev):
    import torch
    n, d, l = 2048, 128, 272
    X = dh.fill_mixture(3, d, n, 16, 0.25)
    O = np.asfortranarray(dh.fill_normal(4, n, l))
    Xd, Od = dev_pts(X), dh.to_dev(O)
    ref = orc.rbf_project(X, X, O, 0.5 / d)
    Yb = torch.empty((l, n), dtype=torch.float64, device="cuda")
    Yh = torch.empty_like(Yb)
    dev.rbf_project_bf16(Xd, Xd, Od, 0.5 / d, Yb)
    dev.rbf_project_f16(Xd, Xd, Od, 0.5 / d, Yh)
    eb, eh = rel(dh.from_dev(Yb), ref), rel(dh.from_dev(Yh), ref)
    print(f"rel err vs FP64: bf16 {eb:.3e}  f16 {eh:.3e}")
    assert eh < eb / 4


@pytest.mark.parametrize("q", [1, 2])
@pytest.mark.parametrize("labels", [False, True])
def test_lra_mixed_error_close_to_fp64(dev, labels, q):
    """Same seeded inputs through the FP64 and the mixed LRA: the mixed
    approximation error stays within a small factor of the FP64 one."""
    import torch
    n, d, l = 4096, 64, 74
    X, y = dh.fill_mixture(5, d, n, 4, 0.3, labels=True)
    y = y if labels else None
    Om = dh.fill_normal(6, n, l)
    Xd, Omd, yd = dev_pts(X), dh.to_dev(Om), dev_vec(y)
    errs = {}
    for prec in (dh.DH_PREC_F64, dh.DH_PREC_BF16, dh.DH_PREC_F16):
        Q = torch.empty((l, n), dtype=torch.float64, device="cuda")
        W = torch.empty((l, n), dtype=torch.float64, device="cuda")
        assert dev.lra(Xd, Omd, 1.0 / d, q, Q, W, yd, prec=prec) == 0
        QtQ = (Q @ Q.T).cpu().numpy()
        assert np.linalg.norm(QtQ - np.eye(l)) < 1e-12
        e2, k2 = dev.lra_error(Xd, 1.0 / d, Q, W, yd)
        errs[prec] = np.sqrt(e2 / k2)
    Qo, Wo, _ = orc.lra(X, Om, 1.0 / d, q, y)
    eo, ko = orc.lra_error(X, 1.0 / d, Qo, Wo, y)
    assert abs(errs[dh.DH_PREC_F64] - np.sqrt(eo / ko)) <= 1e-6 * np.sqrt(eo / ko) + 1e-12
    assert errs[dh.DH_PREC_BF16] <= 1.5 * errs[dh.DH_PREC_F64] + 3e-3
    assert errs[dh.DH_PREC_F16] <= 1.2 * errs[dh.DH_PREC_F64] + 5e-4


@pytest.mark.parametrize("q", [0, 1, 2])
@pytest.mark.parametrize("prec", ["BF16", "F16"])
def test_lra_mixed_breakdown_redoes_the_queued_pass(dev, prec, q):
    """Two equal columns of Omega make the first Y exactly rank deficient in
    any precision, so the first orthogonalisation breaks down.  The 16-bit
    range finder checks that status only after queuing the next pass: the
    fallback then completes the basis and the pass is redone.  q = 0 covers
    the final pass, q = 1 the copy of W into Q, q = 2 a power iteration."""
    import torch
    n, d, l = 3000, 16, 24
    X, _ = dh.fill_mixture(12, d, n, 4, 0.3, labels=True)
    Om = np.array(dh.fill_normal(13, n, l))
    Om[:, 5] = Om[:, 2]
    Xd, Omd = dev_pts(X), dh.to_dev(Om)
    errs = {}
    for p in (dh.DH_PREC_F64, getattr(dh, "DH_PREC_" + prec)):
        Q = torch.empty((l, n), dtype=torch.float64, device="cuda")
        W = torch.empty((l, n), dtype=torch.float64, device="cuda")
        assert dev.lra(Xd, Omd, 1.0 / d, q, Q, W, prec=p) == 0
        QtQ = (Q @ Q.T).cpu().numpy()
        assert np.linalg.norm(QtQ - np.eye(l)) < 1e-12
        e2, k2 = dev.lra_error(Xd, 1.0 / d, Q, W)
        errs[p] = np.sqrt(e2 / k2)
    e64, emix = errs[dh.DH_PREC_F64], errs[getattr(dh, "DH_PREC_" + prec)]
    print(f"{prec} q={q}: error {emix:.3e} (FP64 {e64:.3e})", flush=True)
    assert emix <= 1.5 * e64 + 3e-3


def test_kernel_timing_counts_projection_launches(dev):
    import torch
    n, d, l = 1024, 64, 40
    X = dh.fill_mixture(9, d, n, 4, 0.3)
    O = np.asfortranarray(dh.fill_normal(10, n, l))
    Xd, Od = dev_pts(X), dh.to_dev(O)
    Y = torch.empty((l, n), dtype=torch.float64, device="cuda")
    dev.rbf_project_bf16(Xd, Xd, Od, 1.0 / d, Y)  # schedule settled before timing
    dev.timing(True)
    dev.timing_read()
    for _ in range(3):
        dev.rbf_project_bf16(Xd, Xd, Od, 1.0 / d, Y)
    dev.rbf_project_f64(Xd, Xd, Od, 1.0 / d, Y)
    ms, launches = dev.timing_read()
    dev.timing(False)
    assert launches == 4 and ms > 0.0
    assert dev.timing_read() == (0.0, 0)


def e4m3(v):
    """Round to the nearest OCP FP8 e4m3 value (ties to even); |v| < 448 assumed."""
    v = np.asarray(v, dtype=np.float64)
    a = np.abs(v)
    _, e = np.frexp(a)
    E = np.maximum(e - 1, -6)            # exponent of the leading bit; subnormals below 2^-6
    quantum = np.ldexp(1.0, E - 3)
    return np.sign(v) * np.round(a / quantum) * quantum


def emulate_fp8(Xi, Xj, O, gamma, yi=None, yj=None):
    """FP8 distance (per-row power-of-two scales) and BF16 K / O, as the device computes."""
    mu = Xj.mean(axis=0)

    def quant(X):
        C = np.asarray(X - mu, dtype=np.float32).astype(np.float64)
        mx = np.abs(C).max(axis=1)
        _, e = np.frexp(mx.astype(np.float32))
        sc = np.where(mx > 0, np.ldexp(1.0, 8 - e), 1.0)
        return e4m3(C * sc[:, None]) / sc[:, None]

    a, b = quant(Xi), quant(Xj)
    na, nb = (a * a).sum(1), (b * b).sum(1)
    d2 = np.maximum(na[:, None] + nb[None, :] - 2 * a @ b.T, 0)
    K = bf16(np.exp(-gamma * d2))
    Ow = O * (yj[:, None] if yj is not None else 1.0)
    Y = K @ bf16(Ow)
    return Y * yi[:, None] if yi is not None else Y


def test_e4m3_rounding_known_values():
    # Values the device conversion was seen to produce (OCP e4m3): 1.0625 -> 1, 1.1 -> 1.125,
    # 448 stays, 0.001 -> 2^-9 (subnormal step), 0.0703125 exact.
    got = e4m3([1.0625, 1.1, 448.0, 0.001, 0.0703125, -3.0, 240.0])
    assert list(got) == [1.0, 1.125, 448.0, 2.0**-9, 0.0703125, -3.0, 240.0]


@pytest.mark.parametrize("m,n,d,l", [(64, 64, 13, 16), (300, 257, 64, 138), (200, 333, 128, 40),
                                     (513, 1000, 130, 272), (96, 500, 256, 300),
                                     (200, 700, 64, 544)])
def test_rbf_project_fp8(dev, m, n, d, l):
    import torch
    rng = np.random.default_rng(11 * m + n + d + l)
    Xi, Xj = rng.uniform(-1, 1, (m, d)), rng.uniform(-1, 1, (n, d))
    yi, yj = rng.choice([-1.0, 1.0], m), rng.choice([-1.0, 1.0], n)
    O = rng.standard_normal((n, l))
    gamma = 1.0 / d
    Y = torch.empty((l, m), dtype=torch.float64, device="cuda")
    dev.rbf_project_fp8(dev_pts(Xi), dev_pts(Xj), dh.to_dev(O), gamma, Y, dev_vec(yi), dev_vec(yj))
    Yg = dh.from_dev(Y)
    e_emu = rel(Yg, emulate_fp8(Xi, Xj, O, gamma, yi, yj))
    e_orc = rel(Yg, orc.rbf_project(Xi, Xj, O, gamma, yi, yj))
    print(f"fp8 m={m} n={n} d={d} l={l}: vs emulation {e_emu:.2e}, vs FP64 oracle {e_orc:.2e}")
    assert e_emu < 5e-4
    assert e_orc < 3e-2


def test_lra_fp8_error_close_to_fp64(dev):
    """C4-style data (mixture, gamma = 0.25/d, d = 256) at reduced n: the FP8-distance
    range finder stays within a small factor of the FP64 approximation error."""
    import torch
    n, d, l, q = 4096, 256, 80, 1
    X = dh.fill_mixture(21, d, n, 32, 0.2)
    Om = dh.fill_normal(22, n, l)
    Xd, Omd = dev_pts(X), dh.to_dev(Om)
    errs = {}
    for prec in (dh.DH_PREC_F64, dh.DH_PREC_FP8):
        Q = torch.empty((l, n), dtype=torch.float64, device="cuda")
        W = torch.empty((l, n), dtype=torch.float64, device="cuda")
        assert dev.lra(Xd, Omd, 0.25 / d, q, Q, W, prec=prec) == 0
        e2, k2 = dev.lra_error(Xd, 0.25 / d, Q, W)
        errs[prec] = np.sqrt(e2 / k2)
    print(f"LRA errors: fp64 {errs[dh.DH_PREC_F64]:.4e}  fp8 {errs[dh.DH_PREC_FP8]:.4e}")
    assert errs[dh.DH_PREC_FP8] <= 1.5 * errs[dh.DH_PREC_F64] + 2e-3
