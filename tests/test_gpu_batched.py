"""Batched range finder (dh_lra_batched) against dh_lra run problem by problem.

The BF16 and FP16 batches share every kernel body with the single-problem path; only
the CholeskyQR Gram matrices are summed in another order, so the batched
Q W^T agrees with the single-problem one to BF16 re-rounding level and the
approximation errors agree closely.  The FP64 batch calls dh_lra per problem
and must agree bit for bit.  Graph replays are checked against the first run
and against a fresh plan after the inputs change in place."""
import numpy as np
import pytest

import darray_optimizer_amd as dh

pytestmark = pytest.mark.gpu


def make_problem(rng, n, d, l, q, labels, seed):
    import torch
    X, y = dh.fill_mixture(seed, d, n, 4, 0.3, labels=True)
    p = {"X": torch.from_numpy(X).cuda(), "Omega": dh.to_dev(rng.standard_normal((n, l))),
         "gamma": 1.0 / d, "q": q,
         "Q": torch.empty((l, n), dtype=torch.float64, device="cuda"),
         "W": torch.empty((l, n), dtype=torch.float64, device="cuda")}
    if labels:
        p["y"] = torch.from_numpy(y).cuda()
    return p


SHAPES = [(300, 16, 20), (1000, 64, 42), (4000, 256, 74), (2048, 16, 138), (1500, 64, 266),
          (777, 256, 42), (3000, 64, 138), (512, 16, 74)]


def problems(q=1, shapes=SHAPES):
    rng = np.random.default_rng(11)
    return [make_problem(rng, n, d, l, q, labels=(i % 2 == 1), seed=100 + i)
            for i, (n, d, l) in enumerate(shapes)]


def single(dev, p, prec):
    import torch
    Q = torch.empty_like(p["Q"])
    W = torch.empty_like(p["W"])
    st = dev.lra(p["X"], p["Omega"], p["gamma"], p["q"], Q, W, y=p.get("y"), prec=prec)
    return st, Q, W


def rel_err(dev, p, Q, W):
    e2, k2 = dev.lra_error(p["X"], p["gamma"], Q, W, y=p.get("y"))
    return float(np.sqrt(max(e2, 0.0) / k2))


@pytest.mark.parametrize("prec", [dh.DH_PREC_BF16, dh.DH_PREC_F16], ids=["bf16", "f16"])
def test_16bit_batch_matches_single_problems(dev, prec):
    """The FP16 batch finds each problem's power-of-two scale on the device
    (the single-problem path reads it back to the host): same scale, same
    rounding, so the same bar as BF16 holds."""
    import torch
    probs = problems()
    print("batch:", [(p["X"].shape[0], p["X"].shape[1], p["Omega"].shape[0]) for p in probs],
          flush=True)
    r, st = dev.lra_batched(probs, prec=prec)
    assert r == 0 and st == [0] * len(probs), st
    for i, p in enumerate(probs):
        s1, Q1, W1 = single(dev, p, prec)
        assert s1 == 0
        Q2, W2 = p["Q"], p["W"]
        l = Q2.shape[0]
        orth = float((Q2 @ Q2.T - torch.eye(l, dtype=torch.float64, device="cuda")).abs().max())
        A1, A2 = Q1.T @ W1, Q2.T @ W2
        diff = float((A1 - A2).norm() / A1.norm())
        e1, e2 = rel_err(dev, p, Q1, W1), rel_err(dev, p, Q2, W2)
        print(f"problem {i}: n={p['X'].shape[0]} d={p['X'].shape[1]} l={l}: |QtQ-I|={orth:.1e} "
              f"QW^T diff {diff:.2e}  err single {e1:.4e} batched {e2:.4e}", flush=True)
        assert orth < 1e-12
        assert diff < 2e-3
        assert abs(e2 - e1) <= 0.02 * e1 + 1e-6


@pytest.mark.parametrize("prec", [dh.DH_PREC_BF16, dh.DH_PREC_F16], ids=["bf16", "f16"])
def test_16bit_graph_replay_and_inputs_changed_in_place(dev, prec):
    """Replays read the current inputs: for FP16 the scale is recomputed inside
    the graph, so rescaled points (X * 3) give the same result as a fresh plan."""
    import torch
    probs = problems(shapes=SHAPES[:5])
    dev.lra_batched(probs, prec=prec)
    first = [(p["Q"].clone(), p["W"].clone()) for p in probs]
    for rep in range(3):  # replays of the captured graph, with eager work in between
        torch.cuda.synchronize()
        junk = torch.randn(1000, device="cuda").sum().item()
        r, st = dev.lra_batched(probs, prec=prec)
        assert r == 0
        for (Q0, W0), p in zip(first, probs):
            assert torch.equal(Q0, p["Q"]) and torch.equal(W0, p["W"]), f"replay {rep} differs"
    # New data in the same buffers: the replay must read it.
    for p in probs:
        p["Omega"].mul_(-0.5).add_(0.25)
        p["X"].mul_(3.0)
    dev.lra_batched(probs, prec=prec)
    replayed = [(p["Q"].clone(), p["W"].clone()) for p in probs]
    fresh = dh.Handle(0)
    try:
        fresh.lra_batched(probs, prec=prec)
        for (Qr, Wr), p in zip(replayed, probs):
            assert torch.equal(Qr, p["Q"]) and torch.equal(Wr, p["W"])
    finally:
        fresh.close()


def test_fp64_batch_is_per_problem_lra(dev):
    import torch
    probs = problems(q=2, shapes=SHAPES[:4])
    r, st = dev.lra_batched(probs, prec=dh.DH_PREC_F64)
    assert r == 0
    for p in probs:
        s1, Q1, W1 = single(dev, p, dh.DH_PREC_F64)
        assert s1 == 0
        assert torch.equal(Q1, p["Q"]) and torch.equal(W1, p["W"])


def test_breakdown_is_redone_alone(dev):
    """A problem with 3 distinct points (K of rank 3 < l) breaks CholeskyQR;
    the batch redoes it with dh_lra and leaves the other problems alone."""
    import torch
    probs = problems(shapes=SHAPES[:3])
    bad = probs[1]
    X = bad["X"]
    X.copy_(X[torch.arange(X.shape[0], device="cuda") % 3])
    r, st = dev.lra_batched(probs)
    s1, Q1, W1 = single(dev, bad, dh.DH_PREC_BF16)
    print("statuses", st, "single", s1, flush=True)
    assert st[1] == s1 and st[0] == 0 and st[2] == 0
    assert torch.equal(Q1, bad["Q"]) and torch.equal(W1, bad["W"])


def test_batched_argument_errors(dev):
    probs = problems(shapes=SHAPES[:2])
    probs[1]["q"] = 2
    with pytest.raises(dh.DHError):
        dev.lra_batched(probs)
    with pytest.raises(dh.DHError):
        dev.lra_batched(probs, prec=dh.DH_PREC_F16)


SINGLE_OFF = pytest.mark.skip(
    reason="single-launch batch schedule disabled: its kernel hung on the device "
           "(a wall-clock bound on the stage waits did not end it, so the hang is inside a "
           "work item; cause not found); DH_BATCH_SCHEDULE=single now reports "
           "DH_ERR_UNSUPPORTED")


@SINGLE_OFF
@pytest.mark.parametrize("prec", [dh.DH_PREC_BF16, dh.DH_PREC_F16], ids=["bf16", "f16"])
def test_single_launch_schedule_matches_staged(prec, monkeypatch):
    """The whole batch as one persistent launch (every stage a work item with
    per-problem stage dependencies) against the staged graph: the same kernels'
    arithmetic, so Q and W agree bit for bit; replays agree too."""
    import torch
    out = {}
    for mode in ("staged", "single"):
        monkeypatch.setenv("DH_BATCH_SCHEDULE", mode)
        probs = problems(q=1)
        h = dh.Handle(0)
        try:
            print(f"schedule {mode}", flush=True)
            r, st = h.lra_batched(probs, prec=prec)
            assert r == 0 and st == [0] * len(probs), st
            first = [(p["Q"].clone(), p["W"].clone()) for p in probs]
            r, st = h.lra_batched(probs, prec=prec)  # graph replay
            assert r == 0
            for (Q0, W0), p in zip(first, probs):
                assert torch.equal(Q0, p["Q"]) and torch.equal(W0, p["W"])
            out[mode] = first
        finally:
            h.close()
    for (Qa, Wa), (Qb, Wb) in zip(out["staged"], out["single"]):
        assert torch.equal(Qa, Qb) and torch.equal(Wa, Wb)


@SINGLE_OFF
def test_single_launch_schedule_redoes_breakdown(monkeypatch):
    """A problem whose CholeskyQR breaks down inside the single launch is
    redone alone by dh_lra, as in the staged graph."""
    import torch
    monkeypatch.setenv("DH_BATCH_SCHEDULE", "single")
    probs = problems(shapes=SHAPES[:3])
    bad = probs[1]
    X = bad["X"]
    X.copy_(X[torch.arange(X.shape[0], device="cuda") % 3])
    h = dh.Handle(0)
    try:
        r, st = h.lra_batched(probs)
        s1, Q1, W1 = single(h, bad, dh.DH_PREC_BF16)
        print("statuses", st, "single", s1, flush=True)
        assert st[1] == s1 and st[0] == 0 and st[2] == 0
        assert torch.equal(Q1, bad["Q"]) and torch.equal(W1, bad["W"])
    finally:
        h.close()


def test_single_launch_schedule_is_refused(dev, monkeypatch):
    """Asking for the disabled single-launch schedule is an error, not a
    silent run of the staged graph."""
    monkeypatch.setenv("DH_BATCH_SCHEDULE", "single")
    with pytest.raises(dh.DHError):
        dev.lra_batched(problems(shapes=SHAPES[:2]))
