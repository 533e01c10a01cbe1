"""LIBSVM data reader and model writer (host code in libdarray_hip.so; no GPU).

Reader known answers come from a hand-written fixture (tests/golden/tiny.libsvm)
covering the reference's label rule (1 -> +1, anything else -> -1,
read_input.h:23-25), tabs, CRLF, an empty line, a skipped index 0, comments
and a short file.  The model writer is checked by parsing its output back and
recomputing every field, with the bias against its definition on the
kernel itself, y_j - sum_i a_i y_i K_ij, where Z is an exact factor of
diag(y) K diag(y)."""
from pathlib import Path

import numpy as np
import pytest

import darray_optimizer_amd as dh

FIX = Path(__file__).resolve().parent / "golden" / "tiny.libsvm"


def test_reader_known_answers():
    X, y, got = dh.read_libsvm(FIX, 8, 4)
    assert got == 8
    np.testing.assert_array_equal(y, [1, -1, -1, -1, -1, 1, 1, -1])
    want = np.zeros((8, 4))
    want[0, [0, 2]] = [0.5, -2.25]
    want[1, [1, 3]] = [1e-3, 7]
    want[2] = [1, 2, 3, 4]
    want[3, 3] = -0.125  # index 0 is skipped
    want[6, 2] = 1.5
    want[7, 0] = -1
    np.testing.assert_array_equal(X, want)


def test_reader_short_file_and_padding():
    X, y, got = dh.read_libsvm(FIX, 12, 6)
    assert got == 8
    assert np.all(X[8:] == 0) and np.all(y[8:] == 0)
    assert np.all(X[:, 4:] == 0)
    X2, y2, got2 = dh.read_libsvm(FIX, 3, 4)
    assert got2 == 3 and np.array_equal(X2, dh.read_libsvm(FIX, 8, 4)[0][:3])


def test_reader_errors(tmp_path):
    with pytest.raises(dh.DHError):
        dh.read_libsvm(FIX, 8, 3)  # index 4 > d
    bad = tmp_path / "bad.libsvm"
    bad.write_text("1 2:0.5 junk\n")
    with pytest.raises(dh.DHError):
        dh.read_libsvm(bad, 1, 4)
    with pytest.raises(dh.DHError):
        dh.read_libsvm(tmp_path / "missing.libsvm", 1, 4)


def parse_model(path):
    lines = Path(path).read_text().splitlines()
    head, i = {}, 0
    while lines[i] != "SV":
        key, _, val = lines[i].partition(" ")
        head[key] = val
        i += 1
    svs = []
    for ln in lines[i + 1:]:
        tok = ln.split()
        feats = {int(t.split(":")[0]): float(t.split(":")[1]) for t in tok[1:]}
        svs.append((float(tok[0]), feats))
    return head, svs


def test_model_writer_round_trip(tmp_path):
    rng = np.random.default_rng(5)
    n, d, gamma, Cb = 60, 5, 0.7, 2.0
    X = rng.uniform(-1, 1, (n, d))
    X[rng.random((n, d)) < 0.3] = 0.0  # sparse rows
    y = np.where(rng.random(n) < 0.4, 1.0, -1.0)
    D2 = ((X[:, None, :] - X[None, :, :]) ** 2).sum(-1)
    K = np.exp(-gamma * D2)
    # Exact factor Z Z^T = diag(y) K diag(y) (eigendecomposition of a PSD matrix).
    w, U = np.linalg.eigh((y[:, None] * y[None, :]) * K)
    Z = U * np.sqrt(np.maximum(w, 0))
    a = np.zeros(n)
    a[rng.permutation(n)[:35]] = rng.uniform(0, Cb, 35)
    a[:4] = Cb  # bounded support vectors
    a[4:6] = 5e-7  # below the support threshold
    path = tmp_path / "model.txt"
    nsv, b = dh.write_libsvm_model(path, a, Cb, Z, y, X, gamma)
    sv = np.flatnonzero(a > 1e-6)
    free = sv[a[sv] < Cb - 1e-6][:50]
    b_ref = np.mean([y[j] - np.sum(a[sv] * y[sv] * K[sv, j]) for j in free])
    assert nsv == len(sv)
    assert abs(b - b_ref) <= 1e-12 * max(1.0, abs(b_ref))
    head, svs = parse_model(path)
    assert head["svm_type"] == "c_svc" and head["kernel_type"] == "rbf"
    assert float(head["gamma"]) == gamma and head["nr_class"] == "2"
    assert int(head["total_sv"]) == len(sv) and float(head["rho"]) == -b
    assert head["label"] == "1 -1"
    npos = int(np.sum(y[sv] > 0))
    assert head["nr_sv"] == f"{npos} {len(sv) - npos}"
    order = list(sv[y[sv] > 0]) + list(sv[y[sv] < 0])  # class +1 first
    for (coef, feats), i in zip(svs, order):
        assert coef == y[i] * a[i]
        assert feats == {f + 1: X[i, f] for f in range(d) if X[i, f] != 0.0}
