"""dh_rbf_project_f64_noalloc (the allocation-free FP64 K*Omega pass) through
the C ABI against dh_rbf_project_f64: the two share every kernel and differ
only in where the squared norms live (a static device array instead of a
handle workspace), so the outputs must be bit-identical -- on two calls in a
row, the second smaller, with labels and beta.  The call must not grow any
handle workspace, and sizes beyond the static array are refused."""
import ctypes as C

import numpy as np
import pytest

import darray_optimizer_amd as dh
from darray_optimizer_amd import _capi

pytestmark = pytest.mark.gpu


def _p(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def project(h, fn, Xi, Xj, O, gamma, Y, yi=None, yj=None, beta=0.0):
    m, d = Xi.shape
    n, l = Xj.shape[0], O.shape[0]
    return _capi.call(fn, h.h, m, n, d, l, _p(Xi), d, _p(Xj), d, _p(yi), _p(yj), float(gamma),
                      _p(O), n, float(beta), _p(Y), m)


def case(rng, m, n, d, l, labels):
    import torch
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    Xi, Xj = dev(rng.uniform(-1, 1, (m, d))), dev(rng.uniform(-1, 1, (n, d)))
    O = dev(rng.standard_normal((l, n)))
    yi = dev(rng.choice([-1.0, 1.0], m)) if labels else None
    yj = dev(rng.choice([-1.0, 1.0], n)) if labels else None
    Y0 = dev(rng.standard_normal((l, m)))
    return Xi, Xj, O, yi, yj, Y0


def test_noalloc_matches_allocating_pass_on_two_calls():
    import torch
    h = dh.Handle(0)
    ref = dh.Handle(0)
    try:
        rng = np.random.default_rng(77)
        for (m, n, d, l, labels, beta) in [(3000, 2500, 100, 200, True, 0.5),
                                           (700, 900, 30, 20, False, 0.0)]:
            print(f"noalloc case m={m} n={n} d={d} l={l} labels={labels} beta={beta}", flush=True)
            Xi, Xj, O, yi, yj, Y0 = case(rng, m, n, d, l, labels)
            gamma = 1.0 / d
            Ya, Yb = Y0.clone(), Y0.clone()
            assert project(h, "dh_rbf_project_f64_noalloc", Xi, Xj, O, gamma, Ya, yi, yj, beta) == 0
            assert project(ref, "dh_rbf_project_f64", Xi, Xj, O, gamma, Yb, yi, yj, beta) == 0
            torch.cuda.synchronize()
            assert torch.equal(Ya, Yb)
            assert h.workspace_bytes() == 0  # nothing allocated by the handle
    finally:
        h.close()
        ref.close()


def test_noalloc_refuses_sizes_beyond_the_static_array():
    import torch
    h = dh.Handle(0)
    try:
        limit = 1 << 20
        Y = torch.zeros((1, 8), dtype=torch.float64, device="cuda")
        X = torch.zeros((8, 1), dtype=torch.float64, device="cuda")
        # m + n = limit + 1 points; nothing is read (the call fails before any launch)
        st = _capi.lib().dh_rbf_project_f64_noalloc(
            h.h, 8, limit - 7, 1, 1, _p(X), 1, _p(X), 1, None, None, 1.0, _p(X), limit - 7, 0.0,
            _p(Y), 8)
        assert st == _capi.DH_ERR_UNSUPPORTED
        assert "static norm buffer" in _capi.last_error()
        assert torch.count_nonzero(Y) == 0
        # d > 256 would need a K block workspace: refused before any launch
        st = _capi.lib().dh_rbf_project_f64_noalloc(
            h.h, 8, 1, 300, 1, _p(X), 300, _p(X), 300, None, None, 1.0, _p(X), 1, 0.0, _p(Y), 8)
        assert st == _capi.DH_ERR_UNSUPPORTED
        assert torch.count_nonzero(Y) == 0
        assert h.workspace_bytes() == 0
    finally:
        h.close()
