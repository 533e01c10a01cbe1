"""Writes tests/golden/golden.npz: small input/output vectors for the RBF LRA path.

Every expected output here comes from numpy / scipy (extended-precision
numpy where noted), never from the oracle or the HIP library, so the
fixtures pin both independently.  The reference (DArray) ships no golden
vectors or fixtures for this path; its two known-answer Cholesky tests
(perftest_cholesky.cpp:46-53,76-84 and test3.cpp:23-46) are closed forms and
are checked directly in the tests.

    python tests/golden/make_golden.py
"""
from pathlib import Path

import numpy as np
import scipy.linalg as sla

OUT = Path(__file__).resolve().with_name("golden.npz")


def rbf_longdouble(Xi, Xj, gamma, yi=None, yj=None):
    """exp(-gamma * |xi - xj|^2) with the distance formed directly in long double."""
    a = Xi.astype(np.longdouble)
    b = Xj.astype(np.longdouble)
    d2 = ((a[:, None, :] - b[None, :, :]) ** 2).sum(-1)
    K = np.exp(-np.longdouble(gamma) * d2)
    if yi is not None:
        K = K * yi.astype(np.longdouble)[:, None]
    if yj is not None:
        K = K * yj.astype(np.longdouble)[None, :]
    return K


def main():
    rng = np.random.default_rng(20240601)
    g = {}
    # RBF tile, rectangular, labelled (rbf_kernel, dkernel.cu:116-129).
    Xi = rng.uniform(-1, 1, (19, 7))
    Xj = rng.uniform(-1, 1, (23, 7))
    yi = rng.choice([-1.0, 1.0], 19)
    yj = rng.choice([-1.0, 1.0], 23)
    g["tile_Xi"], g["tile_Xj"], g["tile_yi"], g["tile_yj"] = Xi, Xj, yi, yj
    g["tile_gamma"] = np.array(0.37)
    g["tile_K"] = rbf_longdouble(Xi, Xj, 0.37, yi, yj).astype(np.float64)

    # Fused projection Y = K*O + beta*Y0 (DArray::LRA chunk loop, dkernel.h:20-42).
    X = rng.uniform(0, 1, (77, 13))
    O = rng.standard_normal((77, 9))
    Y0 = rng.standard_normal((77, 9))
    K = rbf_longdouble(X, X, 1.0 / 13)
    g["proj_X"], g["proj_O"], g["proj_Y0"] = X, O, Y0
    g["proj_gamma"] = np.array(1.0 / 13)
    g["proj_beta"] = np.array(0.5)
    g["proj_Y"] = (K @ O.astype(np.longdouble) + np.longdouble(0.5) * Y0).astype(np.float64)

    # Cholesky (spotrf_ "U" contract): random SPD, scipy upper factor.
    M = rng.standard_normal((70, 70))
    A = M.T @ M + 70 * np.eye(70)
    g["chol_A"] = A
    g["chol_U"] = sla.cholesky(A, lower=False)

    # Triangular solves against scipy.
    U = np.triu(rng.standard_normal((45, 45))) + 6 * np.eye(45)
    B = rng.standard_normal((45, 8))
    Br = rng.standard_normal((11, 45))
    g["trsm_U"], g["trsm_B"], g["trsm_Br"] = U, B, Br
    g["trsm_LUT"] = sla.solve_triangular(U, B, trans="T", lower=False)
    g["trsm_LUN"] = sla.solve_triangular(U, B, trans="N", lower=False)
    g["trsm_RUN"] = sla.solve_triangular(U, Br.T, trans="T", lower=False).T  # X U = Br

    # QR of a tall panel, normalised to a positive R diagonal.
    Yq = rng.standard_normal((200, 12))
    Qn, Rn = np.linalg.qr(Yq)
    s = np.sign(np.diag(Rn))
    g["qr_Y"], g["qr_Q"], g["qr_R"] = Yq, Qn * s, s[:, None] * Rn

    # Range finder with one power iteration, projector Q Q^T K in long double.
    Xl = rng.uniform(0, 1, (150, 6))
    Om = rng.standard_normal((150, 14))
    gam = 1.0 / 6
    Kl = rbf_longdouble(Xl, Xl, gam).astype(np.float64)
    Q, _ = np.linalg.qr(Kl @ Om)
    Q, _ = np.linalg.qr(Kl @ Q)
    W = Kl @ Q
    g["lra_X"], g["lra_Omega"] = Xl, Om
    g["lra_gamma"] = np.array(gam)
    g["lra_QWt"] = Q @ W.T
    g["lra_err2"] = np.array([np.sum((Kl - Q @ W.T) ** 2), np.sum(Kl**2)])
    np.savez_compressed(OUT, **g)
    print("wrote", OUT, sum(v.nbytes for v in g.values()), "bytes")


if __name__ == "__main__":
    main()
