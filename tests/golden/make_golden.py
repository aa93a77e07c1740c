"""Writes the committed fixtures of tests/golden from LAPACK (numpy/scipy), independent of the
oracle: singular values of seeded bidiagonals (dgesdd on the dense matrix), Householder QR in
LAPACK's convention (dgeqrf/dorgqr), and the SVD of small seeded matrices.

Run: python tests/golden/make_golden.py  (deterministic; outputs are small .npz files)."""
import os

import numpy as np
import scipy.linalg.lapack as lapack

HERE = os.path.dirname(os.path.abspath(__file__))


def bidiag():
    out = {}
    rng = np.random.default_rng(2024)
    cases = {
        "gauss16": (rng.standard_normal(16), rng.standard_normal(15)),
        "pos33": (rng.uniform(0.5, 2, 33), rng.uniform(0.1, 1, 32)),
        "graded24": (10.0 ** -rng.uniform(0, 6, 24), 10.0 ** -rng.uniform(0, 6, 23)),
        "linear21": (np.arange(1, 22, dtype=float), np.full(20, 0.5)),
    }
    for tag, (a, b) in cases.items():
        B = np.diag(a) + np.diag(b, 1)
        out["alpha_" + tag] = a
        out["beta_" + tag] = b
        out["sigma_" + tag] = np.linalg.svd(B, compute_uv=False)
    np.savez(os.path.join(HERE, "bidiag_lapack.npz"), **out)


def qr():
    rng = np.random.default_rng(77)
    out = {}
    for tag, (m, n) in {"tall40x7": (40, 7), "sq9": (9, 9), "tall100x20": (100, 20)}.items():
        A = rng.standard_normal((m, n))
        qr_, tau, work, info = lapack.dgeqrf(A)
        q, work, info = lapack.dorgqr(qr_, tau)
        out["A_" + tag] = A
        out["qr_" + tag] = qr_
        out["tau_" + tag] = tau
        out["q_" + tag] = q
    np.savez(os.path.join(HERE, "qr_lapack.npz"), **out)


def svd():
    rng = np.random.default_rng(0)
    out = {}
    A = rng.standard_normal((1000, 100))  # 2 peers x 500 x 100 row blocks, entries N(0,1)
    U, s, Vt = np.linalg.svd(A, full_matrices=False)
    out["sigma_c1"] = s
    np.savez(os.path.join(HERE, "svd_lapack.npz"), **out)


if __name__ == "__main__":
    bidiag()
    qr()
    svd()
