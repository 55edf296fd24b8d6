"""TestCompression's SVD variants (LAPACK through torch on the host, like the
reference's LapackSVD) against numpy's singular values of the same matrix;
the point clouds come from the library's host entry ``bem_test_points``."""
import numpy as np

import ngbem_amd as nb
from oracle import oracle as O


def test_svd_and_tsvd():
    X, Y = O.test_points(150, 120, 2.0)
    S = np.linalg.svd(1.0 / np.linalg.norm(X[:, None] - Y[None], axis=-1), compute_uv=False)
    for eps in (1e-4, 1e-8):
        r, err, _ = nb.TestCompression(150, 120, 2.0, eps, "svd")
        assert r == int((S > eps).sum())
        assert abs(err - np.sqrt((S[r:] ** 2).sum())) <= 1e-10 + 1e-6 * err
        rt, errt, _ = nb.TestCompression(150, 120, 2.0, eps, "tsvd")
        assert abs(rt - r) <= 2 and errt <= 10 * eps * np.linalg.norm(S)
    assert nb.TestCompression(10, 10, 2.0, 1e-6, "other") == (0, 0.0)
