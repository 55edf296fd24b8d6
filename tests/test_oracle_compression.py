"""TestCompression pins (src/test_compression.cpp) on the CPU.

- The clouds come from the C library's rand() sequence with seed 1, which is
  what the reference's first call in a process sees: its first values are the
  published glibc ones (1804289383, 846930886, 1681692777, ... / RAND_MAX).
- The restated CalcACA is a cross approximation: the residual M - U V vanishes
  on every pivot row and column (to rounding), its error is never below the
  SVD's at the same rank (Eckart-Young), and it falls with eps.
"""
import numpy as np

from oracle import oracle as O


def test_points_follow_the_c_library_sequence():
    X, Y = O.test_points(3, 2, 2.0, seed=1)
    first = np.array([1804289383, 846930886, 1681692777]) / 2147483647.0
    assert np.allclose(X[0], first, rtol=0, atol=1e-15)
    assert np.all(Y[:, 0] >= 1.0 + np.sqrt(3.0) / 2.0) and np.all(Y[:, 1:] <= 1.0)


def test_aca_is_a_cross_approximation():
    nx, ny, eta = 120, 90, 2.0
    X, Y = O.test_points(nx, ny, eta)
    M = 1.0 / np.linalg.norm(X[:, None] - Y[None], axis=-1)
    S = np.linalg.svd(M, compute_uv=False)
    prev = None
    for eps in (1e-3, 1e-6, 1e-9):
        r, err, U, V, pi, pj = O.test_compression_aca(nx, ny, eta, eps, factors=True)
        R = M - U @ V
        assert np.abs(R[pi[:-1], :]).max() < 1e-12 * np.abs(M).max()  # all pivot rows but the last row pivot
        assert np.abs(R[:, pj]).max() < 1e-12 * np.abs(M).max()
        assert abs(np.linalg.norm(R) - err) <= 1e-12 * np.linalg.norm(M)
        assert err >= np.sqrt((S[r:] ** 2).sum()) * (1 - 1e-9)
        assert err <= 10 * eps * np.linalg.norm(M)
        if prev is not None:
            assert r > prev[0] and err < prev[1]
        prev = (r, err)
