"""TestCompression on the GPU against the oracle (aca) and numpy (svd)."""
import numpy as np
import pytest

import ngbem_amd as nb
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nx,ny,eps", [(100, 100, 1e-10), (300, 200, 1e-6), (64, 500, 1e-4)])
def test_aca_matches_oracle(gpu, nx, ny, eps):
    r, err, sec = nb.TestCompression(nx, ny, 2.0, eps, "aca")
    r_ref, err_ref = O.test_compression_aca(nx, ny, 2.0, eps)
    print(f"aca {nx}x{ny} eps {eps}: rank {r} (oracle {r_ref}), err {err:.3e} (oracle {err_ref:.3e}), {sec:.4f} s",
          flush=True)
    assert r == r_ref
    assert abs(err - err_ref) <= 1e-6 * err_ref + 1e-14


def test_svd_and_tsvd(gpu):
    X, Y = O.test_points(150, 120, 2.0)
    S = np.linalg.svd(1.0 / np.linalg.norm(X[:, None] - Y[None], axis=-1), compute_uv=False)
    for eps in (1e-4, 1e-8):
        r, err, _ = nb.TestCompression(150, 120, 2.0, eps, "svd")
        assert r == int((S > eps).sum())
        assert abs(err - np.sqrt((S[r:] ** 2).sum())) <= 1e-10 + 1e-6 * err
        rt, errt, _ = nb.TestCompression(150, 120, 2.0, eps, "tsvd")
        assert abs(rt - r) <= 2 and errt <= 10 * eps * np.linalg.norm(S)
    assert nb.TestCompression(10, 10, 2.0, 1e-6, "other") == (0, 0.0)
