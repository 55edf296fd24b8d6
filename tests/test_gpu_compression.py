"""TestCompression's cross approximation on the GPU against the oracle's."""
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

