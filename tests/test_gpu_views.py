"""Vectors given as torch's lazy views (x.conj(), x.imag of a conjugate,
non-contiguous slices) are read and written as the values they show: the
library works on raw device memory, so such views are materialised first."""
import numpy as np
import pytest
import torch

import ngbem_amd as nb

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ops():
    mesh = nb.icosphere(2)
    fes = nb.SurfaceL2(mesh)
    kw = dict(intorder=3, leafsize=16, eps=1e-8)
    return (nb.HelmholtzCombinedFieldOperator(fes, fes, kappa=3.0, **kw), nb.SingleLayerPotentialOperator(fes, **kw))


def test_conjugate_and_negative_views(gpu, ops):
    C, V = ops
    n = C.trial_space.ndof
    g = torch.Generator(device="cpu").manual_seed(0)
    x = torch.complex(torch.randn(n, generator=g, dtype=torch.float64),
                      torch.randn(n, generator=g, dtype=torch.float64)).to(gpu)
    xc = x.conj()                        # lazy conjugate view
    assert xc.is_conj()
    ref = C.mat * xc.resolve_conj()
    assert torch.equal(C.mat * xc, ref)
    assert torch.equal(C.mat.T * xc, C.mat.T * xc.resolve_conj())
    # (A^H)^T = conj(A): applied through a conjugate view inside AdjointMatrix
    assert torch.allclose(C.mat.H.T * x, (C.mat * x.conj()).conj(), rtol=0, atol=1e-13 * float(ref.abs().max()))
    # a real operator on the imaginary part of a conjugate (a negative view)
    xi = xc.imag
    assert torch.equal(V.mat * xi, V.mat * xi.clone())
    # output into a conjugate view: y.conj() += A x must land as conj(A x) in y
    y = torch.zeros(n, dtype=torch.complex128, device=gpu)
    C.mat.MultAdd(1.0, x, y.conj())
    assert torch.allclose(y, (C.mat * x).conj(), rtol=0, atol=0)
    # non-contiguous input
    X = torch.stack([x, 2 * x], 1)
    assert torch.equal(C.mat * X[:, 1], C.mat * (2 * x))


def test_matrix_algebra_adjoints(gpu, ops):
    """Scaled, summed, transposed and adjoint combinations against their
    definitions: (c A + B)^H x = conj(c) A^H x + B^H x, and (M^H)^T = conj(M)."""
    C, V = ops
    n = C.trial_space.ndof
    g = torch.Generator(device="cpu").manual_seed(1)
    x = torch.complex(torch.randn(n, generator=g, dtype=torch.float64),
                      torch.randn(n, generator=g, dtype=torch.float64)).to(gpu)
    c = 0.3 - 2.0j
    M = c * C.mat + V.mat
    ref = np.conj(c) * (C.mat.H * x) + V.mat.T * x
    tol = 1e-13 * float(ref.abs().max())
    assert torch.allclose(M.H * x, ref, rtol=0, atol=tol)
    assert torch.allclose(M.T * x, c * (C.mat.T * x) + V.mat.T * x, rtol=0, atol=tol)
    assert torch.allclose(M.H.T * x, (M * x.conj()).conj(), rtol=0, atol=tol)
