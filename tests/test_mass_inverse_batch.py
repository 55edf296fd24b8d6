"""Batched mass inverse with a zero right-hand side (CPU tensors): the
Jacobi-CG of the H1 / edge-space mass inverse stops per row, so a zero row
(converged from the start) neither divides 0 by 0 nor keeps the others
iterating into NaN; GMResBatch with that inverse as preconditioner solves the
nonzero rows and returns zero for the zero row."""
import numpy as np
import pytest
import torch

import ngbem_amd as nb
import ngbem_amd.krylov as K
import ngbem_amd.operators as ops
from ngbem_amd.spaces import mass_matrix_coo


@pytest.fixture
def cpu_device(monkeypatch):
    monkeypatch.setattr(ops, "device", lambda: torch.device("cpu"))


def _dense_mass(sp):
    r, c, v = mass_matrix_coo(sp, sp)
    D = np.zeros((sp.ndof, sp.ndof))
    np.add.at(D, (r, c), v)
    return D


SPACES = [(nb.H1, 1), (nb.H1, 3), (nb.HDivSurface, 0), (nb.SurfaceL2, 1)]


@pytest.mark.parametrize("cls,order", SPACES, ids=[f"{c.__name__}{o}" for c, o in SPACES])
def test_batched_inverse_with_zero_row(cpu_device, cls, order):
    sp = cls(nb.icosphere(2), order=order)
    Minv = nb.MassMatrix(sp).Inverse()
    D = _dense_mass(sp)
    x = np.random.default_rng(5).standard_normal(sp.ndof)
    X = torch.from_numpy(np.stack([x, np.zeros_like(x), 3 * x]))
    Z = (Minv * X).numpy()
    assert np.isfinite(Z).all()
    assert np.all(Z[1] == 0)
    ref = np.linalg.solve(D, x)
    for row, s in ((0, 1.0), (2, 3.0)):
        assert np.linalg.norm(Z[row] - s * ref) <= 1e-11 * np.linalg.norm(s * ref)


class _Dense(ops.BaseMatrix):
    def __init__(self, A):
        self.A = A
        self.height, self.width = A.shape
        self.is_complex = torch.is_complex(A)

    def MultAdd(self, s, x, y):
        y.add_(s * (x.to(self.A.dtype) @ self.A.T))


def test_gmres_batch_with_mass_preconditioner_and_zero_rhs(cpu_device):
    sp = nb.H1(nb.icosphere(2), order=1)
    D = _dense_mass(sp)
    n = sp.ndof
    rng = np.random.default_rng(6)
    A = torch.from_numpy(D @ (np.eye(n) * 2 + rng.standard_normal((n, n)) / (4 * np.sqrt(n))))
    b = torch.from_numpy(rng.standard_normal(n))
    B = torch.stack([b, torch.zeros_like(b)])
    X = K.GMResBatch(_Dense(A), B, pre=nb.MassMatrix(sp).Inverse(), tol=1e-10, maxsteps=200)
    assert torch.isfinite(X).all()
    assert torch.all(X[1] == 0)
    ref = torch.linalg.solve(A, b)
    assert (torch.linalg.norm(X[0] - ref) / torch.linalg.norm(ref)).item() < 1e-8
    assert np.all(np.isfinite(K.GMResBatch.residuals))


@pytest.mark.parametrize("free", [None, "some"])
def test_order0_inverse_is_the_diagonal(cpu_device, free):
    """SurfaceL2 order 0: the inverse is the diagonal 1 / |T| (zero on fixed
    dofs), exposed as ``_diagonal`` for GMRES's fused preconditioner and
    applied as one product; it equals the dense inverse on the free dofs."""
    sp = nb.SurfaceL2(nb.icosphere(2), order=0)
    fd = None
    mask = np.ones(sp.ndof, dtype=bool)
    if free == "some":
        mask[::7] = False
        fd = mask
    Minv = nb.MassMatrix(sp).Inverse(freedofs=fd)
    assert Minv._diagonal is not None and Minv._diagonal.dtype == torch.float64
    area = sp.mesh.areas()
    np.testing.assert_array_equal(Minv._diagonal.numpy(), np.where(mask, 1.0 / area, 0.0))
    x = np.random.default_rng(3).standard_normal(sp.ndof) + 1j * np.random.default_rng(4).standard_normal(sp.ndof)
    z = (Minv * torch.from_numpy(x)).numpy()
    D = _dense_mass(sp)
    ref = np.zeros_like(x)
    ref[mask] = np.linalg.solve(D[np.ix_(mask, mask)], x[mask])
    assert np.abs(z - ref).max() <= 1e-13 * np.abs(ref).max()
    assert nb.MassMatrix(nb.SurfaceL2(nb.icosphere(1), order=1)).Inverse()._diagonal is None
