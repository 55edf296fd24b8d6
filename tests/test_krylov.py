"""Krylov drivers on CPU tensors with a dense stand-in operator."""
import numpy as np
import torch

import ngbem_amd.krylov as K
from ngbem_amd.operators import BaseMatrix


class Dense(BaseMatrix):
    def __init__(self, A):
        self.A = A
        self.height, self.width = A.shape
        self.is_complex = torch.is_complex(A)

    def MultAdd(self, s, x, y):
        y.add_(s * (x @ self.A.T))


def _system(n=60, seed=0):
    rng = np.random.default_rng(seed)
    A = np.eye(n) * 3 + (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    return torch.from_numpy(A), rng


def test_gmres_batch_matches_direct_solve():
    A, rng = _system()
    B = torch.from_numpy(rng.standard_normal((7, 60)) + 1j * rng.standard_normal((7, 60)))
    X_ref = torch.linalg.solve(A, B.T).T
    for kw in (dict(), dict(restart=10, pre=lambda r: r / 3.0)):
        X = K.GMResBatch(Dense(A), B, tol=1e-12, maxsteps=200, **kw)
        assert (torch.linalg.norm(X - X_ref) / torch.linalg.norm(X_ref)).item() < 1e-10
        assert np.all(K.GMResBatch.residuals <= 1e-12)


def test_gmres_batch_equals_single_gmres():
    A, rng = _system(seed=1)
    B = torch.from_numpy(rng.standard_normal((3, 60)) + 0j)
    X = K.GMResBatch(Dense(A), B, tol=1e-13, maxsteps=100)
    for r in range(3):
        x = K.GMRes(Dense(A), B[r], tol=1e-13, maxsteps=100, printrates=False)
        assert (torch.linalg.norm(X[r] - x) / torch.linalg.norm(x)).item() < 1e-11


def test_cg_and_refined_solve():
    rng = np.random.default_rng(2)
    Q = rng.standard_normal((40, 40))
    A = torch.from_numpy(Q @ Q.T + 40 * np.eye(40))
    b = torch.from_numpy(rng.standard_normal(40))
    x = K.CG(Dense(A), b, tol=1e-12, maxsteps=200, printrates=False)
    assert (torch.linalg.norm(A @ x - b) / torch.linalg.norm(b)).item() < 1e-11
    A32 = Dense(A.to(torch.float32).to(torch.float64))
    x2, hist = K.RefinedSolve(Dense(A), A32, b, tol=1e-12, inner_tol=1e-4)
    assert hist[-1] <= 1e-12


def test_gmres_left_preconditioner_drops_fixed_rows():
    """A preconditioner that is zero on some dofs (a mass inverse on the free
    dofs) makes GMRes solve the free rows for the free unknowns and leave the
    fixed unknowns at zero, as ngsolve's left-preconditioned GMRes does; a
    right-preconditioned iteration would instead minimise the residual of all
    rows and miss the free-rows solution."""
    A, rng = _system(n=50, seed=3)
    b = torch.from_numpy(rng.standard_normal(50) + 1j * rng.standard_normal(50))
    free = torch.from_numpy(rng.random(50) < 0.6)
    fi = torch.nonzero(free).reshape(-1)
    ref = torch.zeros_like(b)
    ref[fi] = torch.linalg.solve(A[fi][:, fi], b[fi])
    D = torch.diag(free.to(torch.complex128) / torch.from_numpy(1.0 + rng.random(50)))
    for pre, kw in ((lambda r: r @ D.T, {}), (Dense(D), dict(restart=8))):
        x = K.GMRes(Dense(A), b, pre=pre, tol=1e-13, maxsteps=300, printrates=False, **kw)
        assert (torch.linalg.norm(x - ref) / torch.linalg.norm(ref)).item() < 1e-11
        assert torch.all(x[~free] == 0)
        X = K.GMResBatch(Dense(A), torch.stack([b, 2 * b]), pre=pre, tol=1e-13, maxsteps=300, **kw)
        assert (torch.linalg.norm(X[1] - 2 * ref) / torch.linalg.norm(2 * ref)).item() < 1e-11
        assert torch.all(X[:, ~free] == 0)
