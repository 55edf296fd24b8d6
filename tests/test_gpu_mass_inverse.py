"""MassMatrix(...).Inverse(freedofs): the inverse NGSolve gives
``BilinearForm(u * v * ds).Assemble().mat.Inverse(freedofs)``, as the
reference's demos use it for preconditioners (e.g.
demos/HelmholtzBurtonMiller.ipynb).  Exact per element for SurfaceL2 (block
diagonal), CG for the others; checked against numpy's dense solve on the
free dofs."""
import numpy as np
import pytest
import torch

import ngbem_amd as nb
from ngbem_amd.spaces import mass_matrix_coo

pytestmark = pytest.mark.gpu

SPACES = [(nb.SurfaceL2, 0), (nb.SurfaceL2, 2), (nb.H1, 1), (nb.H1, 3), (nb.HDivSurface, 0), (nb.HDivSurface, 1),
          (nb.HCurl, 1)]


@pytest.mark.parametrize("cls,order", SPACES, ids=[f"{c.__name__}{o}" for c, o in SPACES])
def test_mass_inverse(gpu, cls, order):
    mesh = nb.icosphere(2)
    sp = cls(mesh, order=order)
    M = nb.MassMatrix(sp)
    r, c, v = mass_matrix_coo(sp, sp)
    D = np.zeros((sp.ndof, sp.ndof))
    np.add.at(D, (r, c), v)
    rng = np.random.default_rng(3)
    x = rng.standard_normal(sp.ndof) + 1j * rng.standard_normal(sp.ndof)
    z = (M.Inverse() * torch.from_numpy(x).to(gpu)).cpu().numpy()
    ref = np.linalg.solve(D, x)
    assert np.linalg.norm(z - ref) <= 1e-11 * np.linalg.norm(ref)
    # on a subset of free dofs: M_ff z_f = x_f, zero elsewhere
    free = rng.random(sp.ndof) < 0.8
    zf = (M.Inverse(freedofs=free) * torch.from_numpy(x).to(gpu)).cpu().numpy()
    ref = np.zeros_like(x)
    ref[free] = np.linalg.solve(D[np.ix_(free, free)], x[free])
    assert np.linalg.norm(zf - ref) <= 1e-11 * np.linalg.norm(ref)
    # many right-hand sides at once
    X = torch.from_numpy(np.stack([x.real, x.imag])).to(gpu)
    Z = (M.Inverse() * X).cpu().numpy()
    assert np.linalg.norm(Z - np.linalg.solve(D, np.stack([x.real, x.imag]).T).T) <= 1e-11 * np.linalg.norm(Z)
