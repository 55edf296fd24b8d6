"""Dirichlet dof flags and block systems on the CPU: ``FreeDofs`` against the
mesh's region membership, ``BlockMatrix`` / ``BlockVector`` against the dense
block matrix (NGSolve's classes of the same names, as the reference's
tests/Laplace_Convergence.py:27-61 builds its Calderon system)."""
import numpy as np
import torch

import ngbem_amd as nb
from ngbem_amd.operators import BaseMatrix


class Dense(BaseMatrix):
    def __init__(self, A):
        self.A = A
        self.height, self.width = A.shape
        self.is_complex = torch.is_complex(A)

    def MultAdd(self, s, x, y):
        y.add_(s * (x @ self.A.T))

    def MultTransAdd(self, s, x, y):
        y.add_(s * (x @ self.A))


def _split_sphere(level):
    m = nb.icosphere(level)
    c = m.xyz[m.tri].mean(axis=1)
    return nb.Mesh(m.xyz, m.tri, np.where(c[:, 2] < 0, 1, 2), {1: "dirichlet", 2: "neumann"})


def test_free_dofs_follow_the_dirichlet_region():
    mesh = _split_sphere(2)
    dir_el = mesh.Boundaries("dirichlet").mask
    h1 = nb.H1(mesh, order=1, dirichlet="dirichlet")
    # H1 p1: a vertex is fixed iff it belongs to a Dirichlet triangle
    fixed_v = np.zeros(mesh.xyz.shape[0], dtype=bool)
    fixed_v[mesh.tri[dir_el].reshape(-1)] = True
    assert np.array_equal(h1.FreeDofs(), ~fixed_v)
    # SurfaceL2 p0 fixed on the Neumann triangles; order 1: three dofs each
    for order in (0, 1):
        l2 = nb.SurfaceL2(mesh, order=order, dirichlet="neumann")
        f = l2.FreeDofs()
        assert f.sum() == dir_el.sum() * l2.elem_dofs.shape[1]
        assert np.all(f[l2.elem_dofs[dir_el]]) and not np.any(f[l2.elem_dofs[~dir_el]])
    # no flags: all free; a Region works like its name
    assert nb.H1(mesh, order=2).FreeDofs().all()
    h2 = nb.H1(mesh, order=2, dirichlet=mesh.Boundaries("dirichlet"))
    assert np.array_equal(h2.FreeDofs(), nb.H1(mesh, order=2, dirichlet="dirichlet").FreeDofs())


def test_block_matrix_matches_dense_blocks():
    rng = np.random.default_rng(0)
    A, B, C = (torch.from_numpy(rng.standard_normal(s)) for s in ((5, 5), (5, 3), (3, 5)))
    Mb = nb.BlockMatrix([[Dense(A), Dense(B)], [Dense(C), None]])
    Md = torch.cat([torch.cat([A, B], 1), torch.cat([C, torch.zeros(3, 3, dtype=A.dtype)], 1)], 0)
    assert (Mb.height, Mb.width) == (8, 8)
    x = nb.BlockVector([torch.from_numpy(rng.standard_normal(5)), torch.from_numpy(rng.standard_normal(3))])
    assert x.shape == (8,)
    y = torch.zeros(8, dtype=torch.float64)
    Mb.MultAdd(2.0, x, y)
    assert torch.allclose(y, 2.0 * Md @ x, rtol=1e-14, atol=1e-14)
    yt = torch.zeros(8, dtype=torch.float64)
    Mb.MultTransAdd(1.0, x, yt)
    assert torch.allclose(yt, Md.T @ x, rtol=1e-14, atol=1e-14)
    # a rectangular block layout and the size checks
    R = nb.BlockMatrix([[Dense(B), None], [None, Dense(C)]])
    assert (R.height, R.width) == (8, 8)
    for bad in ([[Dense(A), Dense(C)]], [[None, None]]):
        try:
            nb.BlockMatrix(bad)
        except ValueError:
            continue
        raise AssertionError("inconsistent block sizes accepted")
