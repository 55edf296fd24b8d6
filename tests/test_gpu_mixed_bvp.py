"""The mixed Dirichlet / Neumann Laplace problem of the reference's
tests/Laplace_Convergence.py (order 1 row): unit sphere, Dirichlet data on
the lower half, Neumann data on the upper half, exact solution
u = 1 / |x - (1, 1, 1)| (harmonic inside).  The Calderon system

    [ V   -K ] [ t_D ]   [ (1/2 M + K) u_D - V t_N   ]
    [ K^T  W ] [ u_N ] = [ (1/2 M^T - K^T) t_N - W u_D ]

is solved by GMRES with the block-diagonal mass-inverse preconditioner on the
free dofs (H1 with dirichlet="dirichlet", SurfaceL2 with dirichlet="neumann"),
as the reference builds it from BlockMatrix, FreeDofs and mat.Inverse.  The
trace errors fall at O(h^2) (Dirichlet, H1 p1) and O(h) (Neumann, SurfaceL2 p0),
and the representation formula V t - K u reproduces u inside the sphere."""
import numpy as np
import pytest
import torch

import ngbem_amd as nb
from oracle import oracle as O

pytestmark = pytest.mark.gpu

X0 = np.array([1.0, 1.0, 1.0])


def uexa(X):
    return 1.0 / np.linalg.norm(X - X0, axis=-1)


def dn_uexa(X, n):
    D = X - X0
    return -(D * n).sum(-1) / np.linalg.norm(D, axis=-1) ** 3


def split_sphere(level):
    m = nb.icosphere(level)
    c = m.xyz[m.tri].mean(axis=1)
    return nb.Mesh(m.xyz, m.tri, np.where(c[:, 2] < 0, 1, 2), {1: "dirichlet", 2: "neumann"})


def solve(level, gpu, parity=False):
    mesh = split_sphere(level)
    h1 = nb.H1(mesh, order=1, dirichlet="dirichlet")
    l2 = nb.SurfaceL2(mesh, order=0, dirichlet="neumann")
    dir_el = mesh.Boundaries("dirichlet").mask
    # u_D: vertex values on the Dirichlet triangles; t_N: element means of du/dn on the Neumann ones
    ud = np.zeros(h1.ndof)
    vd = np.unique(mesh.tri[dir_el])
    ud[vd] = uexa(mesh.xyz[vd])
    rule = O.trig_rule(8)
    bary = np.stack([rule[:, 0], rule[:, 1], 1 - rule[:, 0] - rule[:, 1]], 1)
    X = np.einsum("qk,tkc->tqc", bary, mesh.xyz[mesh.tri])
    n = mesh.normals()[:, None, :]
    un = np.where(~dir_el, (rule[:, 2][None] * dn_uexa(X, n)).sum(1) / rule[:, 2].sum(), 0.0)
    kw = dict(intorder=6, leafsize=40, eps=1e-10)
    V = nb.SingleLayerPotentialOperator(l2, **kw)
    K = nb.DoubleLayerPotentialOperator(h1, l2, **kw)
    W = nb.HypersingularOperator(h1, **kw)
    M = nb.MassMatrix(h1, l2)
    udv, unv = torch.from_numpy(ud).to(gpu), torch.from_numpy(un).to(gpu)
    fd = (0.5 * M + K.mat) * udv - V.mat * unv
    fn = (0.5 * M.T - K.mat.T) * unv - W.mat * udv
    lhs = nb.BlockMatrix([[V.mat, -K.mat], [K.mat.T, W.mat]])
    rhs = nb.BlockVector([fd, fn])
    pre = nb.BlockMatrix([[nb.MassMatrix(l2).Inverse(freedofs=l2.FreeDofs()), None],
                          [None, nb.MassMatrix(h1).Inverse(freedofs=h1.FreeDofs())]])
    sol = nb.GMRes(lhs, rhs, pre=pre, tol=1e-12, maxsteps=2000, printrates=False).cpu().numpy()
    tD, uN = sol[:l2.ndof], sol[l2.ndof:]
    if parity:
        # the same system from the oracle's dense matrices, solved directly on the free rows
        from ngbem_amd.spaces import mass_matrix_coo
        Vd = O.OracleOp(mesh, O.LAPLACE_SL, 0.0, trial=O.P0, test=O.P0, intorder=6).dense().real
        Kd = O.OracleOp(mesh, O.LAPLACE_DL, 0.0, trial=O.P1, test=O.P0, intorder=6).dense().real
        Wd = O.OracleOp(mesh, O.LAPLACE_HS, 0.0, trial=O.P1, test=O.P1, intorder=6).dense().real
        r, c, v = mass_matrix_coo(h1, l2)
        Md = np.zeros((l2.ndof, h1.ndof))
        np.add.at(Md, (r, c), v)
        Ad = np.block([[Vd, -Kd], [Kd.T, Wd]])
        bd = np.concatenate([(0.5 * Md + Kd) @ ud - Vd @ un, (0.5 * Md.T - Kd.T) @ un - Wd @ ud])
        free = np.concatenate([l2.FreeDofs(), h1.FreeDofs()])
        ref = np.zeros_like(bd)
        ref[free] = np.linalg.solve(Ad[np.ix_(free, free)], bd[free])
        assert np.linalg.norm(sol - ref) / np.linalg.norm(ref) < 1e-8
    assert np.abs(tD[~l2.FreeDofs()]).max() == 0.0 and np.abs(uN[~h1.FreeDofs()]).max() == 0.0
    # trace errors over the whole sphere
    w = rule[:, 2][None, :] * 2 * mesh.areas()[:, None]
    u_h = (ud + uN)[mesh.tri] @ bary.T                  # P1 trace at the points
    t_h = (un + tD)[:, None]                            # P0 flux
    errd = np.sqrt(np.sum(w * (uexa(X) - u_h) ** 2))
    errn = np.sqrt(np.sum(w * (dn_uexa(X, n) - t_h) ** 2))
    # representation formula inside: u = V t - K u
    P = np.array([[0.1, 0.2, -0.1], [0.3, -0.2, 0.25], [-0.35, 0.1, 0.0], [0.0, 0.0, 0.4]])
    Pt = torch.from_numpy(P).to(gpu)
    pot = (V.GetPotential(torch.from_numpy(un + tD).to(gpu))(Pt)
           - K.GetPotential(torch.from_numpy(ud + uN).to(gpu))(Pt)).cpu().numpy().reshape(-1)
    errp = np.abs(pot - uexa(P)).max() / np.abs(uexa(P)).max()
    print(f"level {level}: ndof L2 {l2.ndof} H1 {h1.ndof}  err_dir {errd:.3e}  err_neu {errn:.3e}  "
          f"err_pot {errp:.3e}", flush=True)
    return errd, errn, errp


def test_mixed_dirichlet_neumann_convergence(gpu):
    e2, e3 = solve(2, gpu, parity=True), solve(3, gpu)
    assert e2[0] < 0.02 and e2[1] < 0.15 and e2[2] < 1e-3
    assert e3[0] < 0.35 * e2[0]          # Dirichlet trace, H1 p1: O(h^2)
    assert e3[1] < 0.6 * e2[1]           # Neumann trace, SurfaceL2 p0: O(h)
    assert e3[2] < 0.5 * e2[2] and e3[2] < 1e-2
