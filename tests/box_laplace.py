"""Interior Laplace Dirichlet problem on the unit box with an exact solution
(test helper): u = 1 / |x - x0| with x0 outside the box is harmonic inside,
so its Neumann trace t = du/dn solves V t = (1/2 M + K) g with g = u on the
boundary (the interior Calderon identity in the reference's sign convention,
K 1 = -1/2; the reference's Laplace tests solve the same first-kind equation,
tests/test_Laplace.py).  Flat triangles represent the box exactly, so the
error of t_h in L2 on the surface reflects only the spaces: O(h^p) for H1
order p and SurfaceL2 order p - 1.  `neumann_error` takes the two operator
matrices (dense, numpy) from whoever assembled them."""
import numpy as np

import ngbem_amd as nb

X0 = np.array([1.2, 0.9, 0.7])


def u(X):
    return 1.0 / np.linalg.norm(X - X0, axis=-1)


def dudn(X, n):
    d = X - X0
    return -np.einsum("...c,c->...", d, n) / np.linalg.norm(d, axis=-1) ** 3


def rule(k):
    """Collapsed Gauss rule on the reference triangle (weights sum to 1/2)."""
    g, w = np.polynomial.legendre.leggauss(k)
    g, w = 0.5 * (g + 1), 0.5 * w
    U, V = np.meshgrid(g, g, indexing="ij")
    W = (w[:, None] * w[None]) * (1 - U)
    x, y = U.ravel(), (V * (1 - U)).ravel()
    return np.stack([x, y, 1 - x - y], 1), W.ravel()


def spaces(n, p):
    m = nb.box(n)
    return m, nb.H1(m, order=p), nb.SurfaceL2(m, order=p - 1)


def _mass(tr, te):
    r, c, v = nb.spaces.mass_matrix_coo(tr, te)
    M = np.zeros((te.ndof, tr.ndof))
    np.add.at(M, (r, c), v)
    return M


def neumann_error(m, h1, l2, V, K):
    """L2 error of the Neumann trace from the dense V (l2 x l2) and K (l2 x h1)."""
    bary, w = rule(8)
    P = m.xyz[m.tri]
    J = 2 * m.areas()
    nrm = m.normals()
    sg, ph = h1.elem_signs(), h1.shapes(bary)
    b = np.zeros(h1.ndof)
    for t in range(m.nt):
        X = bary @ P[t]
        np.add.at(b, h1.elem_dofs[t], J[t] * np.einsum("q,q,qa->a", w, u(X), ph * sg[t]))
    g = np.linalg.solve(_mass(h1, h1), b)
    tN = np.linalg.solve(V, 0.5 * _mass(h1, l2) @ g + K @ g)
    pl = l2.shapes(bary)
    err = 0.0
    for t in range(m.nt):
        X = bary @ P[t]
        err += J[t] * np.sum(w * (pl @ tN[l2.elem_dofs[t]] - dudn(X, nrm[t])) ** 2)
    return np.sqrt(err), tN
