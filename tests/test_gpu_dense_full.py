"""method="dense" at the headline size.

The reference maps method="dense" to one block over the whole space at any
size (leafsize and eta chosen so that the root is a leaf,
src/python_bem.cpp:30-33).  On the icosphere of level 6 (81,920 triangles)
the Laplace single layer on SurfaceL2 0 is then one block of 6.7e9 entries
(54 GB in FP64): its element pairs run in chunks numbered per block and in
several launches.  Checked: columns of the dense operator against the
oracle's entries (1e-12), the dense apply against the H-matrix of the same
operator (ACA eps 1e-8), and the transposed apply through the adjoint
identity (y, D x) = (D^T y, x).
"""
import gc

import numpy as np
import pytest
import torch

import ngbem_amd as nb
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_dense_method_level6(gpu):
    mesh = nb.icosphere(6)
    fes = nb.SurfaceL2(mesh, order=0)
    n = fes.ndof
    kw = dict(intorder=4, leafsize=64, eta=2.0)
    print(f"dense Laplace single layer, icosphere L6, {n} dofs ({n * n:.3e} entries)", flush=True)
    D = nb.SingleLayerPotentialOperator(fes, method="dense", **kw)
    assert D.num_blocks() == 1 and D.block_info(0)["rank"] == -1
    assert D.mat.nze == n * n
    rng = np.random.default_rng(0)
    ref = O.OracleOp(mesh, O.LAPLACE_SL, trial=O.P0, test=O.P0, intorder=4, leafsize=10 ** 9, eta=0.0)
    # columns of D through unit vectors, against the oracle's entries
    for j in rng.choice(n, size=3, replace=False):
        e = torch.zeros(n, dtype=torch.float64, device=gpu)
        e[int(j)] = 1.0
        col = (D.mat * e).cpu().numpy()
        col_ref = ref.dense([int(j)], np.arange(n))[:, 0].real
        err = np.abs(col - col_ref).max() / np.abs(col_ref).max()
        print(f"column {j}: max rel err {err:.2e}", flush=True)
        assert err < 1e-12
    H = nb.SingleLayerPotentialOperator(fes, eps=1e-8, **kw)
    x = torch.from_numpy(rng.standard_normal(n)).to(gpu)
    y = torch.from_numpy(rng.standard_normal(n)).to(gpu)
    yd = D.mat * x
    yh = H.mat * x
    rel = float(torch.linalg.norm(yd - yh) / torch.linalg.norm(yd))
    print(f"dense vs H-matrix apply: {rel:.2e}", flush=True)
    assert rel < 1e-6
    lhs = float(torch.dot(y, yd))
    rhs = float(torch.dot(D.mat.T * y, x))
    print(f"adjoint identity: {abs(lhs - rhs) / abs(lhs):.2e}", flush=True)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    del D, H
    gc.collect()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
