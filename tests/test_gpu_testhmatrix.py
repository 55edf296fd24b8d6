"""testhmatrix=True (src/ngbem.cpp:332-369): the factories print the
compression rate and the spectral-norm error of the H-matrix against the
dense operator, estimated by 30 power steps on (D - H)^H (D - H).  Checked
against numpy's 2-norm of the densified difference (the power estimate is a
lower bound that is close after 30 steps) and against nze / (m n)."""
import re

import numpy as np
import pytest
import torch

import ngbem_amd as nb

pytestmark = pytest.mark.gpu


def _dense_of(M, n, gpu, dtype):
    E = torch.eye(n, dtype=dtype, device=gpu)
    return torch.stack([M * E[i] for i in range(n)], 1).cpu().numpy()


@pytest.mark.parametrize("which", ["LSL", "HCF"])
def test_testhmatrix_report(gpu, capsys, which):
    mesh = nb.icosphere(2)
    fes = nb.SurfaceL2(mesh)
    kw = dict(intorder=3, leafsize=16, eta=2.0, eps=1e-4)
    if which == "LSL":
        op = nb.SingleLayerPotentialOperator(fes, testhmatrix=True, **kw)
        dense = nb.SingleLayerPotentialOperator(fes, method="dense", intorder=3)
        dt = torch.float64
    else:
        op = nb.HelmholtzCombinedFieldOperator(fes, fes, kappa=3.0, testhmatrix=True, **kw)
        dense = nb.HelmholtzCombinedFieldOperator(fes, fes, kappa=3.0, method="dense", intorder=3)
        dt = torch.complex128
    out = capsys.readouterr().out
    m = re.search(r"compression rate ([0-9.eE+-]+)\s+2-norm error ([0-9.eE+-]+)", out)
    assert m, out
    rate, err = float(m.group(1)), float(m.group(2))
    n = fes.ndof
    H = _dense_of(op.mat, n, gpu, dt)
    D = _dense_of(dense.mat, n, gpu, dt)
    ref = np.linalg.norm(D - H, 2)
    assert 0 < ref and 0.8 * ref <= err <= 1.0000001 * ref
    assert err < 1e-2 * np.linalg.norm(D, 2)
    assert rate == pytest.approx(op.stats()["nze"] / n ** 2)
