"""HDivSurface / HCurl order 1 on the GPU, through the C-ABI, against the
oracle (same meshes and parameters), and the order-1 EFIE against the Mie
series.

Bars: near-field entries within 1e-12 of the oracle's (max over the block's
largest entry); each low-rank block as close to the exact block as the
oracle's cross approximation (twice its error plus 10 eps); the apply of
identical blocks (forward and transpose) within 1e-13; the potential within
1e-12; the EFIE current within 10% of the oracle's dense-solve error at the
same level.
"""
import numpy as np
import pytest
import torch

import ngbem_amd as nb
from oracle import oracle as O

pytestmark = pytest.mark.gpu

# (name, trial (class, order, oracle kind), test (class, order, oracle kind))
CASES = [
    ("MSL", (nb.HDivSurface, 1, O.HDIV1), (nb.HDivSurface, 1, O.HDIV1)),
    ("MSLC", (nb.HCurl, 1, O.HCURL1), (nb.HCurl, 1, O.HCURL1)),
    ("MDL", (nb.HCurl, 1, O.HCURL1), (nb.HDivSurface, 1, O.HDIV1)),
    ("MDL", (nb.HCurl, 1, O.HCURL1), (nb.HDivSurface, 0, O.HDIV)),
    ("MDL", (nb.HCurl, 0, O.HCURL), (nb.HDivSurface, 1, O.HDIV1)),
]


def _build(case, eps=1e-8, kappa=3.0):
    name, (tc, to, tk), (sc, so, sk) = case
    m = nb.icosphere(2)
    tr, te = tc(m, order=to), sc(m, order=so)
    kw = dict(intorder=4, leafsize=24, eta=2.0, eps=eps)
    print(f"maxwell order-1 case {name} {tr} {te}", flush=True)
    if name == "MSL":
        op = nb.MaxwellSingleLayerPotentialOperator(tr, kappa, **kw)
    elif name == "MSLC":
        op = nb.MaxwellSingleLayerPotentialOperatorCurl(tr, kappa, **kw)
    else:
        op = nb.MaxwellDoubleLayerPotentialOperator(tr, te, kappa, **kw)
    kid = O.MAXWELL_DL if name == "MDL" else O.MAXWELL_SL
    ref = O.OracleOp(m, kid, kappa=kappa, trial=tk, test=sk, **kw)
    return op, ref


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}-{c[1][1]}{c[2][1]}" for c in CASES])
def test_order1_blocks_and_apply_match_oracle(gpu, case):
    eps = 1e-8
    op, ref = _build(case, eps=eps)
    assert (op.trial_space.ndof, op.test_space.ndof) == (ref.ntrial, ref.ntest)
    ob = ref.blocks()
    assert op.num_blocks() == len(ob)
    nfar = 0
    for b, rb in enumerate(ob):
        info = op.block_info(b)
        assert (info["row_cluster"], info["col_cluster"]) == (rb["row"], rb["col"])
        rows, cols = op.block_dofs(b)
        data = op.block_data(b)
        D = ref.dense(cols, rows)
        if info["rank"] < 0:
            assert np.abs(data - D).max() <= 1e-12 * max(np.abs(D).max(), 1e-300), f"block {b}"
            ref.set_block(b, data.ravel(), -1)
        else:
            nfar += 1
            U, V = ref.aca(cols, rows)
            A = data[0] @ data[1]
            err, err_ref = np.linalg.norm(A - D), np.linalg.norm(U @ V - D)
            assert err <= 2 * err_ref + 10 * eps * np.linalg.norm(D), f"far block {b}: {err} vs {err_ref}"
            ref.set_block(b, np.concatenate([data[0].ravel(), data[1].ravel()]), data[0].shape[1])
    assert nfar > 0
    rng = np.random.default_rng(2)
    for trans in (False, True):
        n_in = op.test_space.ndof if trans else op.trial_space.ndof
        x = rng.standard_normal(n_in) + 1j * rng.standard_normal(n_in)
        y_ref = ref.apply(x.astype(np.complex128), trans=trans)
        y = ((op.mat.T if trans else op.mat) * torch.from_numpy(x).to(gpu)).cpu().numpy()
        assert np.linalg.norm(y - y_ref) <= 1e-13 * np.linalg.norm(y_ref)


@pytest.mark.parametrize("which", ["MSL", "MDL"])
def test_order1_potential_matches_oracle(gpu, which):
    case = CASES[0] if which == "MSL" else CASES[2]
    op, _ = _build(case)
    m = op.trial_space.mesh
    kid = O.MAXWELL_SL if which == "MSL" else O.MAXWELL_DL
    ref = O.OracleOp(m, kid, 3.0, trial=case[1][2], test=case[2][2], intorder=4, leafsize=24)
    rng = np.random.default_rng(9)
    c = rng.standard_normal(op.trial_space.ndof) + 1j * rng.standard_normal(op.trial_space.ndof)
    pts = rng.normal(size=(200, 3)) * 2.0
    pts = pts[np.linalg.norm(pts, axis=1) > 1.3]
    nrm = pts / np.linalg.norm(pts, axis=1)[:, None]
    pot = op.GetPotential(torch.from_numpy(c).to(gpu))
    u = pot(torch.from_numpy(pts).to(gpu), torch.from_numpy(nrm).to(gpu)).cpu().numpy().reshape(len(pts), -1)
    u_ref = ref.potential(c, pts, nrm)
    assert u.shape == u_ref.shape == (len(pts), pot.dim)
    assert np.abs(u - u_ref).max() <= 1e-12 * np.abs(u_ref).max()


def test_efie_order1_vs_mie(gpu):
    """EFIE, kappa 2, HDivSurface order 1 at icosphere level 3: the H-matrix
    GMRES current against the Mie current, beside the order-0 current."""
    from test_oracle_maxwell import dense_mass
    from test_oracle_maxwell1 import efie_error, efie_rhs

    k = 2.0
    d, p = np.array([[0.0, 0.0, 1.0], [0.6, 0.0, 0.8]]), np.array([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    mesh = nb.icosphere(3)
    errs = {}
    for order in (0, 1):
        hd = nb.HDivSurface(mesh, order=order)
        V = nb.MaxwellSingleLayerPotentialOperator(hd, k, intorder=5, leafsize=40, eta=2.0, eps=1e-8)
        Minv = torch.linalg.inv(torch.from_numpy(dense_mass(hd, hd)).to(gpu)).to(torch.complex128)
        B = efie_rhs(hd, k, d, p)
        J = np.stack([nb.GMRes(V.mat, torch.from_numpy(B[r]).to(gpu), pre=lambda x: Minv @ x, tol=1e-10,
                               maxsteps=3000, printrates=False).cpu().numpy() for r in range(len(d))])
        errs[order] = efie_error(hd, J, k, d, p)
        print(f"EFIE L3 order {order}: ndof {hd.ndof}, error vs Mie {errs[order]}", flush=True)
    # the oracle pins give 0.143 / 0.068 at level 2 (order 0 / 1)
    assert np.all(errs[0] < 0.08) and np.all(errs[1] < 0.035)
    assert np.all(errs[1] < 0.6 * errs[0])


def test_order1_on_region_matches_oracle(gpu):
    """HDivSurface / HCurl order 1 on part of the mesh (compressed dofs)."""
    from test_gpu_regions import capped_sphere

    m = capped_sphere(3, zcut=-0.3)
    cap = m.Boundaries("cap")
    hd, hc = nb.HDivSurface(m, order=1, definedon=cap), nb.HCurl(m, order=1, definedon=cap)
    kw = dict(intorder=3, leafsize=24, eta=2.0, eps=1e-8)
    for name, op, kid, tk, sk in (
            ("MSL", nb.MaxwellSingleLayerPotentialOperator(hd, 3.0, **kw), O.MAXWELL_SL, O.HDIV1, O.HDIV1),
            ("MDL", nb.MaxwellDoubleLayerPotentialOperator(hc, hd, 3.0, **kw), O.MAXWELL_DL, O.HCURL1, O.HDIV1)):
        print(f"region order-1 {name}: ndof {op.trial_space.ndof}", flush=True)
        ref = O.OracleOp(m, kid, kappa=3.0, trial=tk, test=sk, trial_mask=cap.mask, test_mask=cap.mask, **kw)
        assert (op.trial_space.ndof, op.test_space.ndof) == (ref.ntrial, ref.ntest)
        for b in range(op.num_blocks()):
            d = op.block_data(b)
            if isinstance(d, tuple):
                ref.set_block(b, np.concatenate([d[0].ravel(), d[1].ravel()]), d[0].shape[1])
            else:
                rows, cols = op.block_dofs(b)
                D = ref.dense(cols, rows)
                assert np.abs(d - D).max() <= 1e-12 * max(np.abs(D).max(), 1e-300), f"{name} block {b}"
                ref.set_block(b, d.ravel(), -1)
        x = np.random.default_rng(4).standard_normal(op.trial_space.ndof) + 0j
        y = (op.mat * torch.from_numpy(x).to(gpu)).cpu().numpy()
        y_ref = ref.apply(x)
        assert np.linalg.norm(y - y_ref) <= 1e-13 * np.linalg.norm(y_ref)
