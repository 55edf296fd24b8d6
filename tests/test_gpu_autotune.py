"""The apply's autotuner and the hipGraph replay of one operator's apply.

Autotune: the report lists every candidate with its time and the choices;
after tuning the FP64 apply equals the untuned one to 1e-13 (the work shapes
change the summation order only) and the FP32 copy to its own rounding;
transposed applies are untouched.  Graph replay: bit-identical to the direct
launches on every replay, with other GPU work between the replays, for the
forward, transposed, adjoint and mixed applies, and again after a tuning run
replaced the tables the graphs read (they are recaptured)."""
import numpy as np
import pytest
import torch

import ngbem_amd as nb

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cf_op(gpu):
    fes = nb.SurfaceL2(nb.icosphere(4))
    return nb.HelmholtzCombinedFieldOperator(fes, fes, kappa=6.0, intorder=4, leafsize=40, eps=1e-6, mixed=True)


def test_autotune_keeps_the_apply(gpu, cf_op):
    C = cf_op
    g = torch.Generator(device="cpu").manual_seed(5)
    x = torch.randn(C.mat.width, dtype=torch.complex128, generator=g).to(gpu)
    y64, y32, yt = C.mat * x, C.mat.mixed * x, C.mat.T * x
    rep = C.Autotune(iters=5)
    print(rep, flush=True)
    for key in ("fp64", "fp32"):
        r = rep[key]
        assert len(r["far_t"]) >= 5 and len(r["rows"]) >= 5
        assert r["far_t_choice"] in [c["rows8_up_to_len"] for c in r["far_t"]]
        assert r["rows_choice"] in [c["columns_per_piece"] for c in r["rows"]]
        best = min(c["ms"] for c in r["rows"])
        assert [c["ms"] for c in r["rows"] if c["columns_per_piece"] == r["rows_choice"]][0] == best
        assert r["apply_ms_after"] > 0
        # the persistent grids run the same pieces: bit-identical outputs
        assert len(r["rows_grid"]) >= 4 and r["rows_grid_identical"] is True
        assert r["rows_grid_choice"] in [c["workgroups"] for c in r["rows_grid"]]
    rel = lambda a, b: (torch.linalg.norm(a - b) / torch.linalg.norm(b)).item()  # noqa: E731
    assert rel(C.mat * x, y64) <= 1e-13
    assert rel(C.mat.mixed * x, y32) <= 1e-6
    assert torch.equal(C.mat.T * x, yt)


def test_graph_replay_equals_direct(gpu, cf_op):
    C = cf_op
    g = torch.Generator(device="cpu").manual_seed(6)
    views = {"fwd": (C.mat, C.mat.graph), "trans": (C.mat.T, C.mat.graph.T), "adj": (C.mat.H, C.mat.graph.H),
             "mixed": (C.mat.mixed, C.mat.mixed.graph)}
    for rep in range(4):
        if rep == 2:
            C.Autotune(iters=3)  # replaces the tables: graphs are dropped and recaptured
        x = torch.randn(C.mat.width, dtype=torch.complex128, generator=g).to(gpu)
        for name, (direct, graph) in views.items():
            ref = direct * x
            torch.cuda.synchronize()
            _ = torch.randn(2048, 2048, device=gpu).sum()  # other work between replays
            y = graph * x
            y2 = torch.randn(C.mat.height, dtype=torch.complex128, generator=g).to(gpu)
            y2c = y2.clone()
            graph.MultAdd(0.5 + 2.0j, x, y2)
            torch.cuda.synchronize()
            assert torch.equal(y, ref), (rep, name)
            assert (torch.linalg.norm(y2 - (y2c + (0.5 + 2.0j) * ref)) / torch.linalg.norm(y2)).item() < 1e-15


def test_gmres_through_graph_view(gpu, cf_op):
    C = cf_op
    fes = C.trial_space
    area = torch.from_numpy(fes.mesh.areas()).to(gpu)
    M = nb.MassMatrix(fes)
    b = torch.randn(fes.ndof, dtype=torch.complex128, device=gpu)
    u = nb.GMRes(0.5 * M + C.mat, b, pre=lambda r: r / area, tol=1e-10, maxsteps=200, printrates=False)
    ug = nb.GMRes(0.5 * M + C.mat.graph, b, pre=lambda r: r / area, tol=1e-10, maxsteps=200, printrates=False)
    assert torch.equal(u, ug)
