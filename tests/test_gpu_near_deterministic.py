"""The near field of spaces with several functions per element (H1, edge
spaces, SurfaceL2 of order >= 1, the wide evaluators) is summed in a fixed
order: every element pair of a block's patch product stores its element
matrix in its own slot and a gather adds, per block entry, the slots of the
pairs that carry the entry's dofs.  Two assemblies of one operator therefore
give bit-identical blocks (with atomic adds the last bits followed the order
in which pairs finished), and so does an assembly whose element matrices go
through many small batches (NGBEM_NEAR_EM_GB)."""

import numpy as np
import pytest

import ngbem_amd as nb

pytestmark = pytest.mark.gpu


def _near_blocks(op):
    out = []
    for b in range(op.num_blocks()):
        if op.block_info(b)["rank"] == -1:
            out.append(np.asarray(op.block_data(b)).copy())
    return out


CASES = [
    ("LHS H1 1", lambda m: nb.HypersingularOperator(nb.H1(m, order=1), intorder=4, leafsize=24, eps=1e-8)),
    ("HHS H1 2", lambda m: nb.HelmholtzHypersingularOperator(nb.H1(m, order=2), nb.H1(m, order=2), kappa=3.0,
                                                             intorder=4, leafsize=24, eps=1e-8)),
    ("LDL H1 1 to L2 0", lambda m: nb.DoubleLayerPotentialOperator(nb.H1(m, order=1), nb.SurfaceL2(m, order=0),
                                                                   intorder=4, leafsize=24, eps=1e-8)),
    ("MSL HDiv 1", lambda m: nb.MaxwellSingleLayerPotentialOperator(nb.HDivSurface(m, order=1), 2.0, intorder=4,
                                                                     leafsize=24, eps=1e-8)),
    ("LHS H1 4 wide", lambda m: nb.HypersingularOperator(nb.H1(m, order=4), intorder=6, leafsize=24, eps=1e-8)),
]


def _same(a, b, name):
    assert len(a) == len(b) and len(a) > 0
    nz = 0
    for x, y in zip(a, b):
        assert x.shape == y.shape
        assert np.array_equal(x, y), f"{name}: a near block differs"
        nz += int(np.count_nonzero(x))
    assert nz > 0
    return nz


@pytest.mark.parametrize("name,make", CASES, ids=[c[0].replace(" ", "_") for c in CASES])
def test_near_field_is_bit_identical_on_repeat(gpu, name, make):
    mesh = nb.icosphere(1 if "wide" in name else 3)
    a = _near_blocks(make(mesh))
    b = _near_blocks(make(mesh))
    nz = _same(a, b, name)
    print(f"{name}: {len(a)} near blocks, {nz} nonzero entries, bit-identical", flush=True)


@pytest.mark.parametrize("name,make", CASES[:4], ids=[c[0].replace(" ", "_") for c in CASES[:4]])
def test_near_field_in_small_batches_is_identical(gpu, name, make, monkeypatch):
    mesh = nb.icosphere(3)
    a = _near_blocks(make(mesh))
    monkeypatch.setenv("NGBEM_NEAR_EM_GB", "0.002")  # 2 MiB of element matrices per batch
    b = _near_blocks(make(mesh))
    _same(a, b, name)


def test_atomic_fallback_for_blocks_beyond_the_budget(gpu, monkeypatch):
    """A block whose element matrices alone exceed the budget sends the
    operator to the per-pair atomic adds; the entries agree with the
    element-matrix form to rounding."""
    mesh = nb.icosphere(3)
    make = CASES[0][1]
    a = _near_blocks(make(mesh))
    monkeypatch.setenv("NGBEM_NEAR_EM_GB", "0.0001")  # 100 KiB: below one block's element matrices
    b = _near_blocks(make(mesh))
    assert len(a) == len(b)
    worst = max(float(np.abs(x - y).max() / np.abs(x).max()) for x, y in zip(a, b))
    assert worst < 1e-13


@pytest.mark.parametrize("name,make", CASES[:4], ids=[c[0].replace(" ", "_") for c in CASES[:4]])
def test_once_per_pair_equals_once_per_block(gpu, name, make, monkeypatch):
    """Each distinct element pair of a batch of blocks computed once (its first
    block's slot, the default) against once per block holding it
    (NGBEM_NEAR_DEDUP=0): the same element matrices summed in the same order,
    so bit-identical blocks."""
    mesh = nb.icosphere(3)
    a = _near_blocks(make(mesh))
    monkeypatch.setenv("NGBEM_NEAR_DEDUP", "0")
    b = _near_blocks(make(mesh))
    _same(a, b, name)
