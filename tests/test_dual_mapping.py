"""SurfaceL2 with NGSolve's dual_mapping (CPU): every shape divided by its
triangle's Jacobian 2|e|, so the space's dof factors are 1 / J and its mass
matrix is D^-1 M D^-1; only SurfaceL2 spaces accept the flag."""
import ctypes

import numpy as np
import pytest

import ngbem_amd as nb
from ngbem_amd import _capi
from ngbem_amd.spaces import mass_matrix_coo


def _dense(sp_a, sp_b):
    r, c, v = mass_matrix_coo(sp_a, sp_b)
    D = np.zeros((sp_b.ndof, sp_a.ndof))
    np.add.at(D, (r, c), v)
    return D


@pytest.mark.parametrize("order", [0, 2])
def test_dual_mapping_scales_and_mass(order):
    m = nb.icosphere(2)
    sp, spd = nb.SurfaceL2(m, order=order), nb.SurfaceL2(m, order=order, dual_mapping=True)
    J = 2 * m.areas()
    assert sp.dof_scale is None
    assert np.allclose(spd.dof_scale[spd.elem_dofs], (1 / J)[:, None], rtol=1e-15, atol=0)
    M, Md = _dense(sp, sp), _dense(spd, spd)
    s = spd.dof_scale
    assert np.abs(Md - s[:, None] * M * s[None, :]).max() <= 1e-15 * np.abs(Md).max()
    if order == 0:  # int (1/J)^2 over e = |e| / (2|e|)^2
        assert np.allclose(np.diag(Md), 1 / (4 * m.areas()), rtol=1e-13)
    # mixed pairs scale one side only
    h1 = nb.H1(m)
    assert np.allclose(_dense(h1, spd), s[:, None] * _dense(h1, sp), rtol=0, atol=1e-15)


def test_dual_mapping_only_for_surface_l2():
    m = nb.icosphere(1)
    with pytest.raises((TypeError, NotImplementedError)):
        nb.H1(m, dual_mapping=True)
    L = _capi.lib()
    h1 = nb.H1(m)  # keep the space alive across the call
    rc = L.bem_space_set_dual_mapping(h1.handle, 1)
    assert rc == -1 and b"SurfaceL2" in L.bem_last_error()
