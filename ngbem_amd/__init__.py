"""ngbem_amd: MI355X-native boundary element operator assembly and H-matrix apply.

Drop-in for the operator hot path of Weggler/ngbem: the same factory names and
keyword arguments (src/python_bem.cpp:27-173), an H-matrix ``.mat`` with the
BaseMatrix calls NGSolve's solvers use, and Krylov solvers with
ngsolve.krylovspace's argument names.  Assembly (Sauter-Schwab near field, ACA
far field) and apply run in hand-written HIP kernels for gfx950 behind the
C-ABI in include/ngbem_amd.h.
"""
from .mesh import Mesh, Region, icosphere, ellipsoid, box, read_mesh, read_stl, read_gmsh, read_netgen_vol, orient_outward
from .spaces import SurfaceL2, H1, HDivSurface, HCurl
from .operators import (BaseMatrix, BlockMatrix, BlockVector, HMatrix, IntegralOperator, IntegralOperatorC, MassMatrix,
                        SingleLayerPotentialOperator, DoubleLayerPotentialOperator, HypersingularOperator,
                        HelmholtzSingleLayerPotentialOperator, HelmholtzDoubleLayerPotentialOperator,
                        HelmholtzCombinedFieldOperator, HelmholtzHypersingularOperator,
                        MaxwellSingleLayerPotentialOperator, MaxwellSingleLayerPotentialOperatorCurl,
                        MaxwellDoubleLayerPotentialOperator, HelmholtzSingleLayerPotentialOperatorDKappa,
                        MultiGeometryOperator, union_mesh)
from .krylov import CG, GMRes, GMResBatch, RefinedSolve
from .batch import OperatorBatch, SingleLayerBatch, assemble_many
from .compression import TestCompression


def reserve(gib: float) -> None:
    """Map ``gib`` GiB of device memory into the library's memory pool
    (NGBEM_ALLOC=pool) so later assemblies reuse it (bem_reserve)."""
    from . import _capi

    _capi.check(_capi.lib().bem_reserve(float(gib) * 2 ** 30), "bem_reserve")

__all__ = ["Mesh", "Region", "icosphere", "ellipsoid", "box", "read_mesh", "read_stl", "read_gmsh", "read_netgen_vol", "orient_outward", "SurfaceL2", "H1", "HDivSurface", "HCurl", "BaseMatrix", "HMatrix", "IntegralOperator",
           "IntegralOperatorC", "MassMatrix", "BlockMatrix", "BlockVector", "SingleLayerPotentialOperator", "DoubleLayerPotentialOperator",
           "HypersingularOperator", "HelmholtzSingleLayerPotentialOperator", "HelmholtzDoubleLayerPotentialOperator",
           "HelmholtzCombinedFieldOperator", "HelmholtzHypersingularOperator", "MaxwellSingleLayerPotentialOperator",
           "MaxwellSingleLayerPotentialOperatorCurl", "MaxwellDoubleLayerPotentialOperator",
           "HelmholtzSingleLayerPotentialOperatorDKappa", "CG", "GMRes", "GMResBatch", "RefinedSolve",
           "OperatorBatch", "SingleLayerBatch", "assemble_many", "reserve", "MultiGeometryOperator", "union_mesh",
           "TestCompression"]
