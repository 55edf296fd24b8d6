"""Python operator API, mirroring the reference's factories (src/python_bem.cpp:27-173)
and the BaseMatrix interface NGSolve's solvers drive (src/hmat.hpp:139-197).

Every factory assembles eagerly on the GPU, like the reference's constructor
(src/ngbem.cpp:314-330).  ``op.mat`` is an :class:`HMatrix`: Mult / MultAdd /
MultTrans / MultTransAdd, ``*`` with a vector, ``.T`` (plain transpose, as
src/hmat.cpp:547-548), ``.H`` (adjoint), ``.nze``, ``CreateRowVector`` /
``CreateColVector``, and algebra with scalars and other matrices
(``0.5 * M + C.mat``).  Vectors are torch tensors on the GPU (numpy arrays are
accepted and moved there).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _capi
from .mesh import Mesh, Region
from .spaces import FESpace, H1, HCurl, HDivSurface, SurfaceL2, mass_matrix_coo

__all__ = ["BaseMatrix", "HMatrix", "IntegralOperator", "IntegralOperatorC", "MassMatrix", "BlockMatrix", "BlockVector",
           "SingleLayerPotentialOperator", "DoubleLayerPotentialOperator", "HypersingularOperator",
           "HelmholtzSingleLayerPotentialOperator", "HelmholtzDoubleLayerPotentialOperator",
           "HelmholtzCombinedFieldOperator", "HelmholtzHypersingularOperator",
           "MaxwellSingleLayerPotentialOperator", "MaxwellSingleLayerPotentialOperatorCurl",
           "MaxwellDoubleLayerPotentialOperator", "HelmholtzSingleLayerPotentialOperatorDKappa", "MultiGeometryOperator", "union_mesh", "device"]


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("ngbem_amd needs a GPU (HIP device); none is visible")
    return torch.device("cuda", torch.cuda.current_device())


_SCALARS = {}


def _scalar(v):
    """(re, im) of a scalar as a float64 array; the common ones are cached."""
    v = complex(v)
    a = _SCALARS.get(v)
    if a is None:
        a = np.array([v.real, v.imag], dtype=np.float64)
        if len(_SCALARS) < 64:
            _SCALARS[v] = a
    return a


def _as_vec(x, dtype, dev):
    if isinstance(x, np.ndarray):
        x = torch.from_numpy(np.ascontiguousarray(x))
    if not isinstance(x, torch.Tensor):
        raise TypeError("vectors are torch tensors or numpy arrays")
    return _materialized(x.to(device=dev, dtype=dtype))


def _materialized(t):
    """t with torch's lazy conjugate / negative views resolved: the library
    reads and writes the raw memory, which such a view does not hold."""
    if t.is_conj():
        t = t.resolve_conj()
    if t.is_neg():
        t = t.resolve_neg()
    return t


def _raw_view(t):
    """True when t's memory holds its values as they read (no lazy bits)."""
    return t.is_contiguous() and not t.is_conj() and not t.is_neg()


# ------------------------------------------------------------- BaseMatrix --

class BaseMatrix:
    """Linear operator with NGSolve's BaseMatrix calling conventions."""

    height: int
    width: int
    is_complex: bool

    def dtype(self):
        return torch.complex128 if self.is_complex else torch.float64

    # y += s * A x
    def MultAdd(self, s, x, y):
        raise NotImplementedError

    # y += s * A^T x
    def MultTransAdd(self, s, x, y):
        raise NotImplementedError

    # y += s * A^H x = s * conj(A^T conj(x))
    def MultAdjAdd(self, s, x, y):
        tmp = torch.zeros_like(y)
        self.MultTransAdd(1.0, x.conj() if torch.is_complex(x) else x, tmp)
        y.add_(s * tmp.conj() if torch.is_complex(tmp) else s * tmp)

    def Mult(self, x, y):
        y.zero_()
        self.MultAdd(1.0, x, y)

    # y = s * A x, whatever y held (an extension used by the Krylov solvers:
    # an H-matrix writes it in its own scatter, without a separate clearing of
    # y, and a sum of terms starts with the term that can do that)
    _native_scale = False

    def MultScale(self, s, x, y):
        y.zero_()
        self.MultAdd(s, x, y)

    def MultTrans(self, x, y):
        y.zero_()
        self.MultTransAdd(1.0, x, y)

    def CreateRowVector(self):
        return torch.zeros(self.width, dtype=self.dtype(), device=device())

    def CreateColVector(self):
        return torch.zeros(self.height, dtype=self.dtype(), device=device())

    def _out_dtype(self, x):
        if self.is_complex or torch.is_complex(x):
            return torch.complex128
        return torch.float64

    def __mul__(self, other):
        if isinstance(other, (int, float, complex)):
            return ScaledMatrix(other, self)
        x = _as_vec(other, self._out_dtype(other if isinstance(other, torch.Tensor) else torch.from_numpy(other)),
                    device())
        # a (k, width) batch of row vectors gives a (k, height) batch
        y = torch.zeros(x.shape[:-1] + (self.height,), dtype=x.dtype, device=x.device)
        self.MultAdd(1.0, x, y)
        return y

    def __rmul__(self, s):
        if isinstance(s, (int, float, complex)):
            return ScaledMatrix(s, self)
        return NotImplemented

    def __add__(self, other):
        return SumMatrix(self, other, 1.0)

    def __sub__(self, other):
        return SumMatrix(self, other, -1.0)

    def __neg__(self):
        return ScaledMatrix(-1.0, self)

    @property
    def T(self):
        return TransposeMatrix(self)

    @property
    def H(self):
        return AdjointMatrix(self)

    def Height(self):
        return self.height

    def Width(self):
        return self.width


class ScaledMatrix(BaseMatrix):
    def __init__(self, s, A):
        self.s, self.A = s, A
        self.height, self.width = A.height, A.width
        self.is_complex = A.is_complex or isinstance(s, complex)

    def MultAdd(self, s, x, y):
        self.A.MultAdd(s * self.s, x, y)

    def MultTransAdd(self, s, x, y):
        self.A.MultTransAdd(s * self.s, x, y)

    @property
    def _native_scale(self):
        return self.A._native_scale

    def MultScale(self, s, x, y):
        self.A.MultScale(s * self.s, x, y)


class SumMatrix(BaseMatrix):
    def __init__(self, A, B, sign):
        if A.height != B.height or A.width != B.width:
            raise ValueError("matrix sum: shapes differ")
        self.A, self.B, self.sign = A, B, sign
        self.height, self.width = A.height, A.width
        self.is_complex = A.is_complex or B.is_complex

    def MultAdd(self, s, x, y):
        self.A.MultAdd(s, x, y)
        self.B.MultAdd(s * self.sign, x, y)

    @property
    def _native_scale(self):
        return self.A._native_scale or self.B._native_scale

    def MultScale(self, s, x, y):
        if self.B._native_scale and not self.A._native_scale:
            self.B.MultScale(s * self.sign, x, y)
            self.A.MultAdd(s, x, y)
        else:
            self.A.MultScale(s, x, y)
            self.B.MultAdd(s * self.sign, x, y)

    def MultTransAdd(self, s, x, y):
        self.A.MultTransAdd(s, x, y)
        self.B.MultTransAdd(s * self.sign, x, y)


class TransposeMatrix(BaseMatrix):
    def __init__(self, A):
        self.A = A
        self.height, self.width = A.width, A.height
        self.is_complex = A.is_complex

    def MultAdd(self, s, x, y):
        self.A.MultTransAdd(s, x, y)

    def MultTransAdd(self, s, x, y):
        self.A.MultAdd(s, x, y)


class AdjointMatrix(BaseMatrix):
    def __init__(self, A):
        self.A = A
        self.height, self.width = A.width, A.height
        self.is_complex = A.is_complex

    def MultAdd(self, s, x, y):
        if isinstance(self.A, HMatrix):
            self.A._apply(_capi.APPLY_TRANS | _capi.APPLY_CONJ, s, x, y)
        else:
            self.A.MultAdjAdd(s, x, y)

    def MultTransAdd(self, s, x, y):
        # (A^H)^T = conj(A)
        tmp = torch.zeros_like(y)
        self.A.MultAdd(1.0, x.conj() if torch.is_complex(x) else x, tmp)
        y.add_((s * tmp.conj()) if torch.is_complex(tmp) else s * tmp)


class BlockMatrix(BaseMatrix):
    """NGSolve's ``BlockMatrix([[A, B], [C, D]])`` (entries may be None) on
    vectors that are the concatenation of the blocks' vectors
    (:func:`BlockVector`); each block's height / width is read from any
    entry of its block row / column."""

    def __init__(self, blocks):
        rows, cols = len(blocks), len(blocks[0])
        if any(len(r) != cols for r in blocks):
            raise ValueError("BlockMatrix: rows of different lengths")
        hs, ws = [None] * rows, [None] * cols
        for i, r in enumerate(blocks):
            for j, b in enumerate(r):
                if b is None:
                    continue
                if (hs[i] not in (None, b.height)) or (ws[j] not in (None, b.width)):
                    raise ValueError("BlockMatrix: block sizes do not line up")
                hs[i], ws[j] = b.height, b.width
        if None in hs or None in ws:
            raise ValueError("BlockMatrix: every block row and column needs one entry")
        self.blocks = blocks
        self.roff = np.concatenate([[0], np.cumsum(hs)]).astype(int)
        self.coff = np.concatenate([[0], np.cumsum(ws)]).astype(int)
        self.height, self.width = int(self.roff[-1]), int(self.coff[-1])
        self.is_complex = any(b.is_complex for r in blocks for b in r if b is not None)

    def MultAdd(self, s, x, y):
        for i, r in enumerate(self.blocks):
            for j, b in enumerate(r):
                if b is not None:
                    b.MultAdd(s, x[..., self.coff[j]:self.coff[j + 1]], y[..., self.roff[i]:self.roff[i + 1]])

    def MultTransAdd(self, s, x, y):
        for i, r in enumerate(self.blocks):
            for j, b in enumerate(r):
                if b is not None:
                    b.MultTransAdd(s, x[..., self.roff[i]:self.roff[i + 1]], y[..., self.coff[j]:self.coff[j + 1]])


def BlockVector(parts):
    """NGSolve's ``BlockVector([a, b])``: the concatenated vector a BlockMatrix
    acts on (slice it by the block sizes to get the parts back), on the device
    of the first torch part (the GPU when all parts are numpy arrays)."""
    dev = next((p.device for p in parts if isinstance(p, torch.Tensor)), None) or device()
    parts = [p if isinstance(p, torch.Tensor) else torch.from_numpy(np.asarray(p)) for p in parts]
    dt = torch.complex128 if any(torch.is_complex(p) for p in parts) else torch.float64
    return torch.cat([p.to(device=dev, dtype=dt) for p in parts], dim=-1)


class MassMatrix(BaseMatrix):
    """Surface mass matrix  int u . v ds  (test x trial) of any two spaces of
    the same family (scalar or edge), the ``Id`` of the reference's
    formulations (e.g. demos/HelmholtzBurtonMiller.ipynb)."""

    def __init__(self, trial: FESpace, test: FESpace | None = None):
        test = test or trial
        r, c, v = mass_matrix_coo(trial, test)
        dev = device()
        self.trial, self.test = trial, test
        self.height, self.width = test.ndof, trial.ndof
        self.is_complex = False
        self.rows = torch.from_numpy(r).to(dev)
        self.cols = torch.from_numpy(c).to(dev)
        self.vals = torch.from_numpy(v).to(dev)
        # SurfaceL2 order 0 on one mesh: one entry per dof, on the diagonal
        self._diag = None
        if self.height == self.width and len(r) == self.height and np.array_equal(r, c) and np.array_equal(
                np.sort(r), np.arange(self.height)):
            d = np.zeros(self.height)
            d[r] = v
            self._diag = {torch.float64: torch.from_numpy(d).to(dev)}

    def _diag_as(self, dtype):
        d = self._diag.get(dtype)
        if d is None:
            d = self._diag[dtype] = self._diag[torch.float64].to(dtype)
        return d

    def MultAdd(self, s, x, y):
        x = x.to(y.dtype)
        if self._diag is not None:
            y.addcmul_(x, self._diag_as(y.dtype), value=s)
            return
        y.index_add_(-1, self.rows, (s * self.vals).to(y.dtype) * x[..., self.cols])

    def MultTransAdd(self, s, x, y):
        x = x.to(y.dtype)
        if self._diag is not None:
            y.addcmul_(x, self._diag_as(y.dtype), value=s)
            return
        y.index_add_(-1, self.cols, (s * self.vals).to(y.dtype) * x[..., self.rows])

    def Inverse(self, freedofs=None, tol=1e-14, maxsteps=2000):
        """The inverse of a square mass matrix (NGSolve's ``mat.Inverse(freedofs)``,
        as the reference's demos use it for preconditioners): on the free dofs
        (all when ``freedofs`` is None; else a boolean mask) the solution of
        M_ff z_f = x_f, zero elsewhere.  SurfaceL2 spaces (element-local dofs,
        block-diagonal M) are inverted exactly block by block; the others by
        Jacobi-preconditioned CG to the relative residual ``tol``."""
        if self.trial is not self.test:
            raise ValueError("Inverse needs a square mass matrix (one space)")
        return _MassInverse(self, freedofs, tol, maxsteps)


class _MassInverse(BaseMatrix):
    def __init__(self, M: MassMatrix, freedofs, tol, maxsteps):
        dev = M.vals.device
        n = M.height
        self.height = self.width = n
        self.is_complex = False
        self._M, self._tol, self._maxsteps = M, tol, maxsteps
        free = np.ones(n, dtype=bool) if freedofs is None else np.asarray(
            freedofs.cpu() if isinstance(freedofs, torch.Tensor) else freedofs, dtype=bool).reshape(-1)
        if free.shape != (n,):
            raise ValueError(f"freedofs must have {n} entries")
        self._free = torch.from_numpy(free.astype(np.float64)).to(dev)
        self._blocks = None
        self._diagonal = None  # float64 vector when the inverse is diagonal
        self._diag_cache = {}
        sp = M.trial
        if isinstance(sp, SurfaceL2):
            # element-local dofs: M is block-diagonal with one nloc x nloc block
            # per element; restricted to the free dofs (identity on the others,
            # masked out again after the solve)
            ed = torch.from_numpy(sp.elem_dofs).to(dev)
            on = (ed[:, 0] >= 0)
            ed = ed[on]
            ne, nl = ed.shape
            pos = torch.full((n,), -1, dtype=torch.int64, device=dev)
            pos[ed.reshape(-1)] = torch.arange(ne * nl, device=dev)
            B = torch.zeros(ne * nl * nl, dtype=torch.float64, device=dev)
            pr, pc = pos[M.rows], pos[M.cols]
            B.index_add_(0, (pr // nl) * nl * nl + (pr % nl) * nl + (pc % nl), M.vals)
            B = B.view(ne, nl, nl)
            f = self._free[ed]                                   # ne x nl
            B = B * f[:, :, None] * f[:, None, :] + torch.diag_embed(1.0 - f)
            self._blocks = (ed, torch.linalg.inv(B))
            if nl == 1:
                # order 0: one dof per element, the inverse is the diagonal 1 / |T| on the free dofs
                d = torch.zeros(n, dtype=torch.float64, device=dev)
                d[ed[:, 0]] = self._blocks[1][:, 0, 0]
                self._diagonal = (d * self._free).contiguous()
        else:
            d = torch.zeros(n, dtype=torch.float64, device=dev)
            d.index_add_(0, M.rows[M.rows == M.cols], M.vals[M.rows == M.cols])
            self._dinv = torch.where(self._free > 0, 1.0 / d, torch.zeros_like(d))

    def _solve(self, x):
        if self._diagonal is not None:
            rdt = x.real.dtype if torch.is_complex(x) else x.dtype
            d = self._diag_cache.get(rdt)
            if d is None:
                d = self._diag_cache[rdt] = self._diagonal.to(rdt)
            return x * d
        f = self._free.to(x.dtype)
        b = x * f
        if self._blocks is not None:
            ed, Binv = self._blocks
            z = torch.zeros_like(b)
            loc = b[..., ed]                                      # (..., ne, nl)
            z[..., ed] = torch.einsum("eij,...ej->...ei", Binv.to(b.dtype), loc)
            return z * f
        M, dinv = self._M, self._dinv.to(x.dtype)
        Mf = lambda v: M * (v * f) * f  # noqa: E731  (M_ff, zero on the fixed dofs)
        z = torch.zeros_like(b)
        r = b.clone()
        p = r * dinv
        rz = (r.conj() * p).sum(-1, keepdim=True)
        nb = torch.linalg.norm(b, dim=-1, keepdim=True)
        one = torch.ones_like(rz)
        for _ in range(self._maxsteps):
            # every right-hand side stops on its own: a converged row (a zero
            # row from the start included) keeps z and r from then on, so its
            # zero rz never enters a division
            done = torch.linalg.norm(r, dim=-1, keepdim=True) <= self._tol * nb
            if bool(done.all()):
                break
            q = Mf(p)
            pq = (p.conj() * q).sum(-1, keepdim=True)
            alpha = torch.where(done, torch.zeros_like(rz), rz / torch.where(done, one, pq))
            z = z + alpha * p
            r = r - alpha * q
            w = r * dinv
            rz1 = (r.conj() * w).sum(-1, keepdim=True)
            p = w + torch.where(done, torch.zeros_like(rz), rz1 / torch.where(done, one, rz)) * p
            rz = torch.where(done, rz, rz1)
        return z * f

    def MultAdd(self, s, x, y):
        y += s * self._solve(x.to(y.dtype))

    def MultTransAdd(self, s, x, y):  # symmetric
        y += s * self._solve(x.to(y.dtype))


# ---------------------------------------------------------------- HMatrix --

class HMatrix(BaseMatrix):
    """The assembled H-matrix of one operator (src/hmat.hpp:163-201)."""

    def __init__(self, op, mixed=False, graph=False):
        self._op = op
        self.height, self.width = op.test_space.ndof, op.trial_space.ndof
        self.is_complex = op.is_complex
        self._mixed = mixed
        self._graph = graph

    @property
    def nze(self):
        return int(self._op.stats()["nze"])

    @property
    def mixed(self):
        """View of this matrix applied with the FP32 / complex64 copy."""
        if not self._op.has_mixed:
            raise RuntimeError("operator built without mixed=True")
        return HMatrix(self._op, mixed=True, graph=self._graph)

    @property
    def graph(self):
        """View of this matrix whose single-vector applies replay the apply
        captured once in a hipGraph (BEM_APPLY_GRAPH); no counterpart in the
        reference."""
        return HMatrix(self._op, mixed=self._mixed, graph=True)

    def IsComplex(self):
        return self.is_complex

    _native_scale = True

    def MultScale(self, s, x, y):
        """y = s A x in one apply (beta = 0 in the scatter)."""
        s = complex(s)
        if (isinstance(x, torch.Tensor) and x.is_cuda and y.is_cuda and x.dim() == y.dim() and x.dim() <= 2
                and x.shape[-1] == self.width and y.shape[-1] == self.height
                and (x.dim() == 1 or x.shape[0] == y.shape[0]) and _raw_view(x) and _raw_view(y)
                and x.is_contiguous() and y.is_contiguous()):
            flags = (_capi.APPLY_MIXED if self._mixed else 0) | (
                _capi.APPLY_GRAPH if self._graph and x.dim() == 1 else 0)
            if self.is_complex and x.dtype == y.dtype == torch.complex128:
                self._op._call_apply(flags, s, x, 0.0, y)
                return
            if not self.is_complex and x.dtype == y.dtype == torch.float64 and s.imag == 0.0:
                self._op._call_apply(flags, s.real, x, 0.0, y)
                return
        y.zero_()
        self.MultAdd(s, x, y)

    def _apply(self, flags, s, x, y):
        op = self._op
        trans = flags & _capi.APPLY_TRANS
        nin, nout = (self.height, self.width) if trans else (self.width, self.height)
        if x.shape[-1] != nin or y.shape[-1] != nout or x.dim() != y.dim() or x.dim() > 2 or (
                x.dim() == 2 and x.shape[0] != y.shape[0]):
            raise ValueError(f"vector sizes {tuple(x.shape)}, {tuple(y.shape)} do not match {nout}x{nin}")
        shape_out = y.shape
        if not y.is_cuda:
            raise ValueError("output vector must live on the GPU")
        if self._mixed:
            flags |= _capi.APPLY_MIXED
        if self._graph and x.dim() == 1:
            flags |= _capi.APPLY_GRAPH
        s = complex(s)
        if not self.is_complex:
            # real operator: real and imaginary parts separately
            if torch.is_complex(y):
                xr = _as_vec(x.real if torch.is_complex(x) else x, torch.float64, y.device)
                xi = _as_vec(x.imag, torch.float64, y.device) if torch.is_complex(x) else None
                yr = torch.zeros(shape_out, dtype=torch.float64, device=y.device)
                op._call_apply(flags, 1.0, xr.contiguous(), 0.0, yr)
                res = yr.to(torch.complex128)
                if xi is not None:
                    xi = xi.contiguous()
                    yi = torch.zeros(shape_out, dtype=torch.float64, device=y.device)
                    op._call_apply(flags, 1.0, xi, 0.0, yi)
                    res = res + 1j * yi
                y.add_(s * res)
                return
            if s.imag != 0.0:
                raise TypeError("complex scalar with a real vector")
            xr = _as_vec(x, torch.float64, y.device).contiguous()
            if _raw_view(y):
                op._call_apply(flags, s.real, xr, 1.0, y)
            else:
                tmp = _materialized(y).contiguous().clone()
                op._call_apply(flags, s.real, xr, 1.0, tmp)
                y.copy_(tmp)
            return
        if not torch.is_complex(y):
            raise TypeError("complex operator needs a complex output vector")
        xc = _as_vec(x, torch.complex128, y.device).contiguous()
        if _raw_view(y):
            op._call_apply(flags, s, xc, 1.0, y)
        else:
            tmp = _materialized(y).contiguous().clone()
            op._call_apply(flags, s, xc, 1.0, tmp)
            y.copy_(tmp)

    def MultAdd(self, s, x, y):
        self._apply(0, s, x, y)

    def Mult(self, x, y):
        # y = A x in one call (beta = 0): no separate clearing of y
        if (self.is_complex and torch.is_complex(x) and x.dtype == y.dtype == torch.complex128 and x.is_cuda
                and y.is_cuda and x.dim() == 1 and x.shape[0] == self.width and y.shape[0] == self.height
                and _raw_view(x) and _raw_view(y)):
            self._op._call_apply((_capi.APPLY_MIXED if self._mixed else 0) | (_capi.APPLY_GRAPH if self._graph else 0),
                                 1.0, x, 0.0, y)
            return
        BaseMatrix.Mult(self, x, y)

    def MultTransAdd(self, s, x, y):
        self._apply(_capi.APPLY_TRANS, s, x, y)

    def MultAdjAdd(self, s, x, y):
        self._apply(_capi.APPLY_TRANS | _capi.APPLY_CONJ, s, x, y)

    def blocks(self):
        return self._op.blocks()


# ------------------------------------------------- several geometries --

def union_mesh(meshes, names=None):
    """One mesh holding the given surfaces one after another: (mesh, element
    offsets, vertex offsets).  Surface i becomes region i, named names[i]
    when given (so that mesh.Boundaries(names[i]) selects it)."""
    xyz, tri, eoff, voff = [], [], [0], [0]
    for m in meshes:
        tri.append(m.tri.astype(np.int64) + voff[-1])
        xyz.append(m.xyz)
        eoff.append(eoff[-1] + m.nt)
        voff.append(voff[-1] + m.nv)
    regions = np.repeat(np.arange(len(meshes)), [m.nt for m in meshes])
    rn = {} if names is None else {i: str(n) for i, n in enumerate(names)}
    return Mesh(np.concatenate(xyz), np.concatenate(tri), regions, rn), np.array(eoff), np.array(voff)


def MultiGeometryOperator(factory_kernel, meshes, kappas=None, trial_space=SurfaceL2, test_space=None, intorder=3,
                          leafsize=40, eta=2., eps=1e-6, mixed=False):
    """One operator over several geometries (block diagonal): every geometry
    gets the blocks its own operator would have, with its own wavenumber, and
    assembly and apply of all of them run in the launches of one operator
    (bem_op_create_multi).  ``factory_kernel``: a kernel id of ``_capi``
    (LAPLACE_SL, HELMHOLTZ_SL, HELMHOLTZ_SL_DKAPPA, ...); the spaces are
    classes (SurfaceL2, H1, ...) instantiated on the union mesh.  Vectors are
    the concatenation of the geometries' vectors.  Returns the operator; its
    ``dof_offsets`` split vectors per geometry.  No counterpart in the
    reference."""
    mesh, eoff, _ = union_mesh(meshes)
    n = len(meshes)
    kap = np.zeros(n) if kappas is None else np.asarray(kappas, dtype=np.float64)
    if kap.shape != (n,):
        raise ValueError("one wavenumber per geometry")
    trial = trial_space(mesh)
    test = trial if test_space is None else test_space(mesh)
    cls = IntegralOperatorC if factory_kernel >= _capi.HELMHOLTZ_SL else IntegralOperator
    op = cls(factory_kernel, kap, trial, test, intorder, leafsize, eta, eps, False, mixed, False, _multi=(kap, eoff))
    op.dof_offsets = np.cumsum([0] + [trial_space(m).ndof for m in meshes])
    op.meshes = list(meshes)
    return op


# -------------------------------------------------------------- potential --

class PotentialCF:
    """u_c(x) = sum over the trial side's quadrature points y of
    J w K(x, y, n_x, n_y) vals(y) per kernel term (PotentialCF::T_Evaluate,
    src/ngbem.cpp:1187-1235; the term factors are not applied, as there).
    Call with points (npts x 3) and optional normals n_x (zero otherwise);
    returns (npts,) for scalar evaluators, else (npts, dim)."""

    def __init__(self, op, coef):
        self._op = op
        dt = torch.complex128 if op.is_complex else torch.float64
        self.coef = _as_vec(coef, dt, device()).contiguous()
        if self.coef.shape != (op.trial_space.ndof,):
            raise ValueError("coefficient vector does not match the trial space")
        self.dim = int(_capi.lib().bem_op_potential_dim(op._h))

    def __call__(self, points, normals=None):
        op = self._op
        dev = self.coef.device
        P = _as_vec(points, torch.float64, dev).reshape(-1, 3).contiguous()
        N = None if normals is None else _as_vec(normals, torch.float64, dev).reshape(-1, 3).contiguous()
        out = torch.zeros(P.shape[0], self.dim, dtype=self.coef.dtype, device=dev)
        reg = getattr(op, "trial_definedon", None)  # src/ngbem.cpp:1205: only the region's elements
        em = None if reg is None else np.ascontiguousarray(reg.mask, dtype=np.uint8)
        with torch.cuda.device(dev):
            stream = torch.cuda.current_stream().cuda_stream
            _capi.check(_capi.lib().bem_op_potential(
                op._h, ctypes.c_void_p(self.coef.data_ptr()), ctypes.c_void_p(P.data_ptr()),
                None if N is None else ctypes.c_void_p(N.data_ptr()), P.shape[0],
                None if em is None else em.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(out.data_ptr()),
                _capi.C128 if op.is_complex else _capi.F64, ctypes.c_void_p(stream)), "bem_op_potential")
        return out[:, 0] if self.dim == 1 else out


# ------------------------------------------------------ integral operator --

class IntegralOperator:
    """Assembled boundary integral operator (real-valued: Laplace)."""

    is_complex = False

    def __init__(self, kernel, kappa, trial_space, test_space, intorder, leafsize, eta, eps, dense=False,
                 mixed=False, testhmatrix=False, _multi=None, definedon=(None, None)):
        if not isinstance(trial_space, FESpace) or not isinstance(test_space, FESpace):
            raise TypeError("spaces must be ngbem_amd SurfaceL2 / H1 / HDivSurface / HCurl spaces")
        self.trial_space, self.test_space = trial_space, test_space
        # operator-level regions (trial_definedon / test_definedon): as in the
        # reference they select the elements of the potential and the dofs of
        # testhmatrix's dense matrix, not the H-matrix (src/ngbem.cpp:37-113,
        # 1170, 1205)
        self.trial_definedon, self.test_definedon = (_region_of(trial_space, definedon[0]),
                                                     _region_of(test_space, definedon[1]))
        self.kernel, self.kappa = kernel, float(np.atleast_1d(kappa)[0])
        self.params = dict(intorder=intorder, leafsize=leafsize, eta=eta, eps=eps, dense=dense)
        self.has_mixed = bool(mixed)
        L = _capi.lib()
        p = _capi.Params(int(intorder), int(leafsize), float(eta), float(eps), int(bool(dense)), int(bool(mixed)))
        out = ctypes.c_void_p()
        if _multi is None:
            _capi.check(L.bem_op_create(kernel, float(kappa), trial_space.handle, test_space.handle, ctypes.byref(p),
                                        ctypes.byref(out)), "bem_op_create")
        else:  # several geometries: per-geometry wavenumbers and element offsets
            kap, eoff = _multi
            kap = np.ascontiguousarray(kap, dtype=np.float64)
            eoff = np.ascontiguousarray(eoff, dtype=np.int32)
            _capi.check(L.bem_op_create_multi(kernel, _capi.dptr(kap), len(kap), _capi.iptr(eoff),
                                              trial_space.handle, test_space.handle, ctypes.byref(p),
                                              ctypes.byref(out)), "bem_op_create_multi")
        self._h = out
        dev = device()
        with torch.cuda.device(dev):
            stream = torch.cuda.current_stream().cuda_stream
            _capi.check(L.bem_op_assemble(self._h, ctypes.c_void_p(stream)), "bem_op_assemble")
        self._mat = HMatrix(self)
        if testhmatrix:
            self._test_hmatrix()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _capi.lib().bem_op_destroy(h)
            self._h = None

    @property
    def mat(self) -> HMatrix:
        return self._mat

    def GetMatrix(self):
        return self._mat

    def GetPotential(self, gf):
        """The potential of the trial-side density ``gf`` (its coefficient
        vector) as a callable ``pot(points, normals=None)`` evaluated on the GPU
        (GenericIntegralOperator::GetPotential, src/ngbem.cpp:1166-1235)."""
        return PotentialCF(self, gf)

    def Autotune(self, mixed=None, iters=10):
        """Time candidate work shapes of the forward apply on this operator and
        keep the fastest (bem_op_autotune): the FP64 storage, and the FP32 copy
        too when the operator has one (``mixed=None``) or only it
        (``mixed=True``).  Returns {"fp64": report, "fp32": report} with each
        candidate's median milliseconds.  No counterpart in the reference."""
        import json

        precs = [False, True] if mixed is None else [bool(mixed)]
        if mixed is None and not self.has_mixed:
            precs = [False]
        out = {}
        L = _capi.lib()
        buf = ctypes.create_string_buffer(1 << 14)
        for m in precs:
            stream = torch.cuda.current_stream().cuda_stream
            rc = L.bem_op_autotune(self._h, _capi.APPLY_MIXED if m else 0, int(iters), buf, len(buf), stream)
            if rc < 0:
                _capi.check(rc, "bem_op_autotune")
            out["fp32" if m else "fp64"] = json.loads(buf.value.decode())
        return out

    def stats(self):
        st = _capi.Stats()
        _capi.check(_capi.lib().bem_op_stats(self._h, ctypes.byref(st)), "bem_op_stats")
        return st.as_dict()

    def _call_apply(self, flags, alpha, x, beta, y):
        """x, y: device vectors, or batches (k, n) of k row vectors (the forward
        product of a batch runs as one block product on the matrix cores)."""
        a = _scalar(alpha)
        b = _scalar(beta)
        dtype = _capi.C128 if self.is_complex else _capi.F64
        nrhs, ld = 1, 0
        xb, yb = x, y
        if x.dim() == 2:
            nrhs, nin, nout = x.shape[0], x.shape[1], y.shape[1]
            ld = max(nin, nout)
            if nin != nout:  # one leading dimension for both sides
                xb = torch.zeros(nrhs, ld, dtype=x.dtype, device=x.device)
                xb[:, :nin] = x
                yb = torch.zeros(nrhs, ld, dtype=y.dtype, device=y.device)
                if b.any():
                    yb[:, :nout] = y
        L = _capi.lib()
        if y.device.index == torch.cuda.current_device():  # no device switch needed
            stream = torch.cuda.current_stream().cuda_stream
            rc = L.bem_op_apply(self._h, flags | _capi.APPLY_ON_DEVICE, a.ctypes.data, xb.data_ptr(), b.ctypes.data,
                                yb.data_ptr(), nrhs, ld, dtype, stream)
        else:
            with torch.cuda.device(y.device):
                stream = torch.cuda.current_stream().cuda_stream
                rc = L.bem_op_apply(self._h, flags | _capi.APPLY_ON_DEVICE, a.ctypes.data, xb.data_ptr(),
                                    b.ctypes.data, yb.data_ptr(), nrhs, ld, dtype, stream)
        _capi.check(rc, "bem_op_apply")
        if yb is not y:
            y.copy_(yb[:, : y.shape[1]])

    # -- introspection (parity tests) --
    def num_blocks(self):
        return _capi.lib().bem_op_num_blocks(self._h)

    def block_info(self, b):
        bi = _capi.BlockInfo()
        _capi.check(_capi.lib().bem_op_block_info(self._h, b, ctypes.byref(bi)))
        return {k: getattr(bi, k) for k, _ in bi._fields_}

    def blocks(self):
        return [self.block_info(b) for b in range(self.num_blocks())]

    def block_dofs(self, b):
        i = self.block_info(b)
        rows = np.zeros(i["m"], dtype=np.int32)
        cols = np.zeros(i["n"], dtype=np.int32)
        _capi.check(_capi.lib().bem_op_block_dofs(self._h, b, _capi.iptr(rows), _capi.iptr(cols)))
        return rows, cols

    def block_data(self, b):
        i = self.block_info(b)
        m, n, r = i["m"], i["n"], i["rank"]
        dt = np.complex128 if self.is_complex else np.float64
        size = m * n if r < 0 else r * (m + n)
        buf = np.zeros(max(size, 1), dtype=dt)
        _capi.check(_capi.lib().bem_op_block_data(self._h, b, ctypes.c_void_p(buf.ctypes.data)))
        if r < 0:
            return buf[: m * n].reshape(m, n)
        return buf[: m * r].reshape(m, r), buf[m * r: m * r + r * n].reshape(r, n)

    def tree(self, which=0):
        n = self.trial_space.ndof if which == 0 else self.test_space.ndof
        perm = np.zeros(n, dtype=np.int32)
        L = _capi.lib()
        ncl = L.bem_op_tree(self._h, which, _capi.iptr(perm), None, None, 0)
        cl = np.zeros(7 * ncl)
        cen = np.zeros(3 * ncl)
        L.bem_op_tree(self._h, which, _capi.iptr(perm), _capi.dptr(cl), _capi.dptr(cen), ncl)
        return perm, cl.reshape(ncl, 7), cen.reshape(ncl, 3)

    def densify(self):
        """Full matrix from the stored blocks (small problems only)."""
        dt = np.complex128 if self.is_complex else np.float64
        A = np.zeros((self.test_space.ndof, self.trial_space.ndof), dtype=dt)
        for b in range(self.num_blocks()):
            rows, cols = self.block_dofs(b)
            blk = self.block_data(b)
            if isinstance(blk, tuple):
                blk = blk[0] @ blk[1]
            A[np.ix_(rows, cols)] += blk
        return A

    def _test_hmatrix(self):
        """Spectral-norm error of H against the dense operator (30 power
        iterations on (D-H)^H (D-H)) and compression rate, as the reference's
        testhmatrix diagnostic (src/ngbem.cpp:332-369)."""
        p = dict(self.params)
        dense = type(self)(self.kernel, self.kappa, self.trial_space, self.test_space, p["intorder"], p["leafsize"],
                           p["eta"], p["eps"], dense=True)
        print(f"dense: {self.test_space.ndof} x {self.trial_space.ndof}")
        D, H = dense.mat, self.mat
        # the reference's dense matrix holds the rows and columns of the
        # definedon regions' dofs only (mapbnd2glob, src/ngbem.cpp:275)
        dt = self.mat.dtype()
        cmask = _dof_mask(self.trial_space, self.trial_definedon).to(dt)
        rmask = _dof_mask(self.test_space, self.test_definedon).to(dt)
        x = torch.ones(self.trial_space.ndof, dtype=dt, device=device())
        for _ in range(30):
            x = x / torch.linalg.norm(x)
            y = rmask * (D * (cmask * x))
            H.MultAdd(-1.0, x, y)
            y = y.conj()
            x2 = cmask * (D.T * (rmask * y))
            H.MultTransAdd(-1.0, y, x2)
            x = x2.conj()
        st = self.stats()
        rate = st["nze"] / (self.test_space.ndof * self.trial_space.ndof)
        print(f"compression rate {rate}  2-norm error {math.sqrt(float(torch.linalg.norm(x)))}")


class IntegralOperatorC(IntegralOperator):
    """Assembled boundary integral operator (complex-valued: Helmholtz)."""

    is_complex = True


def _dense_remap(method, leafsize, eta):
    if method == "dense":
        return True
    if method != "aca":
        raise ValueError(f"method must be 'aca' or 'dense', got {method!r}")
    return False


def _region_of(space, region):
    """A Region of the space's mesh (or a pattern for mesh.Boundaries), or None."""
    if region is None:
        return None
    if isinstance(region, str):
        region = space.mesh.Boundaries(region)
    if not isinstance(region, Region) or region.mesh is not space.mesh:
        raise ValueError("definedon must be a Region of the space's mesh (mesh.Boundaries(...))")
    return region


def _dof_mask(space, region):
    """1 for the dofs of the region's elements (all dofs without a region)."""
    m = torch.ones(space.ndof, dtype=torch.float64)
    if region is not None:
        d = space.elem_dofs[region.mask].reshape(-1)
        m.zero_()
        m[torch.from_numpy(np.unique(d[d >= 0]))] = 1.0
    return m.to(device())


def SingleLayerPotentialOperator(space, intorder=3, leafsize=40, eta=2., eps=1e-6, method="aca", testhmatrix=False,
                                 mixed=False):
    """Laplace single layer V (src/python_bem.cpp:27-38, src/ngbem.hpp:230-246)."""
    dense = _dense_remap(method, leafsize, eta)
    return IntegralOperator(_capi.LAPLACE_SL, 0.0, space, space, intorder, leafsize, eta, eps, dense, mixed,
                            testhmatrix)


def DoubleLayerPotentialOperator(trial_space, test_space, trial_definedon=None, test_definedon=None, intorder=3,
                                 leafsize=40, eta=2., eps=1e-6, method="aca", testhmatrix=False, mixed=False):
    """Laplace double layer K (src/python_bem.cpp:41-60, src/ngbem.hpp:257-274)."""
    dense = _dense_remap(method, leafsize, eta)
    return IntegralOperator(_capi.LAPLACE_DL, 0.0, trial_space, test_space, intorder, leafsize, eta, eps, dense,
                            mixed, testhmatrix, definedon=(trial_definedon, test_definedon))


def HypersingularOperator(space, definedon=None, intorder=3, leafsize=40, eta=2., eps=1e-6, method="aca",
                          testhmatrix=False, mixed=False):
    """Laplace hypersingular W with n x grad evaluators (src/python_bem.cpp:63-75)."""
    if method not in ("aca", "dense"):
        raise ValueError("method must be 'aca' or 'dense'")
    return IntegralOperator(_capi.LAPLACE_HS, 0.0, space, space, intorder, leafsize, eta, eps, False, mixed,
                            testhmatrix, definedon=(definedon, definedon))


def _helmholtz(kernel, trial_space, test_space, kappa, intorder, leafsize, eta, eps, method, testhmatrix, mixed,
               definedon=(None, None)):
    if kappa is None:
        raise TypeError("kappa is required")
    if method not in ("aca", "dense"):
        raise ValueError("method must be 'aca' or 'dense'")
    test_space = test_space if test_space is not None else trial_space
    return IntegralOperatorC(kernel, kappa, trial_space, test_space, intorder, leafsize, eta, eps, False, mixed,
                             testhmatrix, definedon=definedon)


def HelmholtzSingleLayerPotentialOperator(trial_space, test_space=None, kappa=None, intorder=3, leafsize=40, eta=2.,
                                          eps=1e-6, method="aca", testhmatrix=False, mixed=False):
    """Helmholtz single layer, e^{ikr}/(4 pi r) (src/python_bem.cpp:79-87)."""
    return _helmholtz(_capi.HELMHOLTZ_SL, trial_space, test_space, kappa, intorder, leafsize, eta, eps, method,
                      testhmatrix, mixed)


def HelmholtzSingleLayerPotentialOperatorDKappa(trial_space, test_space=None, kappa=None, intorder=3, leafsize=40,
                                                eta=2., eps=1e-6, method="aca", testhmatrix=False, mixed=False):
    """d/dkappa of the Helmholtz single layer: kernel i e^{ikr}/(4 pi), smooth.
    No counterpart in the reference; it gives the kappa-gradient of V(kappa) x
    (ngbem_amd.batch.SingleLayerBatch)."""
    return _helmholtz(_capi.HELMHOLTZ_SL_DKAPPA, trial_space, test_space, kappa, intorder, leafsize, eta, eps,
                      method, testhmatrix, mixed)


def HelmholtzDoubleLayerPotentialOperator(trial_space, test_space=None, kappa=None, intorder=3, leafsize=40, eta=2.,
                                          eps=1e-6, method="aca", testhmatrix=False, mixed=False):
    """Helmholtz double layer (src/python_bem.cpp:91-99)."""
    return _helmholtz(_capi.HELMHOLTZ_DL, trial_space, test_space, kappa, intorder, leafsize, eta, eps, method,
                      testhmatrix, mixed)


def HelmholtzCombinedFieldOperator(trial_space, test_space=None, kappa=None, intorder=3, leafsize=40, eta=2.,
                                   eps=1e-6, method="aca", testhmatrix=False, mixed=False):
    """Combined field K - i kappa V (src/python_bem.cpp:102-110, src/ngbem.hpp:412-434)."""
    return _helmholtz(_capi.HELMHOLTZ_CF, trial_space, test_space, kappa, intorder, leafsize, eta, eps, method,
                      testhmatrix, mixed)


def HelmholtzHypersingularOperator(trial_space, test_space=None, kappa=None, intorder=3, leafsize=40, eta=2.,
                                   eps=1e-6, method="aca", testhmatrix=False, mixed=False):
    """Helmholtz hypersingular with [n x grad, u] evaluators (src/python_bem.cpp:113-119)."""
    return _helmholtz(_capi.HELMHOLTZ_HS, trial_space, test_space, kappa, intorder, leafsize, eta, eps, method,
                      testhmatrix, mixed)


def MaxwellSingleLayerPotentialOperator(space, kappa, definedon=None, intorder=3, leafsize=40, eta=2., eps=1e-6,
                                        method="aca", testhmatrix=False, mixed=False):
    """Maxwell single layer on HDivSurface: kappa <G phi, phi> - 1/kappa <G div, div>
    (src/python_bem.cpp:122-134, src/ngbem.hpp:442-468, DiffOpMaxwellNew)."""
    if not isinstance(space, HDivSurface):
        raise TypeError("MaxwellSingleLayerPotentialOperator needs an HDivSurface space")
    return _helmholtz(_capi.MAXWELL_SL, space, space, kappa, intorder, leafsize, eta, eps, method, testhmatrix, mixed,
                      (definedon, definedon))


def MaxwellSingleLayerPotentialOperatorCurl(space, kappa, definedon=None, intorder=3, leafsize=40, eta=2., eps=1e-6,
                                            method="aca", testhmatrix=False, mixed=False):
    """Maxwell single layer on HCurl with the rotated trace and the surface curl
    (src/python_bem.cpp:137-149, DiffOpMaxwell)."""
    if not isinstance(space, HCurl):
        raise TypeError("MaxwellSingleLayerPotentialOperatorCurl needs an HCurl space")
    return _helmholtz(_capi.MAXWELL_SL, space, space, kappa, intorder, leafsize, eta, eps, method, testhmatrix, mixed,
                      (definedon, definedon))


def MaxwellDoubleLayerPotentialOperator(trial_space, test_space, kappa, trial_definedon=None, test_definedon=None,
                                        intorder=3, leafsize=40, eta=2., eps=1e-6, method="aca", testhmatrix=False,
                                        mixed=False):
    """Maxwell double layer: rotated HCurl trace (trial) against the HDivSurface
    trace (test) (src/python_bem.cpp:154-173, src/ngbem.hpp:482-510)."""
    if not (isinstance(trial_space, HCurl) and isinstance(test_space, HDivSurface)):
        raise TypeError("MaxwellDoubleLayerPotentialOperator needs an HCurl trial and an HDivSurface test space")
    dense = _dense_remap(method, leafsize, eta)
    return IntegralOperatorC(_capi.MAXWELL_DL, kappa, trial_space, test_space, intorder, leafsize, eta, eps, dense,
                             mixed, testhmatrix, definedon=(trial_definedon, test_definedon))
