"""Batches of operators and differentiable single-layer applies.

``OperatorBatch``: the block-diagonal union of many assembled operators
(different geometries, sizes and wavenumbers).  An apply is one launch of each
apply kernel for the whole batch, captured once in a hipGraph and replayed
(``bem_batch_apply``); vectors are the concatenation of the instances'
vectors.

``assemble_many``: build many operators on several host threads, each with its
own GPU stream, so the instances' assemblies overlap on the device.

``SingleLayerBatch``: differentiable y_i = V_i(kappa_i) x_i over a batch of
Helmholtz single layers (one multi-geometry operator for all of them), a
``torch.autograd.Function`` whose backward applies the adjoint (gradient
w.r.t. x) and the d/dkappa operator (gradient w.r.t. kappa).  None of this has a counterpart in the reference,
which builds one ``GenericIntegralOperator`` at a time (src/ngbem.cpp:314-330)
and is not differentiable.
"""
from __future__ import annotations

import ctypes
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import torch

from . import _capi
from .operators import (BaseMatrix, HelmholtzSingleLayerPotentialOperator,
                        HelmholtzSingleLayerPotentialOperatorDKappa, _materialized, _raw_view, device)

__all__ = ["OperatorBatch", "assemble_many", "SingleLayerBatch"]


def assemble_many(build, items, threads=8):
    """``[build(item) for item in items]``, evaluated on ``threads`` host
    threads, each on its own (default-priority) GPU stream."""
    dev = device()

    def run(item):
        with torch.cuda.device(dev):
            s = torch.cuda.Stream(dev)
            with torch.cuda.stream(s):
                op = build(item)
            s.synchronize()
            return op

    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        return list(ex.map(run, items))


class OperatorBatch(BaseMatrix):
    """Block-diagonal batch of assembled operators (all real or all complex).

    ``ops``: IntegralOperator / IntegralOperatorC objects.  Their storage is
    copied into the batch, so they may be dropped afterwards.  ``mixed``
    selects the FP32 / complex64 copies (all operators built with mixed=True).
    ``graph=False`` launches the kernels directly instead of replaying a graph.
    ``consume=True`` frees each operator's block storage as it is copied (the
    operators can then no longer be applied); peak memory is the batch plus
    one operator instead of twice the batch.
    """

    def __init__(self, ops, mixed=False, graph=True, consume=False):
        ops = list(ops)
        L = _capi.lib()
        arr = (ctypes.c_void_p * len(ops))(*[o._h for o in ops])
        out = ctypes.c_void_p()
        _capi.check(L.bem_batch_create(arr, len(ops), _capi.BATCH_CONSUME if consume else 0, ctypes.byref(out)),
                    "bem_batch_create")
        self._h = out
        w, h, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _capi.check(L.bem_batch_sizes(self._h, ctypes.byref(w), ctypes.byref(h), ctypes.byref(c)))
        self.width, self.height, self.is_complex = w.value, h.value, bool(c.value)
        self.col_offsets = np.cumsum([0] + [o.trial_space.ndof for o in ops])
        self.row_offsets = np.cumsum([0] + [o.test_space.ndof for o in ops])
        self._mixed = bool(mixed)
        self._graph = bool(graph)
        self._keep = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _capi.lib().bem_batch_destroy(h)
            self._h = None

    def view(self, mixed=None, graph=None):
        """Same batch, other apply options (shares the device storage)."""
        v = OperatorBatch.__new__(OperatorBatch)
        v.__dict__.update(self.__dict__)
        v._keep = self  # the handle belongs to self
        v._h = None
        v._hh = self._h
        if mixed is not None:
            v._mixed = bool(mixed)
        if graph is not None:
            v._graph = bool(graph)
        return v

    @property
    def handle(self):
        return self._h if self._h is not None else self._hh

    def _apply(self, flags, s, x, y):
        trans = flags & _capi.APPLY_TRANS
        nin, nout = (self.height, self.width) if trans else (self.width, self.height)
        if x.shape != (nin,) or y.shape != (nout,):
            raise ValueError(f"vector sizes {tuple(x.shape)}, {tuple(y.shape)} do not match {nout}x{nin}")
        if self._mixed:
            flags |= _capi.APPLY_MIXED
        if not self._graph:
            flags |= _capi.BATCH_EAGER
        dt = torch.complex128 if self.is_complex else torch.float64
        if not self.is_complex and (torch.is_complex(x) or torch.is_complex(y)):
            raise TypeError("real batch needs real vectors")
        xc = _materialized(x.to(device=y.device, dtype=dt)).contiguous()
        yc = y if _raw_view(y) else _materialized(y).contiguous().clone()
        a = np.array([complex(s).real, complex(s).imag])
        one = np.array([1.0, 0.0])
        with torch.cuda.device(y.device):
            stream = torch.cuda.current_stream().cuda_stream
            _capi.check(_capi.lib().bem_batch_apply(self.handle, flags, _capi.dptr(a), ctypes.c_void_p(xc.data_ptr()),
                                                    _capi.dptr(one), ctypes.c_void_p(yc.data_ptr()),
                                                    _capi.C128 if self.is_complex else _capi.F64,
                                                    ctypes.c_void_p(stream)), "bem_batch_apply")
        if yc is not y:
            y.copy_(yc)

    def MultAdd(self, s, x, y):
        self._apply(0, s, x, y)

    def MultTransAdd(self, s, x, y):
        self._apply(_capi.APPLY_TRANS, s, x, y)

    def MultAdjAdd(self, s, x, y):
        self._apply(_capi.APPLY_TRANS | _capi.APPLY_CONJ, s, x, y)

    @property
    def H(self):
        outer = self

        class _Adj(BaseMatrix):
            height, width, is_complex = outer.width, outer.height, outer.is_complex

            def MultAdd(self, s, x, y):
                outer.MultAdjAdd(s, x, y)

        return _Adj()

    def split(self, v, rows=True):
        """Per-instance views of a concatenated vector."""
        off = self.row_offsets if rows else self.col_offsets
        return [v[off[i]:off[i + 1]] for i in range(len(off) - 1)]


class _BatchApply(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, kappa, batch, dbatch, seg):
        ctx.batch, ctx.dbatch, ctx.seg = batch, dbatch, seg
        ctx.x_real = not torch.is_complex(x)
        xc = x.detach().to(torch.complex128)
        ctx.save_for_backward(xc)
        return batch * xc

    @staticmethod
    def backward(ctx, g):
        (xc,) = ctx.saved_tensors
        gx = gk = None
        g = g.to(torch.complex128).contiguous()
        if ctx.needs_input_grad[0]:
            gx = ctx.batch.H * g
            if ctx.x_real:
                gx = gx.real
        if ctx.needs_input_grad[1]:
            dy = ctx.dbatch * xc
            contrib = (g.conj() * dy).real
            gk = torch.zeros(ctx.seg.max().item() + 1, dtype=torch.float64, device=g.device)
            gk.index_add_(0, ctx.seg, contrib)
        return gx, gk, None, None, None


class SingleLayerBatch:
    """Helmholtz single layers V_i(kappa_i) on spaces_i, with their d/dkappa
    operators; ``__call__(x, kappa)`` is the differentiable y = blockdiag(V_i) x
    for concatenated x (kappa: a tensor of the wavenumbers the batch was built
    with, which receives the gradient).

    By default all geometries go into one operator each for V and dV/dkappa
    (``MultiGeometryOperator``: one assembly, one apply for the whole batch);
    ``multi=False`` assembles one operator per geometry (host threads) and
    merges them into ``OperatorBatch`` objects instead."""

    def __init__(self, spaces, kappas, threads=8, multi=True, **kw):
        self.kappas = [float(k) for k in kappas]
        self.ndofs = [s.ndof for s in spaces]
        if multi:
            from .operators import MultiGeometryOperator

            meshes = [s.mesh for s in spaces]
            cls = type(spaces[0])
            if any(type(s) is not cls for s in spaces):
                raise TypeError("all spaces of a batch must be of one kind")
            self._ops = (MultiGeometryOperator(_capi.HELMHOLTZ_SL, meshes, self.kappas, cls, **kw),
                         MultiGeometryOperator(_capi.HELMHOLTZ_SL_DKAPPA, meshes, self.kappas, cls, **kw))
            self.V, self.dV = self._ops[0].mat, self._ops[1].mat
        else:
            items = list(zip(spaces, self.kappas))
            V = assemble_many(lambda it: HelmholtzSingleLayerPotentialOperator(it[0], it[0], kappa=it[1], **kw),
                              items, threads)
            self.V = OperatorBatch(V, consume=True)
            del V
            dV = assemble_many(lambda it: HelmholtzSingleLayerPotentialOperatorDKappa(it[0], it[0], kappa=it[1],
                                                                                      **kw), items, threads)
            self.dV = OperatorBatch(dV, consume=True)
            del dV
        seg = np.repeat(np.arange(len(spaces)), self.ndofs)
        self.seg = torch.from_numpy(seg).to(device())

    def __call__(self, x, kappa):
        k = torch.as_tensor(kappa, dtype=torch.float64)
        if not np.allclose(k.detach().cpu().numpy(), self.kappas, rtol=0, atol=0):
            raise ValueError("kappa differs from the wavenumbers the batch was assembled with")
        return _BatchApply.apply(x, kappa, self.V, self.dV, self.seg)
