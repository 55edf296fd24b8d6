"""TestCompression: low-rank compression of the 1/r matrix between two random
point clouds, the reference's diagnostic (src/python_bem.cpp:177-193,
src/test_compression.cpp).

The clouds are CreateMatrix's (:16-43): points in the unit cube from the C
library's rand() after srand(seed) (the reference's first call in a process
sees seed 1), the y cloud shifted by 1 + sqrt(3)/eta.  method "aca" runs
TestCompressionACA's cross approximation on the GPU (``bem_test_compression_aca``);
"svd" and "tsvd" are LAPACK SVDs through torch on the host, as the reference's
LapackSVD (the GPU solver's SVD resolves the singular values below ~1e-9 of the
largest less accurately, which changes the rank count), "tsvd" with the
stochastic range finder of MyStochasticTSVD (:81-122), its random sketch from
torch's generator.  Returns (rank, Frobenius error, seconds) like the reference.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np
import torch

from . import _capi


def test_points(nx, ny, eta, seed=1):
    X, Y = np.zeros((nx, 3)), np.zeros((ny, 3))
    _capi.check(_capi.lib().bem_test_points(int(nx), int(ny), float(eta), int(seed), _capi.dptr(X), _capi.dptr(Y)),
                "bem_test_points")
    return X, Y


def _tsvd(A, eps, gen):
    p = min(A.shape)
    rank = 5
    while rank < p:
        V = torch.rand(rank, A.shape[1], dtype=A.dtype, device=A.device, generator=gen)
        U1, _, _ = torch.linalg.svd(A @ V.T, full_matrices=False)
        tmp, S, V2 = torch.linalg.svd(U1.T @ A, full_matrices=False)
        if S[rank - 1] < eps:
            small = torch.nonzero(S[1:] < eps)
            if len(small):
                j = int(small[0]) + 1
                U = U1 @ tmp
                return j - 1, U, S, V2
        rank = int(rank * 1.5)
    U, S, V2 = torch.linalg.svd(A, full_matrices=False)
    small = torch.nonzero(S[1:] < eps)
    return (int(small[0]) if len(small) else p), U, S, V2


def TestCompression(nx=100, ny=100, eta=2.0, eps=1e-10, method="svd", seed=1):
    """(rank, error, seconds) of compressing the 1/r matrix between two point
    clouds (src/python_bem.cpp:177-193)."""
    X, Y = test_points(nx, ny, eta, seed)
    if method == "aca":
        r = ctypes.c_int()
        err = np.zeros(1)
        sec = np.zeros(1)
        _capi.check(_capi.lib().bem_test_compression_aca(_capi.dptr(X), nx, _capi.dptr(Y), ny, float(eps),
                                                          ctypes.byref(r), _capi.dptr(err), _capi.dptr(sec)),
                    "bem_test_compression_aca")
        return r.value, float(err[0]), float(sec[0])
    if method not in ("svd", "tsvd"):
        return (0, 0.0)
    Xt, Yt = torch.from_numpy(X), torch.from_numpy(Y)
    A = 1.0 / torch.linalg.norm(Xt[:, None, :] - Yt[None, :, :], dim=-1)  # exact differences, not |x|^2+|y|^2-2xy
    t0 = time.perf_counter()
    if method == "svd":
        U, S, V = torch.linalg.svd(A, full_matrices=False)
        num = int((S.abs() > eps).sum())
    else:
        gen = torch.Generator().manual_seed(seed)
        num, U, S, V = _tsvd(A, eps, gen)
    t = time.perf_counter() - t0
    approx = (U[:, :num] * S[:num]) @ V[:num]
    return num, float(torch.linalg.norm(A - approx)), t
