"""Device Arnoldi step (bem_arnoldi_step) and the GMRES driver built on it.

The step against classical Gram-Schmidt twice in torch (float64 / complex128):
Hessenberg column to 1e-12, the new basis vector orthogonal to the old ones
and of unit length, bit-identical on a repeat (fixed-order reductions);
GMRes on GPU vectors against a direct solve, restarted and not, real and
complex, with the lagged convergence check stopping after the same number of
steps as the host-synchronous iteration on CPU tensors."""
import numpy as np
import pytest
import torch

import ngbem_amd.krylov as K
from ngbem_amd.operators import BaseMatrix

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dt", [torch.float64, torch.complex128], ids=["f64", "c128"])
@pytest.mark.parametrize("n,k", [(1000, 0), (81920, 7), (5000, 40), (70001, 130)])
def test_arnoldi_step_matches_cgs2(gpu, dt, n, k):
    g = torch.Generator(device="cpu").manual_seed(n + k)
    Q, _ = torch.linalg.qr(torch.randn(n, k + 1, dtype=dt, generator=g))
    w0 = torch.randn(n, dtype=dt, generator=g)
    Vk = Q.T.contiguous()
    h = Vk.conj() @ w0
    w1 = w0 - h @ Vk
    h2 = Vk.conj() @ w1
    w2 = w1 - h2 @ Vk
    hn = torch.linalg.norm(w2)
    ref = torch.cat([h + h2, hn.reshape(1).to(dt)])
    outs = []
    for rep in range(2):
        arn = K._DeviceArnoldi(k + 1, n, dt, gpu)
        arn.V[: k + 1] = Vk.to(gpu)
        arn.step(k, w0.to(gpu).clone())
        col = torch.from_numpy(arn.column(k).copy())
        outs.append((col, arn.V[k + 1].cpu()))
        assert (col - ref).abs().max().item() <= 1e-12 * ref.abs().max().item()
        v = arn.V[k + 1].cpu()
        assert abs(torch.linalg.norm(v).item() - 1.0) < 1e-13
        assert (Vk.conj() @ v).abs().max().item() < 1e-13
        assert (v - w2 / hn).abs().max().item() < 1e-12
    assert torch.equal(outs[0][0], outs[1][0]) and torch.equal(outs[0][1], outs[1][1])


class _Dense(BaseMatrix):
    def __init__(self, A):
        self.A = A
        self.height, self.width = A.shape
        self.is_complex = torch.is_complex(A)

    def MultAdd(self, s, x, y):
        y.add_(s * (x @ self.A.T))


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("restart", [None, 12])
def test_device_gmres_matches_direct_solve(gpu, cplx, restart):
    rng = np.random.default_rng(11)
    n = 300
    A = np.eye(n) * 3 + rng.standard_normal((n, n)) / np.sqrt(n)
    b = rng.standard_normal(n)
    if cplx:
        A = A + 1j * rng.standard_normal((n, n)) / np.sqrt(n)
        b = b + 1j * rng.standard_normal(n)
    At, bt = torch.from_numpy(A), torch.from_numpy(b)
    x_ref = torch.linalg.solve(At, bt)
    x = K.GMRes(_Dense(At.to(gpu)), bt.to(gpu), tol=1e-12, maxsteps=300, restart=restart, printrates=False)
    its_gpu = K.GMRes.iterations
    assert (torch.linalg.norm(x.cpu() - x_ref) / torch.linalg.norm(x_ref)).item() < 1e-10
    xc = K.GMRes(_Dense(At), bt, tol=1e-12, maxsteps=300, restart=restart, printrates=False)
    assert abs(K.GMRes.iterations - its_gpu) <= 1
    assert (torch.linalg.norm(xc - x_ref) / torch.linalg.norm(x_ref)).item() < 1e-10
