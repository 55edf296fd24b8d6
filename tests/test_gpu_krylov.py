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


@pytest.mark.parametrize("dt", [torch.float64, torch.complex128], ids=["f64", "c128"])
def test_arnoldi_step_with_diagonal_scale(gpu, dt):
    """dscale = d: the step orthogonalises d * w (a diagonal left
    preconditioner applied inside the step) exactly as if w had been scaled
    first, to the last bit."""
    n, k = 40000, 5
    g = torch.Generator(device="cpu").manual_seed(7)
    Q, _ = torch.linalg.qr(torch.randn(n, k + 1, dtype=dt, generator=g))
    w0 = torch.randn(n, dtype=dt, generator=g)
    d = torch.rand(n, dtype=torch.float64, generator=g) + 0.5
    cols = []
    for fused in (False, True):
        arn = K._DeviceArnoldi(k + 1, n, dt, gpu)
        arn.V[: k + 1] = Q.T.contiguous().to(gpu)
        if fused:
            arn.step(k, w0.to(gpu).clone(), d.to(gpu))
        else:
            arn.step(k, (w0 * d).to(gpu))
        cols.append((torch.from_numpy(arn.column(k).copy()), arn.V[k + 1].cpu()))
    assert torch.equal(cols[0][0], cols[1][0]) and torch.equal(cols[0][1], cols[1][1])


def test_gmres_fused_mass_inverse_preconditioner(gpu):
    """GMRES with MassMatrix(SurfaceL2 p0).Inverse() as preconditioner (fused
    into the Arnoldi step) against the same preconditioner as a plain
    function: same iterations, same solution."""
    import ngbem_amd as nb

    mesh = nb.icosphere(3)
    fes = nb.SurfaceL2(mesh, order=0)
    C = nb.HelmholtzCombinedFieldOperator(fes, fes, kappa=3.0, intorder=4, leafsize=32, eps=1e-8)
    A = 0.5 * nb.MassMatrix(fes) + C.mat
    Minv = nb.MassMatrix(fes).Inverse()
    assert Minv._diagonal is not None
    area = torch.from_numpy(mesh.areas()).to(gpu)
    rng = np.random.default_rng(2)
    f = torch.from_numpy(rng.standard_normal(fes.ndof) + 1j * rng.standard_normal(fes.ndof)).to(gpu)
    u1 = K.GMRes(A, f, pre=Minv, tol=1e-10, maxsteps=200, printrates=False)
    it1 = K.GMRes.iterations
    u2 = K.GMRes(A, f, pre=lambda r: r / area, tol=1e-10, maxsteps=200, printrates=False)
    it2 = K.GMRes.iterations
    assert it1 == it2
    assert (torch.linalg.norm(u1 - u2) / torch.linalg.norm(u2)).item() < 1e-12


@pytest.mark.parametrize("dt", [torch.float64, torch.complex128], ids=["f64", "c128"])
def test_arnoldi_cycle_start(gpu, dt):
    """k = -1 (the start of a cycle): V[0] = D r / |D r| and |D r| in the
    pinned beta slot, against torch."""
    n = 30001
    g = torch.Generator(device="cpu").manual_seed(3)
    r = torch.randn(n, dtype=dt, generator=g)
    d = torch.rand(n, dtype=torch.float64, generator=g) + 0.5
    for dsc in (None, d):
        arn = K._DeviceArnoldi(4, n, dt, gpu)
        arn.start(r.to(gpu).clone(), None if dsc is None else dsc.to(gpu))
        arn.beta_ev.synchronize()
        ref = r if dsc is None else r * dsc
        nrm = torch.linalg.norm(ref).item()
        assert abs(arn.beta_h[0].real.item() - nrm) <= 1e-13 * nrm
        assert (arn.V[0].cpu() - ref / nrm).abs().max().item() < 1e-14
