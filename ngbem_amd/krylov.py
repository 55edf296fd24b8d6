"""Krylov solvers on GPU vectors, with ngsolve.krylovspace's argument names
(CG: mat, rhs, pre, sol, tol, maxsteps, printrates, initialize; GMRes: A, b,
pre, x, maxsteps, tol, printrates, restart), as the reference's scripts call
them (tests/test_Laplace.py:60, demos/HelmholtzBurtonMiller.ipynb cell 9).

``RefinedSolve`` is the mixed-precision form: an inner GMRES on the FP32 /
complex64 copy of the operator, wrapped in FP64 iterative refinement where the
residual b - A x is recomputed with the FP64 operator.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from .operators import BaseMatrix, device


def _dot(a, b):
    return torch.vdot(a, b) if torch.is_complex(a) else torch.dot(a, b)


def _apply(A, x):
    if isinstance(A, BaseMatrix):
        y = torch.empty(x.shape[:-1] + (A.height,),
                        dtype=torch.complex128 if (A.is_complex or torch.is_complex(x)) else torch.float64,
                        device=x.device)
        A.MultScale(1.0, x, y)
        return y
    return A(x)


def CG(mat, rhs, pre=None, sol=None, tol=1e-12, maxsteps=100, printrates=True, initialize=True, conjugate=False):
    """Preconditioned conjugate gradients (relative residual tolerance)."""
    b = rhs
    x = sol if sol is not None else torch.zeros_like(b)
    if initialize:
        x.zero_()
    r = b - _apply(mat, x)
    z = _apply(pre, r) if pre is not None else r.clone()
    p = z.clone()
    rz = _dot(r, z)
    err0 = math.sqrt(abs(float(rz.real if torch.is_complex(rz) else rz)))
    if err0 == 0.0:
        return x
    for it in range(maxsteps):
        Ap = _apply(mat, p)
        alpha = rz / _dot(p, Ap)
        x.add_(alpha * p)
        r.sub_(alpha * Ap)
        z = _apply(pre, r) if pre is not None else r
        rz_new = _dot(r, z)
        err = math.sqrt(abs(float(rz_new.real if torch.is_complex(rz_new) else rz_new)))
        if printrates:
            print(f"CG iteration {it + 1}, residual = {err}")
        if err < tol * err0:
            break
        beta = rz_new / rz
        rz = rz_new
        p = z + beta * p
    CG.iterations = it + 1
    return x


def _givens_column(H, cs, sn, g, k):
    """Apply the previous rotations to column k of H, make the next one and
    update g; returns |g[k+1]| (the residual norm after step k)."""
    for i in range(k):
        hi, hj = H[i, k], H[i + 1, k]
        H[i, k] = np.conj(cs[i]) * hi + np.conj(sn[i]) * hj
        H[i + 1, k] = -sn[i] * hi + cs[i] * hj
    a, bb = H[k, k], H[k + 1, k]
    t = math.hypot(abs(a), abs(bb))
    cs[k], sn[k] = (a / t, bb / t) if t > 0 else (1.0, 0.0)
    H[k, k] = t
    H[k + 1, k] = 0.0
    g[k + 1] = -sn[k] * g[k]
    g[k] = np.conj(cs[k]) * g[k]
    return abs(g[k + 1])


_ARNOLDI_CACHE = {}


def _arnoldi_for(m, n, dt, dev, basis_dt):
    """The device workspace of a GMRES cycle of up to m steps, kept between
    calls (the basis, the pinned Hessenberg buffer and the events are
    allocated once per vector size and dtype, and again only when a longer
    cycle is asked for)."""
    key = (n, dt, basis_dt, str(dev))
    a = _ARNOLDI_CACHE.get(key)
    if a is None or a.m < m:
        _ARNOLDI_CACHE.pop(key, None)
        a = _ARNOLDI_CACHE[key] = _DeviceArnoldi(m, n, dt, dev, basis_dt)
    return a


class _DeviceArnoldi:
    """Basis, Hessenberg columns and workspace of one GMRES cycle on the GPU;
    step(k, w) orthogonalises w = P A v_k against v_0..v_k and writes v_{k+1}
    (library call bem_arnoldi_step), whose last launch writes column k
    straight into pinned host memory; column(k) waits for that launch only
    (an event recorded after it)."""

    def __init__(self, m, n, dt, dev, basis_dt=None):
        from . import _capi

        self._capi = _capi
        self.cplx = dt == torch.complex128
        basis_dt = dt if basis_dt is None else basis_dt
        if basis_dt not in ((torch.complex128, torch.complex64) if self.cplx else (torch.float64, torch.float32)):
            raise ValueError(f"basis dtype {basis_dt} does not fit vectors of dtype {dt}")
        self.dtype_id = {torch.float64: _capi.F64, torch.complex128: _capi.C128, torch.float32: _capi.F32,
                         torch.complex64: _capi.C64}[basis_dt]
        self.V = torch.empty(m + 1, n, dtype=basis_dt, device=dev)
        self.Hh = torch.zeros(m, m + 1, dtype=dt, pin_memory=True)
        nbytes = int(_capi.lib().bem_arnoldi_work_bytes(m, n, self.dtype_id))
        if nbytes < 0:
            raise RuntimeError(_capi.lib().bem_last_error().decode())
        self.work = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        self.events = [torch.cuda.Event() for _ in range(m)]
        self.beta_h = torch.zeros(1, dtype=dt, pin_memory=True)  # |r| at the start of a cycle
        self.beta_ev = torch.cuda.Event()
        self.y_h = torch.zeros(m, dtype=dt, pin_memory=True)  # the cycle's coefficients, on their way to the device
        self.n = n
        self.m = m

    def start(self, r, dscale=None):
        """V[0] = D r / |D r| (r overwritten with D r) and |D r| on its way
        to beta_h (event beta_ev); D = diag(dscale) or the identity."""
        self.step(-1, r, dscale)

    def step(self, k, w, dscale=None):
        """dscale: None, or a float64 device vector d (w is replaced by d * w
        before the orthogonalisation: a diagonal left preconditioner)."""
        import ctypes

        w = w.contiguous()
        if dscale is not None and (dscale.dtype != torch.float64 or dscale.shape != (self.n,)
                                   or dscale.device != w.device or not dscale.is_contiguous()):
            raise ValueError("dscale must be a contiguous float64 vector of n entries on the device of w")
        stream = torch.cuda.current_stream(w.device).cuda_stream
        self._capi.check(self._capi.lib().bem_arnoldi_step(
            ctypes.c_void_p(self.V.data_ptr()), self.n, k, self.n, ctypes.c_void_p(w.data_ptr()),
            ctypes.c_void_p(dscale.data_ptr() if dscale is not None else None),
            ctypes.c_void_p((self.Hh[k] if k >= 0 else self.beta_h).data_ptr()),
            ctypes.c_void_p(self.work.data_ptr()), self.work.numel(), self.dtype_id, ctypes.c_void_p(stream)),
            "bem_arnoldi_step")
        (self.events[k] if k >= 0 else self.beta_ev).record()

    def column(self, k):
        self.events[k].synchronize()
        return self.Hh[k, : k + 2].numpy()


def GMRes(A, b, pre=None, freedofs=None, x=None, maxsteps=100, tol=None, innerproduct=None, callback=None,
          restart=None, startiteration=0, printrates=True, reltol=None, basis=None):
    """Restarted GMRES, left-preconditioned like ngsolve.krylovspace.GMRes: it
    minimises ||pre (b - A x)|| over the Krylov space of pre A and stops when
    that norm has fallen to tol times its value at the start.  A preconditioner
    that is zero on some dofs (a mass inverse on ``FreeDofs()``) therefore
    drops those rows of the system and leaves x unchanged there, which is how
    the reference's mixed Dirichlet / Neumann scripts impose the given data
    (tests/Laplace_Convergence.py:57-61).  Without ``pre`` the criterion is
    ||b - A x|| <= tol ||b - A x0||.

    On GPU vectors the Krylov basis is one (m+1) x n tensor and each Arnoldi
    step (classical Gram-Schmidt twice, the norm and the next basis vector)
    runs on the device without a read back (bem_arnoldi_step); the host applies
    the Givens rotations to each Hessenberg column one step later, when its
    asynchronous copy has arrived, so the GPU never waits for the host.  When
    step k's column shows convergence, step k + 1 has already been issued and
    is discarded; to avoid that, once the observed convergence rate predicts
    that the step just issued reaches the tolerance, the host waits for its
    column before issuing the next one.  Restart cycles are at most 1022
    steps long.  ``basis`` (an extension, GPU vectors only): the dtype of the
    Krylov basis, e.g. torch.complex64 for the inner solve of a
    mixed-precision refinement (products still accumulate in FP64).
    """
    if freedofs is not None or innerproduct is not None:
        raise NotImplementedError("freedofs / innerproduct are not supported")
    tol = 1e-7 if tol is None else tol
    if reltol is not None:
        tol = reltol
    dt = torch.complex128 if (torch.is_complex(b) or getattr(A, "is_complex", False)) else torch.float64
    b = b.to(dt)
    x_zero = x is None
    if x is None:
        x = torch.zeros_like(b)
    m = min(restart or maxsteps, maxsteps, 1022)
    n = b.shape[0]
    P = (lambda v: _apply(pre, v)) if pre is not None else (lambda v: v)  # noqa: E731
    dev_path = b.is_cuda
    # a diagonal preconditioner (the inverse of a diagonal mass matrix) is
    # applied inside the device Arnoldi step instead of by a launch of its own
    dscale = getattr(pre, "_diagonal", None) if dev_path else None
    if dscale is not None and (dscale.device != b.device or dscale.shape != (b.shape[0],)):
        dscale = None
    if basis is not None and not dev_path:
        raise ValueError("a basis dtype needs GPU vectors")
    arn = _arnoldi_for(m, n, dt, b.device, basis) if dev_path else None
    Vb = arn.V if dev_path else torch.empty(m + 1, n, dtype=dt, device=b.device)
    cdt = np.complex128
    total = 0
    bnorm = None  # preconditioned residual norm at the start
    if not x_zero:
        x_zero = not bool(torch.any(x != 0))
    # (b = 0 with x = 0: beta = 0 below ends the solve at x = 0)
    while total < maxsteps:
        r = b if x_zero else b - _apply(A, x)
        if not (dev_path and dscale is not None):
            r = P(r)  # else applied by the cycle's first Arnoldi launch
        r = r.to(dt)
        x_zero = False
        H = np.zeros((m + 1, m), dtype=cdt)
        cs = np.zeros(m, dtype=cdt)
        sn = np.zeros(m, dtype=cdt)
        g = np.zeros(m + 1, dtype=cdt)
        if dev_path:
            # |r| stays on the device: the first step is issued before the
            # host reads it (written to pinned memory, read at the first column)
            arn.start(r if r is not b else r.clone(), dscale)
            beta = None
        else:
            beta = float(torch.linalg.norm(r))
            if bnorm is None:
                bnorm = beta
            if beta <= tol * bnorm:
                break
            Vb[0].copy_(r / beta)
            g[0] = beta
        k_used = 0
        stop = False
        done_at_start = False

        hist = []

        def take(k):
            # column k has arrived: rotate, report, decide
            nonlocal k_used, total, stop, beta, bnorm, done_at_start
            if beta is None:
                arn.beta_ev.synchronize()
                beta = float(arn.beta_h[0].real)
                g[0] = beta
                if bnorm is None:
                    bnorm = beta
                if beta <= tol * bnorm:
                    # converged before this cycle (or b = 0): the issued steps are dropped
                    done_at_start = stop = True
                    return
            hcol = arn.column(k) if dev_path else hcols[k]
            H[: k + 2, k] = hcol
            hn = float(hcol[k + 1].real)
            res = _givens_column(H, cs, sn, g, k)
            hist.append(res)
            k_used = k + 1
            total += 1
            if printrates:
                print(f"GMRes iteration {total}, residual = {res}")
            if callback is not None:
                callback(res)
            stop = res <= tol * bnorm or hn == 0.0 or total >= maxsteps

        def likely_last():
            # the residual after the step just issued, extrapolated at the last rate
            if len(hist) < 2 or hist[-2] == 0.0:
                return False
            return hist[-1] * (hist[-1] / hist[-2]) <= 2.0 * tol * bnorm

        hcols = {}
        steps = min(m, maxsteps - total)
        for k in range(steps):
            v = Vb[k] if Vb.dtype == dt else Vb[k].to(dt)
            if dscale is not None:
                w = _apply(A, v).to(dt)
            else:
                w = P(_apply(A, v)).to(dt)
            if dev_path:
                arn.step(k, w, dscale)
                if k >= 1 and k_used == k - 1:
                    take(k - 1)
                    if stop:
                        break
                if (likely_last() or total + 1 >= maxsteps) and k + 1 < steps:
                    take(k)
                    if stop:
                        break
            else:
                Vk = Vb[: k + 1]
                h = Vk.conj() @ w
                w = w - h @ Vk
                h2 = Vk.conj() @ w  # second pass (CGS2)
                w = w - h2 @ Vk
                hn_t = torch.linalg.norm(w)
                hcols[k] = torch.cat([h + h2, hn_t.reshape(1).to(h.dtype)]).numpy()
                hn = float(hcols[k][k + 1].real)
                Vb[k + 1] = w / hn if hn > 0 else 0.0
                take(k)
                if stop:
                    break
        else:
            if dev_path and steps > 0 and k_used < steps:
                take(steps - 1)
        if done_at_start:
            break
        y = np.linalg.solve(np.triu(H[:k_used, :k_used]), g[:k_used])
        y = y if torch.is_complex(x) else y.real.copy()
        if dev_path:
            # through pinned memory: an asynchronous copy (the next write to
            # y_h follows a column read, which waits for later launches)
            arn.y_h[:k_used].copy_(torch.from_numpy(y))
            yt = arn.y_h[:k_used].to(device=x.device, non_blocking=True).to(Vb.dtype)
        else:
            yt = torch.from_numpy(y).to(device=x.device, dtype=Vb.dtype)
        x.add_(yt @ Vb[:k_used])
        if abs(g[k_used]) <= tol * bnorm:
            break
    GMRes.iterations = total
    return x


def GMResBatch(A, B, pre=None, maxsteps=100, tol=1e-7, restart=None, printrates=False):
    """GMRES for the k right-hand sides in the rows of B (k x n), each with its
    own Krylov space, all advanced together: every step applies A to the k new
    basis vectors as one block product (the matrix-core path of the H-matrix)
    and orthogonalises them with batched CGS2.  Left-preconditioned like
    :func:`GMRes` (a ``pre`` that is zero on some dofs drops those rows); stops
    when every row has ||pre r_i|| <= tol ||pre b_i||; returns X (k x n)."""
    dt = torch.complex128 if (torch.is_complex(B) or getattr(A, "is_complex", False)) else torch.float64
    B = B.to(dt)
    k, n = B.shape
    X = torch.zeros_like(B)
    m = min(restart or maxsteps, maxsteps)
    P = (lambda v: _apply(pre, v)) if pre is not None else (lambda v: v)  # noqa: E731
    V = torch.empty(m + 1, k, n, dtype=dt, device=B.device)
    total = 0
    res = np.full(k, np.inf)
    bnorm = None  # per row: the preconditioned residual norm at the start
    while total < maxsteps:
        R = P(B - _apply(A, X)).to(dt)
        beta = torch.linalg.norm(R, dim=1)
        beta_h = beta.cpu().numpy()
        if bnorm is None:
            bnorm = beta_h.copy()
            bnorm[bnorm == 0] = 1.0
        if np.all(beta_h <= tol * bnorm):
            res = beta_h / bnorm
            break
        V[0] = R / torch.clamp(beta, min=1e-300)[:, None]
        H = np.zeros((k, m + 1, m), dtype=np.complex128)
        cs = np.zeros((k, m), dtype=np.complex128)
        sn = np.zeros((k, m), dtype=np.complex128)
        g = np.zeros((k, m + 1), dtype=np.complex128)
        g[:, 0] = beta_h
        used = 0
        for j in range(m):
            w = P(_apply(A, V[j])).to(dt)
            Vk = V[: j + 1].transpose(0, 1).contiguous()  # k x (j+1) x n
            h = torch.bmm(Vk.conj(), w.unsqueeze(2)).squeeze(2)  # k x (j+1)
            w = w - torch.bmm(h.unsqueeze(1), Vk).squeeze(1)
            h2 = torch.bmm(Vk.conj(), w.unsqueeze(2)).squeeze(2)
            w = w - torch.bmm(h2.unsqueeze(1), Vk).squeeze(1)
            hn_t = torch.linalg.norm(w, dim=1)
            hcol = torch.cat([h + h2, hn_t.to(h.dtype)[:, None]], 1).cpu().numpy()
            hn = hcol[:, j + 1].real.copy()
            H[:, : j + 2, j] = hcol
            for i in range(j):
                hi, hj = H[:, i, j].copy(), H[:, i + 1, j].copy()
                H[:, i, j] = np.conj(cs[:, i]) * hi + np.conj(sn[:, i]) * hj
                H[:, i + 1, j] = -sn[:, i] * hi + cs[:, i] * hj
            a, bb = H[:, j, j].copy(), H[:, j + 1, j].copy()
            t = np.hypot(np.abs(a), np.abs(bb))
            ok = t > 0
            cs[:, j] = np.where(ok, a / np.where(ok, t, 1.0), 1.0)
            sn[:, j] = np.where(ok, bb / np.where(ok, t, 1.0), 0.0)
            H[:, j, j] = t
            H[:, j + 1, j] = 0.0
            g[:, j + 1] = -sn[:, j] * g[:, j]
            g[:, j] = np.conj(cs[:, j]) * g[:, j]
            used = j + 1
            total += 1
            res = np.abs(g[:, j + 1]) / bnorm
            if printrates:
                print(f"GMResBatch iteration {total}, max relative residual = {res.max():.3e}")
            if np.all(res <= tol) or total >= maxsteps:
                break
            V[j + 1] = w / torch.from_numpy(np.maximum(hn, 1e-300)).to(w.device)[:, None]
        Y = np.zeros((k, used), dtype=np.complex128)
        for i in range(k):
            Hi = np.triu(H[i, :used, :used])
            d = np.abs(np.diag(Hi))
            if np.all(d > 0):
                Y[i] = np.linalg.solve(Hi, g[i, :used])
            else:  # converged early (exact breakdown): use the leading nonsingular part
                q = int(np.argmax(d == 0))
                if q:
                    Y[i, :q] = np.linalg.solve(Hi[:q, :q], g[i, :q])
        Yt = torch.from_numpy(Y if torch.is_complex(X) else Y.real.copy()).to(device=X.device, dtype=dt)
        basis = V[:used].transpose(0, 1).contiguous()  # k x used x n
        X.add_(torch.bmm(Yt.unsqueeze(1), basis).squeeze(1))
        if np.all(res <= tol):
            break
    GMResBatch.iterations = total
    GMResBatch.residuals = res
    return X


def RefinedSolve(A, A_inner, b, pre=None, tol=1e-10, inner_tol=1e-5, maxouter=20, maxsteps=500, printrates=False,
                 basis=None):
    """FP64 iterative refinement around an inner GMRES on A_inner (typically
    ``A.mixed``): x += GMRes(A_inner, b - A x).  Returns (x, history).
    ``basis``: Krylov basis dtype of the inner GMRES (default: complex64 /
    float32 on GPU vectors, the precision of the mixed operator)."""
    x = torch.zeros_like(b)
    bnorm = float(torch.linalg.norm(b))
    hist = []
    inner_total = 0
    if basis is None and b.is_cuda:
        basis = torch.complex64 if (torch.is_complex(b) or getattr(A, "is_complex", False)) else torch.float32
    for outer in range(maxouter):
        r = b - _apply(A, x) if outer > 0 else b.clone()
        rn = float(torch.linalg.norm(r)) / bnorm
        hist.append(rn)
        if printrates:
            print(f"refinement step {outer}: relative residual {rn:.3e}")
        if rn <= tol:
            break
        d = GMRes(A_inner, r, pre=pre, maxsteps=maxsteps, tol=inner_tol, printrates=False, basis=basis)
        inner_total += GMRes.iterations
        x.add_(d)
    RefinedSolve.inner_iterations = inner_total
    return x, hist
