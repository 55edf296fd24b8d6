"""Electromagnetic Mie series for a perfectly conducting sphere (test oracle).

Textbook expansion in vector spherical wave functions (Bohren & Huffman,
"Absorption and Scattering of Light by Small Particles", ch. 4; time factor
exp(-i w t)) of the plane wave E_inc = x exp(i k z) and of its scattered
field for a perfect conductor of radius a:

  E_inc = sum_n E_n (M1_o1n - i N1_e1n),           E_n = i^n (2n+1) / (n (n+1))
  E_sca = sum_n E_n (i a_n N3_e1n - b_n M3_o1n),    a_n = psi_n'(ka) / xi_n'(ka),
                                                      b_n = psi_n(ka) / xi_n(ka)

with psi_n(x) = x j_n(x), xi_n(x) = x h1_n(x); curl M = k N and curl N = k M.
The surface current in the normalisation of the boundary integral equations
V j = -<E_inc, phi> is j = n x curl E_total / k (the reference compares
kappa * j with the Neumann trace n x curl E, tests/test_Maxwell_Mie.py:40-43
and tests/mie.cpp:191-210).  The expansion is checked in the tests against the
plane wave itself, the vanishing tangential total field on the sphere and a
finite-difference curl.
"""
import numpy as np
from scipy.special import spherical_jn, spherical_yn


def _pi_tau(mu, nmax):
    """Angular functions pi_n = P_n^1 / sin, tau_n = d P_n^1 / d theta (B&H 4.47)."""
    pi = np.zeros((nmax + 1,) + mu.shape)
    tau = np.zeros_like(pi)
    pi[1] = 1.0
    for n in range(2, nmax + 1):
        pi[n] = (2 * n - 1) / (n - 1) * mu * pi[n - 1] - n / (n - 1) * pi[n - 2]
    for n in range(1, nmax + 1):
        tau[n] = n * mu * pi[n] - (n + 1) * pi[n - 1]
    return pi, tau


def _radial(kind, n, rho):
    """z_n(rho) and [rho z_n(rho)]' / rho for j_n (kind 1) or h1_n (kind 3)."""
    j, dj = spherical_jn(n, rho), spherical_jn(n, rho, derivative=True)
    if kind == 1:
        z, dz = j + 0j, dj + 0j
    else:
        z = j + 1j * spherical_yn(n, rho)
        dz = dj + 1j * spherical_yn(n, rho, derivative=True)
    return z, (z + rho * dz) / rho


def _mn(kind, n, rho, cos_p, sin_p, sin_t, pi_n, tau_n, rho_rad=None):
    """Spherical components (r, theta, phi) of M_o1n, M_e1n, N_o1n, N_e1n;
    the radial functions at rho_rad (a scalar when every point has one radius)."""
    z, dz = _radial(kind, n, rho if rho_rad is None else rho_rad)
    zeros = np.zeros_like(rho, dtype=complex)
    M_o = np.stack([zeros, cos_p * pi_n * z, -sin_p * tau_n * z])
    M_e = np.stack([zeros, -sin_p * pi_n * z, -cos_p * tau_n * z])
    nr = n * (n + 1) * sin_t * pi_n * z / rho
    N_o = np.stack([sin_p * nr, sin_p * tau_n * dz, cos_p * pi_n * dz])
    N_e = np.stack([cos_p * nr, cos_p * tau_n * dz, -sin_p * pi_n * dz])
    return M_o, M_e, N_o, N_e


def _frame(X):
    r = np.linalg.norm(X, axis=-1)
    ct = np.clip(X[..., 2] / r, -1.0, 1.0)
    st = np.sqrt(1.0 - ct ** 2)
    ph = np.arctan2(X[..., 1], X[..., 0])
    cp, sp = np.cos(ph), np.sin(ph)
    er = np.stack([st * cp, st * sp, ct], -1)
    et = np.stack([ct * cp, ct * sp, -st], -1)
    ep = np.stack([-sp, cp, np.zeros_like(sp)], -1)
    return r, ct, st, cp, sp, er, et, ep


def fields(X, k, a=1.0, nmax=None, parts=("inc", "sca")):
    """E and curl E (cartesian, complex) at points X (..., 3), for the plane
    wave x exp(i k z) on the conducting sphere of radius a."""
    if nmax is None:
        nmax = int(k * max(a, 1.0) + 4 * (k * max(a, 1.0)) ** (1 / 3) + 12)
    r, ct, st, cp, sp, er, et, ep = _frame(X)
    pi, tau = _pi_tau(ct, nmax)
    rho = k * r
    # points on one sphere: the radial functions once per order, not per point
    rho_rad = rho.flat[0] if rho.size and np.ptp(rho) <= 1e-13 * abs(rho.flat[0]) else None
    E = np.zeros((3,) + r.shape, dtype=complex)
    C = np.zeros_like(E)
    x = k * a
    for n in range(1, nmax + 1):
        En = (1j ** n) * (2 * n + 1) / (n * (n + 1))
        if "inc" in parts:
            M_o, M_e, N_o, N_e = _mn(1, n, rho, cp, sp, st, pi[n], tau[n], rho_rad)
            E += En * (M_o - 1j * N_e)
            C += En * k * (N_o - 1j * M_e)
        if "sca" in parts:
            psi, dpsi = x * spherical_jn(n, x), spherical_jn(n, x) + x * spherical_jn(n, x, derivative=True)
            h = spherical_jn(n, x) + 1j * spherical_yn(n, x)
            dh = spherical_jn(n, x, derivative=True) + 1j * spherical_yn(n, x, derivative=True)
            xi, dxi = x * h, h + x * dh
            an, bn = dpsi / dxi, psi / xi
            M_o, M_e, N_o, N_e = _mn(3, n, rho, cp, sp, st, pi[n], tau[n], rho_rad)
            E += En * (1j * an * N_e - bn * M_o)
            C += En * k * (1j * an * M_e - bn * N_o)
    to_cart = lambda F: (F[0][..., None] * er + F[1][..., None] * et + F[2][..., None] * ep)  # noqa: E731
    return to_cart(E), to_cart(C)


def current(X, k, a=1.0):
    """j = n x curl E_total / k on the sphere surface (points X on |x| = a)."""
    _, C = fields(X, k, a)
    n = X / np.linalg.norm(X, axis=-1)[..., None]
    return np.cross(n, C) / k


def rotation_to_z(d, p):
    """Rotation R with R d = z and R p = x (d, p orthonormal)."""
    d = d / np.linalg.norm(d)
    p = p - (p @ d) * d
    p = p / np.linalg.norm(p)
    return np.stack([p, np.cross(d, p), d])


def current_rotated(X, k, d, p, a=1.0):
    """Current for the plane wave p exp(i k d.x): rotate into the frame where
    the wave is x exp(i k z), evaluate, rotate back."""
    R = rotation_to_z(d, p)
    return current(X @ R.T, k, a) @ R
