"""Closed-form solutions of the timestep (src/var_diff.f90 loops 1-14) for
states where the nonlinear terms collapse, used to pin both the oracle and the
GPU modes independently of any implementation of the scheme.

Single c mode (var = 1, uniform phi = a): h = 0 (the h sum runs over the first
var-1 grains), so df/dc = 2 ap (c - c_p) is linear; the mobility is uniform,
m - D_mean = mu0 = 2 (1 - a^2) D_gb, and the variable-mobility correction
(loops 10-12) multiplies the mode by dt mu0 s1 (2 ap s1 - kappa_c k3).  A
cosine mode of wavevector k evolves by the real factor returned below.

Allen-Cahn harmonics (var = 1): with one grain the gamma coupling and the h
derivative vanish, psi = phi^3 - phi, and a + e cos(theta) produces the
harmonics cos(n theta), n <= 3, with known coefficients.

Gradient energy of a single mode: f_int = 1/2 sum_k |kf|^2 (kappa_phi |phi_k|^2
+ kappa_c |c_k|^2) * P, with the Nyquist-zeroed kf (src/k_space.f90:51-59).
"""
import math

import numpy as np


def _axis_a(m, n, zero_nyq):
    if zero_nyq and 2 * m == n:
        return 0.0
    ms = m if m <= n // 2 else m - n
    return ms * (2 * math.pi / n)


def theta(mode, shape):
    nx, ny, nz = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return 2 * np.pi * (mode[0] * x / nx + mode[1] * y / ny + mode[2] * z / nz)


def ch_mode_factor(p, mode, shape, mu0):
    (mx, my, mz), (nx, ny, nz) = mode, shape
    a = [_axis_a(mx, nx, False), _axis_a(my, ny, False), _axis_a(mz, nz, False)]
    ap_ = [_axis_a(mx, nx, True), _axis_a(my, ny, True), _axis_a(mz, nz, True)]
    s1 = sum(ap_)
    k2 = -sum(t * t for t in a)
    k3 = -sum(q * t * t for q, t in zip(ap_, a))
    k4 = sum(t ** 4 for t in a)
    D, dt, kc, apf = p.D_L, p.t_step, p.grad_coeff_c, p.ap
    num = 1 - dt * mu0 * s1 * (2 * apf * s1 - kc * k3) + 2 * apf * D * k2 * dt
    return num / (1 + D * dt * kc * k4)


def ch_mode_case(p, mode, shape, nsteps, c0=0.45, eps=0.01, a0=0.6):
    """-> (phi0, con0, phi_expected, con_expected) after nsteps; arrays (nz,ny,nx[,1])."""
    th = theta(mode, shape)
    nx, ny, nz = shape
    con = c0 + eps * np.cos(th)
    phi = np.full((nz, ny, nx, 1), a0)
    amp, f = eps, a0
    for _ in range(nsteps):
        mu0 = 2 * (1 - f * f) * p.D_gb
        amp *= ch_mode_factor(p, mode, shape, mu0)
        f = f - p.mob_phi * p.t_step * (f ** 3 - f)
    return phi, con, np.full_like(phi, f), c0 + amp * np.cos(th), amp


def ac_harmonics_case(p, mode, shape, a0=0.3, eps=0.2, c=0.5):
    """-> (phi0, con0, phi_expected) after one step, var = 1."""
    th = theta(mode, shape)
    phi = (a0 + eps * np.cos(th))[..., None].copy()
    con = np.full(th.shape, c)
    L, dt, kap = p.mob_phi, p.t_step, p.grad_coeff_phi
    psi = [a0 ** 3 + 1.5 * a0 * eps ** 2 - a0, 3 * a0 ** 2 * eps + 0.75 * eps ** 3 - eps,
           1.5 * a0 * eps ** 2, 0.25 * eps ** 3]
    coef = [a0, eps, 0.0, 0.0]
    expect = np.zeros_like(th)
    for n in range(4):
        k2 = sum(_axis_a(n * mm, N, False) ** 2 for mm, N in zip(mode, shape))
        expect += (coef[n] - L * dt * psi[n]) / (1 + L * kap * dt * k2) * np.cos(n * th)
    return phi, con, expect[..., None]


def single_mode_fint_case(p, mode, shape, eps_c=0.01, eps_p=0.02):
    """-> (phi0 (var 2), con0, f_int expected at the start of step 1)."""
    th = theta(mode, shape)
    con = 0.4 + eps_c * np.cos(th)
    phi = np.stack([0.5 + eps_p * np.cos(th), np.full_like(th, 0.5)], axis=-1)
    kp2 = sum(_axis_a(mm, N, True) ** 2 for mm, N in zip(mode, shape))
    P = th.size
    return phi, con, 0.5 * (p.grad_coeff_phi * eps_p ** 2 + p.grad_coeff_c * eps_c ** 2) * kp2 * P / 2
