"""Initial states on GL nodes, in the reference layout U(5,0:N,0:N,0:N,nElems) (Fortran order).

Shu vortex: ExactFunc CASE(7) (src/equations/navierstokes/idealgas/exactfunc.f90:482-503).
Taylor-Green vortex: ExactFuncTestcase (src/testcase/taylorgreenvortex/testcase.f90:220-262), p0 from
src/testcase/taylorgreenvortex/testcase.f90:140-148.
PrimToCons: src/equations/navierstokes/idealgas/eos.f90 (E = p/(kappa-1) + rho|u|^2/2).
Synthetic waves with a manufactured source: ExactFunc CASE(4), CASE(41), CASE(42), CASE(43)
(src/equations/navierstokes/idealgas/exactfunc.f90:326-420).
"""
from __future__ import annotations

import numpy as np


def prim_to_cons(rho, u, v, w, p, kappa):
    U = np.empty((5,) + np.shape(rho), order="F")
    U[0] = rho
    U[1] = rho * u
    U[2] = rho * v
    U[3] = rho * w
    U[4] = p / (kappa - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
    return np.asfortranarray(U)


def freestream(x, rho=1.0, vel=(0.3, -0.2, 0.1), p=1.0, kappa=1.4):
    s = x[0].shape
    return prim_to_cons(np.full(s, rho), np.full(s, vel[0]), np.full(s, vel[1]), np.full(s, vel[2]),
                        np.full(s, p), kappa)


def shu_vortex(x, t=0.0, center=(0.0, 0.0, 0.0), amplitude=0.2, halfwidth=1.0, axis=(0.0, 0.0, 1.0),
               rho0=1.0, vel0=(0.5, 0.0, 0.0), p0=None, kappa=1.4):
    """Isentropic vortex (exactfunc.f90:482-503) around the given axis."""
    if p0 is None:
        p0 = 1.0 / kappa
    vel0 = np.asarray(vel0, dtype=np.float64)
    RT = p0 / rho0
    cent = np.asarray(center, dtype=np.float64) + vel0 * t
    d = x - cent.reshape((3,) + (1,) * (x.ndim - 1))
    ax = np.asarray(axis, dtype=np.float64).reshape((3,) + (1,) * (x.ndim - 1))
    c = np.cross(ax, d, axis=0) / halfwidth
    r2 = np.sum(c * c, axis=0)
    du = amplitude / (2.0 * np.pi) * np.exp(0.5 * (1.0 - r2))
    dT = -(kappa - 1.0) / (2.0 * kappa * RT) * du ** 2
    rho = rho0 * (1.0 + dT) ** (1.0 / (kappa - 1.0))
    u = vel0[0] + du * c[0]
    v = vel0[1] + du * c[1]
    w = vel0[2] + du * c[2]
    p = p0 * (1.0 + dT) ** (kappa / (kappa - 1.0))
    return prim_to_cons(rho, u, v, w, p, kappa)


def taylor_green(x, Ma=0.1, rho0=1.0, U0=1.0, kappa=1.4, noise=0.0, seed=1600):
    """TGV initial condition, IniConstDens=T, plus optional seeded N(0,1)*noise velocity perturbation."""
    p0 = (U0 / Ma) ** 2 / kappa * rho0
    X, Y, Z = x[0], x[1], x[2]
    u = U0 * np.sin(X) * np.cos(Y) * np.cos(Z)
    v = -U0 * np.cos(X) * np.sin(Y) * np.cos(Z)
    w = np.zeros_like(X)
    p = p0 + (rho0 * U0 ** 2) / 16.0 * (np.cos(2 * X) * np.cos(2.0 * Z) + 2.0 * np.cos(2.0 * Y) +
                                        2.0 * np.cos(2.0 * X) + np.cos(2 * Y) * np.cos(2.0 * Z))
    if noise:
        rng = np.random.default_rng(seed)
        u = u + noise * rng.standard_normal(u.shape)
        v = v + noise * rng.standard_normal(v.shape)
        w = w + noise * rng.standard_normal(w.shape)
    rho = np.full_like(X, rho0)
    return prim_to_cons(rho, u, v, w, p, kappa)


def smooth_random(x, lo, hi, seed=0, kappa=1.4, kmax=2):
    """Seeded smooth periodic state: rho=1+0.1 s1, u_d=0.3 s_{d+1}, p=(1+0.1 s5)/kappa, where s_i are sums
    of low Fourier modes (|k|<=kmax) on the box, normalised to max|s_i|=1 over the nodes."""
    rng = np.random.default_rng(seed)
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    xn = 2.0 * np.pi * (x - lo.reshape((3,) + (1,) * (x.ndim - 1))) / (hi - lo).reshape((3,) + (1,) * (x.ndim - 1))
    ks = [(a, b, c) for a in range(-kmax, kmax + 1) for b in range(-kmax, kmax + 1) for c in range(-kmax, kmax + 1)
          if 0 < a * a + b * b + c * c <= kmax * kmax]
    fields = []
    for _ in range(5):
        s = np.zeros(x.shape[1:])
        for (a, b, c) in ks:
            amp, ph = rng.standard_normal(), rng.uniform(0, 2 * np.pi)
            s += amp * np.cos(a * xn[0] + b * xn[1] + c * xn[2] + ph)
        fields.append(s / np.max(np.abs(s)))
    rho = 1.0 + 0.1 * fields[0]
    return prim_to_cons(rho, 0.3 * fields[1], 0.3 * fields[2], 0.3 * fields[3], (1.0 + 0.1 * fields[4]) / kappa,
                        kappa)


SOURCE_FUNCS = (4, 41, 42, 43)


def synthetic_wave(case, x, t=0.0, AdvVel=(0.3, 0.3, 0.3)):
    """ExactFunc CASE(4) / CASE(41..43) (exactfunc.f90:326-420), 3D, Frequency 1, Amplitude 0.1.
    rho = 2 + 0.1 sin(pi*s - a*t) with s = x+y+z (4) or x, y, z (41, 42, 43) and a = 2*pi*AdvVel(d); the momentum
    equals rho in every direction of the wave (velocity 1 there, 0 elsewhere) and E = rho^2.
    x(3, ...) -> U(5, ...)."""
    if case not in SOURCE_FUNCS:
        raise ValueError(f"exact function {case} is not one of the synthetic waves {SOURCE_FUNCS}")
    omega, amp = np.pi * 1.0, 0.1
    d = 0 if case == 4 else case - 41
    a = AdvVel[d] * 2.0 * np.pi
    s = (x[0] + x[1]) + x[2] if case == 4 else x[d]
    rho = 2.0 + amp * np.sin(omega * s - a * t)
    U = np.zeros((5,) + np.shape(rho), order="F")
    U[0] = rho
    if case == 4:
        U[1] = U[2] = U[3] = rho
    else:
        U[1 + d] = rho
    U[4] = rho * rho
    return np.asfortranarray(U)

