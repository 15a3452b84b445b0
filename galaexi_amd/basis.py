"""1D nodal bases used by the host-side geometry producer.

Restates the reference's basis module (src/interpolation/basis.f90:100-807,
src/interpolation/interpolation.f90:199-395): Legendre/Gauss/Gauss-Lobatto/Chebyshev-Lobatto/
equidistant node sets, barycentric weights, Lagrange interpolation matrices, the polynomial
derivative matrix and the modal (Legendre) truncation Vandermonde.
"""
from __future__ import annotations

import numpy as np

GAUSS = "GAUSS"
GAUSS_LOBATTO = "GAUSS-LOBATTO"
CHEBYSHEV_LOBATTO = "CHEBYSHEV-GAUSS-LOBATTO"
VISU = "VISU"


def legendre(n: int, x):
    """L2-normalised Legendre polynomial sqrt(n+1/2)*L_n(x) and its derivative (basis.f90:195-234)."""
    x = np.asarray(x, dtype=np.float64)
    s = np.sqrt(n + 0.5)
    if n == 0:
        return s * np.ones_like(x), np.zeros_like(x)
    if n == 1:
        return s * x, s * np.ones_like(x)
    lm2, lm1, ldm1 = np.ones_like(x), x.copy(), np.ones_like(x)
    for i in range(2, n + 1):
        l = ((2 * i - 1) * x * lm1 - (i - 1) * lm2) / i
        ld = i * lm1 + x * ldm1
        lm2, lm1, ldm1 = lm1, l, ld
    return s * l, s * ld


def _q_and_l(n, x):
    """q = L_{n+1}-L_{n-1}, q' and L_n (basis.f90 qAndLEvaluation)."""
    lm2, lm1, l = 1.0, x, x
    for i in range(2, n + 1):
        l = ((2 * i - 1) * x * lm1 - (i - 1) * lm2) / i
        lm2, lm1 = lm1, l
    q = (2 * n + 1) / (n + 1) * (x * l - lm2)
    qd = (2 * n + 1) * l
    return q, qd, l


def nodes_and_weights(N: int, node_type: str = GAUSS_LOBATTO):
    """Nodes, quadrature weights and barycentric weights (interpolation.f90:199-318)."""
    x = np.zeros(N + 1)
    w = np.zeros(N + 1)
    if node_type == GAUSS_LOBATTO:
        x[0], x[N] = -1.0, 1.0
        w[0] = w[N] = 2.0 / (N * (N + 1))
        if N > 1:
            c1, c2 = np.pi / N, 3.0 / (8 * N * np.pi)
            for i in range(1, (N + 1) // 2):
                xi = -np.cos(c1 * (i + 0.25) - c2 / (i + 0.25))
                for _ in range(11):
                    q, qd, _l = _q_and_l(N, xi)
                    dx = -q / qd
                    xi += dx
                    if abs(dx) < 1e-15 * abs(xi):
                        break
                q, qd, l = _q_and_l(N, xi)
                x[i], x[N - i] = xi, -xi
                w[i] = w[N - i] = w[0] / (l * l)
        if N % 2 == 0:
            _q, _qd, l = _q_and_l(N, 0.0)
            x[N // 2] = 0.0
            w[N // 2] = w[0] / (l * l)
    elif node_type == GAUSS:
        if N == 0:
            x[0], w[0] = 0.0, 2.0
        elif N == 1:
            x[:] = [-np.sqrt(1 / 3), np.sqrt(1 / 3)]
            w[:] = 1.0
        else:
            ct = 2.0 * np.arctan(1.0) / (N + 1)
            for i in range((N + 1) // 2):
                xi = -np.cos(ct * (2 * i + 1))
                for _ in range(11):
                    l, ld = legendre(N + 1, xi)
                    dx = -l / ld
                    xi += dx
                    if abs(dx) < 1e-15 * abs(xi):
                        break
                l, ld = legendre(N + 1, xi)
                x[i], x[N - i] = xi, -xi
                w[i] = w[N - i] = (2 * N + 3) / ((1 - xi * xi) * ld * ld)
            if N % 2 == 0:
                l, ld = legendre(N + 1, 0.0)
                x[N // 2] = 0.0
                w[N // 2] = (2 * N + 3) / (ld * ld)
    elif node_type == CHEBYSHEV_LOBATTO:
        x[:] = -np.cos(np.arange(N + 1) / N * np.pi)
        w[:] = np.pi / N
        w[0] *= 0.5
        w[N] *= 0.5
    elif node_type == VISU:
        x[:] = 2.0 * np.arange(N + 1) / N - 1.0
        w[:] = 2.0 / N
        w[0] *= 0.5
        w[N] *= 0.5
    else:
        raise ValueError(f"unknown node type {node_type}")
    return x, w, barycentric_weights(x)


def barycentric_weights(x):
    """basis.f90 BarycentricWeights."""
    n = len(x)
    wb = np.ones(n)
    for i in range(1, n):
        for j in range(i):
            wb[j] *= x[j] - x[i]
            wb[i] *= x[i] - x[j]
    return 1.0 / wb


def lagrange_polys(xe: float, x, wb):
    """Lagrange basis at xe (basis.f90 LagrangeInterpolationPolys)."""
    eps = np.finfo(np.float64).eps
    L = np.zeros(len(x))
    hit = False
    for i, xi in enumerate(x):
        if (xe == 0.0 or xi == 0.0):
            eq = abs(xe - xi) <= 2 * eps
        else:
            eq = abs(xe - xi) <= eps * abs(xe) and abs(xe - xi) <= eps * abs(xi)
        if eq:
            L[i] = 1.0
            hit = True
    if hit:
        return L
    L = wb / (xe - x)
    return L / L.sum()


def vandermonde(N_in, type_in, N_out, type_out, modal=False):
    """Interpolation (or modal truncation) matrix V(0:N_out,0:N_in) (interpolation.f90:324-395)."""
    if type_in == type_out and N_in == N_out:
        return np.eye(N_in + 1)
    x_in, _w, wb_in = nodes_and_weights(N_in, type_in)
    x_out, _w2, _wb = nodes_and_weights(N_out, type_out)
    if modal and N_out < N_in:
        Vleg_in = np.stack([legendre(j, x_in)[0] for j in range(N_in + 1)], axis=1)
        Vleg_out = np.stack([legendre(j, x_out)[0] for j in range(N_out + 1)], axis=1)
        sV_in = np.linalg.inv(Vleg_in)
        return Vleg_out @ sV_in[: N_out + 1, :]
    return np.stack([lagrange_polys(xo, x_in, wb_in) for xo in x_out], axis=0)


def derivative_matrix(x):
    """D(i,l) = l_l'(x_i) (basis.f90:664-690)."""
    n = len(x)
    wb = barycentric_weights(x)
    D = np.zeros((n, n))
    for l in range(n):
        for i in range(n):
            if l != i:
                D[i, l] = wb[l] / (wb[i] * (x[i] - x[l]))
                D[i, i] -= D[i, l]
    return D
