"""The FP64 oracle's full operator against the exact compressible Navier-Stokes right-hand side.

For a smooth periodic state given by a few Fourier modes, -div(F_Euler - F_viscous) is evaluated in closed form
(product and quotient rules on the analytic first and second derivatives).  The DG operator (BR1 gradients,
viscous and Euler volume terms, face fluxes, split or weak form) must approach it at the nodes with
high-order accuracy on a curved mesh with rotated element frames.  Planted mistakes in the oracle's viscous
part (viscosity or Prandtl number off by 2 %, the viscous term dropped) must fail the same bound, so the
viscous flux, the heat flux and the lifted gradients are pinned, not only the inviscid part."""
import numpy as np
import pytest

import oracle as O
from galaexi_amd import mesh as M
from galaexi_amd import testcases as T

KAPPA, R, PR, MU = 1.4, 1.0, 0.72, 0.05


class Fourier:
    """f(x) = c + sum_m a_m cos(k_m . x + phi_m) with value, gradient and Hessian at points x(3, ...)."""

    def __init__(self, c, rng, amp, modes=4, kmax=2):
        self.c = c
        ks = [np.array(k) for k in np.ndindex(2 * kmax + 1, 2 * kmax + 1, 2 * kmax + 1)]
        ks = [k - kmax for k in ks if 0 < np.sum((k - kmax) ** 2) <= kmax ** 2]
        idx = rng.choice(len(ks), size=modes, replace=False)
        self.k = [ks[i].astype(np.float64) for i in idx]
        a = rng.standard_normal(modes)
        self.a = amp * a / np.sum(np.abs(a))
        self.ph = rng.uniform(0, 2 * np.pi, modes)

    def eval(self, x):
        f = np.full(x.shape[1:], self.c)
        g = np.zeros(x.shape)
        H = np.zeros((3, 3) + x.shape[1:])
        for a, k, ph in zip(self.a, self.k, self.ph):
            arg = k[0] * x[0] + k[1] * x[1] + k[2] * x[2] + ph
            c, s = np.cos(arg), np.sin(arg)
            f = f + a * c
            for i in range(3):
                g[i] -= a * k[i] * s
                for j in range(3):
                    H[i, j] -= a * k[i] * k[j] * c
        return f, g, H


def exact_rhs(x, fields, mu=MU, kappa=KAPPA, Rg=R, Pr=PR):
    """Ut = -div(F_E) + div(F_V) for the state (rho, u, v, w, p) = fields evaluated at x."""
    (rho, grho, Hrho), *vel, (p, gp, Hp) = [f.eval(x) for f in fields]
    u = [v_[0] for v_ in vel]
    gu = [v_[1] for v_ in vel]  # gu[i][j] = d u_i / d x_j
    Hu = [v_[2] for v_ in vel]
    lam = mu * Rg * kappa / ((kappa - 1.0) * Pr)
    divu = gu[0][0] + gu[1][1] + gu[2][2]
    q2 = u[0] ** 2 + u[1] ** 2 + u[2] ** 2
    E = p / (kappa - 1.0) + 0.5 * rho * q2
    gE = [gp[j] / (kappa - 1.0) + 0.5 * q2 * grho[j] + rho * sum(u[k] * gu[k][j] for k in range(3)) for j in range(3)]
    Ut = np.zeros((5,) + rho.shape)
    # Euler
    Ut[0] = -sum(u[j] * grho[j] + rho * gu[j][j] for j in range(3))
    for i in range(3):
        Ut[1 + i] = -(sum(u[i] * u[j] * grho[j] + rho * u[j] * gu[i][j] for j in range(3)) + rho * u[i] * divu + gp[i])
    Ut[4] = -((E + p) * divu + sum(u[j] * (gE[j] + gp[j]) for j in range(3)))
    # viscous: tau_ij = mu (du_i/dx_j + du_j/dx_i - 2/3 delta_ij div u)
    tau = [[mu * (gu[i][j] + gu[j][i] - (2.0 / 3.0) * (divu if i == j else 0.0)) for j in range(3)] for i in range(3)]
    lap = [Hu[i][0, 0] + Hu[i][1, 1] + Hu[i][2, 2] for i in range(3)]
    gdiv = [Hu[0][0, i] + Hu[1][1, i] + Hu[2][2, i] for i in range(3)]
    dtau = [mu * (lap[i] + gdiv[i] / 3.0) for i in range(3)]  # d tau_ij / dx_j
    for i in range(3):
        Ut[1 + i] += dtau[i]
    # T = p / (R rho): Laplacian by the quotient rule
    lapT = 0.0
    for j in range(3):
        lapT = lapT + (Hp[j, j] / rho - 2.0 * gp[j] * grho[j] / rho ** 2 - p * Hrho[j, j] / rho ** 2
                       + 2.0 * p * grho[j] ** 2 / rho ** 3) / Rg
    Ut[4] += sum(u[i] * dtau[i] for i in range(3)) + sum(tau[i][j] * gu[i][j] for i in range(3) for j in range(3)) \
        + lam * lapT
    return Ut


def state(seed):
    rng = np.random.default_rng(seed)
    return [Fourier(1.0, rng, 0.1), Fourier(0.0, rng, 0.3), Fourier(0.0, rng, 0.3), Fourier(0.0, rng, 0.3),
            Fourier(1.0 / KAPPA, rng, 0.1 / KAPPA)]


def dg_error(ne, N, split, riemann, seed=3, mu=MU, Pr=PR):
    m = M.box_mesh((ne, ne, ne), N, lo=(0, 0, 0), hi=(2 * np.pi,) * 3, deform_amp=0.04,
                   deform_phase=(0.2, 0.5, 0.7), rotations="random", seed=seed)
    f = state(seed)
    x = m.Elem_xGP
    vals = [fi.eval(x)[0] for fi in f]
    U = T.prim_to_cons(vals[0], vals[1], vals[2], vals[3], vals[4], KAPPA)
    Ut = O.Operator(m, split=split, riemann=riemann, kappa=KAPPA, R=R, Pr=Pr, mu0=mu).rhs(U)
    ex = exact_rhs(x, f)
    return np.abs(Ut - ex).max() / np.abs(ex).max(), np.abs(ex).max()


@pytest.mark.parametrize("split,riemann", [(True, O.RIEMANN_LF), (False, O.RIEMANN_ROE)])
def test_operator_converges_to_exact_navier_stokes(split, riemann):
    """Curved, rotated 3^3 and 6^3 meshes at N=7: error <= 2e-3 on the coarse mesh, and h-refinement gains at
    least 2^(0.8 N) (derivative of a degree-N interpolant: order N)."""
    e1, _ = dg_error(3, 7, split, riemann)
    e2, _ = dg_error(6, 7, split, riemann)
    assert e1 <= 2e-3, e1
    assert np.log2(e1 / e2) >= 0.8 * 7, (e1, e2)


def test_viscous_part_is_pinned():
    """Planted mistakes in the viscous part of the oracle (6^3, N=7) are at least 10x the error of the correct
    operator, which stays below 3e-5."""
    N, ne = 7, 6
    e_ok, _ = dg_error(ne, N, True, O.RIEMANN_LF)
    e_mu, _ = dg_error(ne, N, True, O.RIEMANN_LF, mu=MU * 1.02)
    e_pr, _ = dg_error(ne, N, True, O.RIEMANN_LF, Pr=PR * 1.02)
    e_inv, _ = dg_error(ne, N, True, O.RIEMANN_LF, mu=0.0)
    assert e_ok < 3e-5, e_ok
    for e in (e_mu, e_pr, e_inv):
        assert e > 10 * e_ok, (e_ok, e_mu, e_pr, e_inv)
