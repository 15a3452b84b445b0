"""The FP64 oracle against properties the mathematics fixes: discrete conservation, element-frame
equivariance, isentropic-vortex convergence order (regressioncheck/checks/convtest rule: rate >= 0.8 of
N+1), fourth-order LSERK coefficients, the closed-form time step and the errType path."""
import numpy as np
import pytest

import oracle as O
from galaexi_amd import mesh as M
from galaexi_amd import testcases as T


def vol_weights(N):
    w = O.dg_basis(N)["wGP"]
    return w[:, None, None] * w[None, :, None] * w[None, None, :]


FORMS = [(True, O.RIEMANN_LF), (True, O.RIEMANN_ROEEF), (False, O.RIEMANN_ROE), (False, O.RIEMANN_LF)]


@pytest.mark.parametrize("split,riemann", FORMS)
def test_conservation_curved(split, riemann):
    """Periodic mesh: sum_nodes w J Ut = 0 for every conservative variable."""
    m = M.box_mesh((3, 2, 2), 3, lo=(-1, -1, -1), hi=(1, 1, 1), deform_amp=0.05, deform_phase=(0.2, 0.4, 0.6),
                   rotations="random", seed=2)
    U = T.smooth_random(m.Elem_xGP, (-1, -1, -1), (1, 1, 1), seed=3)
    Ut = O.Operator(m, split=split, riemann=riemann, mu0=2e-2, R=1.0).rhs(U)
    W = vol_weights(3)
    tot = np.einsum("ijk,vijke,ijke->v", W, Ut, 1.0 / m.sJ)
    scale = np.einsum("ijk,vijke,ijke->v", W, np.abs(Ut), 1.0 / m.sJ)
    assert np.all(np.abs(tot) <= 1e-13 * scale)


@pytest.mark.parametrize("split,riemann", FORMS[:3])
def test_element_frame_equivariance(split, riemann):
    """Relabelling element frames (all 24 rotations, every flip) maps Ut node-for-node."""
    N = 3
    m0 = M.box_mesh((3, 3, 2), N, lo=(-2, -2, -2), hi=(2, 2, 2))
    m1 = M.box_mesh((3, 3, 2), N, lo=(-2, -2, -2), hi=(2, 2, 2), rotations="random", seed=7)
    kw = dict(center=(0.3, -0.2, 0.1), amplitude=0.8, axis=(0.3, 0.5, 0.8))
    Ut0 = O.Operator(m0, split=split, riemann=riemann, mu0=2e-2, R=1.0).rhs(T.shu_vortex(m0.Elem_xGP, **kw))
    Ut1 = O.Operator(m1, split=split, riemann=riemann, mu0=2e-2, R=1.0).rhs(T.shu_vortex(m1.Elem_xGP, **kw))
    n = N + 1
    c0 = m0.Elem_xGP.reshape(3, n ** 3, -1, order="F").mean(axis=1).T
    c1 = m1.Elem_xGP.reshape(3, n ** 3, -1, order="F").mean(axis=1).T
    err = 0.0
    for e1 in range(m1.nElems):
        e0 = int(np.argmin(np.sum((c0 - c1[e1]) ** 2, axis=1)))
        x0 = m0.Elem_xGP[..., e0].reshape(3, -1, order="F").T
        x1 = m1.Elem_xGP[..., e1].reshape(3, -1, order="F").T
        perm = [int(np.argmin(np.sum((x0 - p) ** 2, axis=1))) for p in x1]
        u0 = Ut0[..., e0].reshape(5, -1, order="F")[:, perm]
        u1 = Ut1[..., e1].reshape(5, -1, order="F")
        err = max(err, np.abs(u0 - u1).max())
    assert err < 1e-12 * np.abs(Ut0).max()


@pytest.mark.parametrize("split,riemann", [(False, O.RIEMANN_ROE), (True, O.RIEMANN_LF)])
def test_isentropic_vortex_convergence(split, riemann):
    """h-convergence of the Shu vortex (exactfunc.f90:482-503) at N=3, inviscid: rate >= 0.8*(N+1)."""
    N, tend = 3, 0.5
    errs = []
    for ne in (4, 8):
        m = M.box_mesh((ne, ne, 1), N, lo=(-5, -5, -5), hi=(5, 5, 5))
        kw = dict(center=(0.3, -0.2, 0.0), amplitude=0.8, axis=(0, 0, 1))
        U = np.asfortranarray(T.shu_vortex(m.Elem_xGP, **kw))
        op = O.Operator(m, split=split, riemann=riemann, kappa=1.4, R=1.0, Pr=0.72, mu0=0.0)
        dt = op.calc_dt(U, 0.5, 0.5)[0]
        nst = int(np.ceil(tend / dt))
        dt = tend / nst
        K = np.zeros_like(U, order="F")
        for s in range(nst):
            op.lserk_step(U, K, s * dt, dt)
        Ue = T.shu_vortex(m.Elem_xGP, t=tend, **kw)
        W, J = vol_weights(N), 1.0 / m.sJ
        errs.append(np.sqrt(np.einsum("ijk,ijke,ijke->", W, (U[0] - Ue[0]) ** 2, J) / np.einsum("ijk,ijke->", W, J)))
    rate = np.log2(errs[0] / errs[1])
    assert rate >= 0.8 * (N + 1), (errs, rate)


def test_lserk_coefficients_fourth_order():
    """The 2N-storage Carpenter-Kennedy RK4-5 coefficients (timedisc_vars.f90:213-227) integrate y'=lambda y
    with order 4, applied with the register algebra of VAXPB_OUT_VAXPB_IN (vector.f90:180-181)."""
    A, b, c = O.lserk_coefs()
    lam = -1.0 + 2.0j

    def run(dt, T_end=1.0):
        y = 1.0 + 0j
        k = 0.0 + 0j
        for _ in range(int(round(T_end / dt))):
            for s in range(5):
                k = k * (-A[s] if s else 0.0) + lam * y
                y = y + b[s] * dt * k
        return y

    e = [abs(run(dt) - np.exp(lam)) for dt in (0.1, 0.05)]
    assert np.log2(e[0] / e[1]) > 3.8
    # the nodes c_s are the accumulated b of the previous stages
    cum, kk = [0.0], 0.0
    for s in range(1, 5):
        kk = kk * (-A[s - 1] if s > 1 else 0.0) + 1.0
        cum.append(cum[-1] + b[s - 1] * kk)
    assert np.allclose(c[1:], cum[1:], rtol=1e-12)


def test_calc_dt_closed_form():
    """Cartesian box, uniform state: dt = min(CFL*alpha_N/(2N+1) * 2/lambda_c, DFL*alpha_v*1.853/(2N+1)^2 *
    4/lambda_v) with lambda_c = sum_d (|u_d| + c) 2/h_d and lambda_v = sum_d mu/rho max(4/3,kappa/Pr) (2/h_d)^2."""
    N, ne, L = 5, 4, 2.0
    m = M.box_mesh((ne, ne, ne), N, lo=(0, 0, 0), hi=(L, L, L))
    vel, rho, p, kappa, Pr, mu = (0.3, -0.2, 0.1), 1.2, 0.9, 1.4, 0.72, 0.05
    U = T.freestream(m.Elem_xGP, rho=rho, vel=vel, p=p, kappa=kappa)
    op = O.Operator(m, split=True, riemann=O.RIEMANN_LF, kappa=kappa, R=1.0, Pr=Pr, mu0=mu)
    dt, err, dtc, dtv = op.calc_dt(U, 0.9, 0.8)
    h = L / ne
    c = np.sqrt(kappa * p / rho)
    lc = sum((abs(v) + c) * 2 / h for v in vel)
    lv = 3 * mu / rho * max(4 / 3, kappa / Pr) * (2 / h) ** 2
    dtc_ref = 0.9 * 2.2027 / (2 * N + 1) * 2 / lc
    dtv_ref = 0.8 * 0.56 * 1.853 / (2 * N + 1) ** 2 * 4 / lv
    assert err == 0
    assert abs(dtc - dtc_ref) < 1e-12 * dtc_ref and abs(dtv - dtv_ref) < 1e-12 * dtv_ref
    assert dt == min(dtc, dtv)


def test_calc_dt_err_type():
    m = M.box_mesh((2, 2, 2), 3)
    U = T.freestream(m.Elem_xGP).copy(order="F")
    U[0, 0, 0, 0, 1] = -0.5
    _dt, err, _c, _v = O.Operator(m).calc_dt(U)
    assert err == 1


def _normal_shock(M1=2.0, rho1=1.0, p1=1.0, kappa=1.4, vt=(0.3, -0.2)):
    """Stationary normal shock in +x (Rankine-Hugoniot), tangential velocity continuous."""
    c1 = np.sqrt(kappa * p1 / rho1)
    u1 = M1 * c1
    rho2 = rho1 * (kappa + 1) * M1 ** 2 / ((kappa - 1) * M1 ** 2 + 2)
    p2 = p1 * (1 + 2 * kappa / (kappa + 1) * (M1 ** 2 - 1))
    u2 = u1 * rho1 / rho2
    return (rho1, u1, vt[0], vt[1], p1), (rho2, u2, vt[0], vt[1], p2)


def _euler_flux_n(prim, nv, kappa=1.4):
    rho, u, v, w, p = prim
    vel = np.array([u, v, w])
    vn = vel @ nv
    E = p / (kappa - 1) + 0.5 * rho * vel @ vel
    return np.array([rho * vn, *(rho * vn * vel + p * nv), vn * (E + p)])


@pytest.mark.parametrize("riemann", [O.RIEMANN_ROE, O.RIEMANN_ROEEF, O.RIEMANN_LF])
def test_riemann_flux_at_stationary_shock(riemann):
    """Elements with x-index 0 hold the upstream state of a Mach-2 stationary shock, x-index 1 the downstream
    state (inviscid, weak form).  Faces between equal states give F(U).n (consistency, all solvers).  Roe's
    linearisation resolves any Rankine-Hugoniot pair exactly, so both x-face families give F(U_up).n; the
    entropy fix stays inactive at the admissible (upstream-master) faces, so RoeEntropyFix is exact there; the
    Lax-Friedrichs flux carries its dissipation -lmax/2 (U_R - U_L) with lmax = max|u_n| + max c."""
    up, dn = _normal_shock()
    m = M.box_mesh((2, 2, 2), 2, lo=(0, 0, 0), hi=(1, 1, 1))
    xi = (m.Elem_xGP[0].reshape(-1, m.nElems, order="F").mean(axis=0) > 0.5).astype(int)
    prims = [np.where(xi[None, None, None, :] == 0, a, b) * np.ones(m.sJ.shape) for a, b in zip(up, dn)]
    U = T.prim_to_cons(*prims, 1.4)
    op = O.Operator(m, split=False, riemann=riemann, mu0=0.0, R=1.0)
    op.rhs(U)
    Fm = op.array(8).reshape((5, 3, 3, m.nSides), order="F")
    checked = {"same": 0, "up": 0, "down": 0}
    for s in range(m.nSides):
        e_m, e_s = m.SideToElem[0, s] - 1, m.SideToElem[1, s] - 1
        nv = m.NormVec[:, 0, 0, s]
        se = m.SurfElem[0, 0, s]
        F = Fm[:, 1, 1, s] / se
        if xi[e_m] == xi[e_s]:
            exact = _euler_flux_n(up if xi[e_m] == 0 else dn, nv)
            assert np.allclose(F, exact, rtol=1e-13, atol=1e-13)
            checked["same"] += 1
            continue
        admissible = xi[e_m] == 0 and nv[0] > 0.5
        exact = _euler_flux_n(up, nv)
        assert np.allclose(exact, _euler_flux_n(dn, nv), rtol=1e-13)
        if riemann == O.RIEMANN_ROE or (riemann == O.RIEMANN_ROEEF and admissible):
            assert np.allclose(F, exact, rtol=1e-12, atol=1e-12), (riemann, s, F, exact)
        elif riemann == O.RIEMANN_LF:
            L, Rr = (up, dn) if xi[e_m] == 0 else (dn, up)
            vnL, vnR = np.array(L[1:4]) @ nv, np.array(Rr[1:4]) @ nv
            lmax = max(abs(vnL), abs(vnR)) + max(np.sqrt(1.4 * L[4] / L[0]), np.sqrt(1.4 * Rr[4] / Rr[0]))
            UL = T.prim_to_cons(*[np.array([v]) for v in L], 1.4)[:, 0]
            UR = T.prim_to_cons(*[np.array([v]) for v in Rr], 1.4)[:, 0]
            assert np.allclose(F, exact - 0.5 * lmax * (UR - UL), rtol=1e-12, atol=1e-12)
        checked["up" if admissible else "down"] += 1
    assert checked["same"] > 0 and checked["up"] > 0 and checked["down"] > 0


@pytest.mark.parametrize("split", [False, True])
def test_lax_friedrichs_flux_jacobian_eigenvalues(split):
    """Riemann_LF (src/equations/navierstokes/riemann.f90:707-733) is the local Lax-Friedrichs (Rusanov) flux
    F*(U_L,U_R) = C(U_L,U_R) - lmax/2 (U_R - U_L), lmax = max(|u_n| + c) of the two sides, with C the central flux
    (weak form: the mean of the Euler fluxes; split form: the Pirozzoli flux, splitflux.f90:735-761).  Any
    consistent symmetric C has dC/dU_R = dC/dU_L = A/2 at U_L = U_R = U (A = dF(U).n/dU), and the jump term adds
    -+lmax/2 I there, so the Jacobians of the oracle's face flux at a uniform state have the eigenvalues
    (a_k -+ lmax)/2 with a_k = u_n - c, u_n (three times), u_n + c.  A wrong dissipation coefficient, sign or wave
    speed moves them.  The Jacobian comes from central differences of the whole operator's master-side flux on a
    two-element-wide box whose second element column carries the perturbed state."""
    kappa = 1.4
    m = M.box_mesh((2, 1, 1), 2, lo=(0, 0, 0), hi=(2, 1, 1))
    xi = (m.Elem_xGP[0].reshape(-1, m.nElems, order="F").mean(axis=0) > 1.0).astype(int)
    rho, vel, p = 1.1, np.array([0.35, -0.2, 0.15]), 0.8
    base = T.freestream(m.Elem_xGP, rho=rho, vel=tuple(vel), p=p, kappa=kappa)
    U0 = base[:, 0, 0, 0, 0].copy()
    op = O.Operator(m, split=split, riemann=O.RIEMANN_LF, kappa=kappa, mu0=0.0, R=1.0)
    # an x-face with master element in column 0, slave in column 1 (normal +x)
    sides = [s for s in range(m.nSides)
             if xi[m.SideToElem[0, s] - 1] == 0 and xi[m.SideToElem[1, s] - 1] == 1 and m.NormVec[0, 1, 1, s] > 0.5]
    assert sides
    s = sides[0]
    nv = m.NormVec[:, 1, 1, s]
    se = m.SurfElem[1, 1, s]

    def face_flux(dL, dR):
        U = base.copy(order="F")
        for e in range(m.nElems):
            U[:, :, :, :, e] += (dL if xi[e] == 0 else dR)[:, None, None, None]
        op.rhs(U)
        return op.array(8).reshape((5, 3, 3, m.nSides), order="F")[:, 1, 1, s] / se

    eps = 1e-6
    JL, JR = np.zeros((5, 5)), np.zeros((5, 5))
    z = np.zeros(5)
    for k in range(5):
        d = np.zeros(5)
        d[k] = eps * max(abs(U0[k]), 1.0)
        JR[:, k] = (face_flux(z, d) - face_flux(z, -d)) / (2 * d[k])
        JL[:, k] = (face_flux(d, z) - face_flux(-d, z)) / (2 * d[k])
    un = vel @ nv
    c = np.sqrt(kappa * p / rho)
    lmax = abs(un) + c
    a = np.array([un - c, un, un, un, un + c])
    eR = np.sort(np.linalg.eigvals(JR).real)
    eL = np.sort(np.linalg.eigvals(JL).real)
    assert np.allclose(eR, np.sort((a - lmax) / 2), atol=1e-6 * lmax), (eR, (a - lmax) / 2)
    assert np.allclose(eL, np.sort((a + lmax) / 2), atol=1e-6 * lmax), (eL, (a + lmax) / 2)
    # and their sum is the flux Jacobian A itself (the dissipation cancels), whose eigenvalues are a_k
    assert np.allclose(np.sort(np.linalg.eigvals(JL + JR).real), np.sort(a), atol=1e-6 * lmax)
