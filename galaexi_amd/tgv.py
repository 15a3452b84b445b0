"""Taylor-Green vortex testcase on the galx engine: the testcase analysis and the reference's time loop.

  TGVAnalysis.analyze   AnalyzeTestcase (src/testcase/taylorgreenvortex/testcase.f90:283-515): 15 integral
                        quantities from U and the lifted velocity gradients, interpolated to NAnalyze Gauss-Lobatto
                        points (InitAnalyzeBasis, src/analyze/analyze.f90:235-273; Vol from :158-167)
  run                   TimeDisc (src/timedisc/timedisc.f90:100-200) with InitTimeStep / UpdateTimeStep /
                        AnalyzeTimeStep (src/timedisc/timedisc_func.f90:205-400) at their defaults
                        (nCalcTimeStepMax=1, no dtmin): dt from CALCTIMESTEP every step, the testcase analysis every
                        nAnalyzeTestCase steps and at analyze times.

Between analyze times the reference evaluates the testcase analysis with the gradients that the last DG
operator call left behind, i.e. those of the input state of the fifth RK stage of the step just taken; at analyze
times it calls the operator on the new state first (timedisc_func.f90:348-349).  run() reproduces both.
"""
from __future__ import annotations

import numpy as np

from . import basis

COLUMNS = ("Time", "Dissipation Rate Incompressible", "Dissipation Rate Compressible", "Ekin incomp", "Ekin comp",
           "Enstrophy comp", "DR_u", "DR_S", "DR_Sd", "DR_p", "Maximum Vorticity", "Mean Temperature", "uprime",
           "Mean Entropy", "ED_S", "ED_D")


def _interp3(V, f):
    """f(..., i, j, k, e) -> (..., a, b, c, e) with V(a, i) along each of the three node axes."""
    return np.einsum("ai,bj,ck,...ijke->...abce", V, V, V, f, optimize=True)


class TGVAnalysis:
    def __init__(self, mesh, kappa=1.4, R=71.42857, mu0=6.25e-4, rho0=1.0, NAnalyze=10):
        N = mesh.N
        self.kappa, self.R, self.mu0, self.rho0 = kappa, R, mu0, rho0
        _x, wGP, _wb = basis.nodes_and_weights(N, basis.GAUSS_LOBATTO)
        _xa, wA, _wba = basis.nodes_and_weights(NAnalyze, basis.GAUSS_LOBATTO)
        self.V = basis.vandermonde(N, basis.GAUSS_LOBATTO, NAnalyze, basis.GAUSS_LOBATTO)
        Wv = wGP[:, None, None] * wGP[None, :, None] * wGP[None, None, :]
        self.Vol = float(np.sum(Wv[..., None] / mesh.sJ))
        WA = wA[:, None, None] * wA[None, :, None] * wA[None, None, :]
        self.IntFactor = WA[..., None] / _interp3(self.V, mesh.sJ)

    def analyze(self, U, gradUx, gradUy, gradUz):
        """U(5,0:N,0:N,0:N,nElems); gradU*(>=3, ...) with the velocity gradients in rows 0..2 (lifting order
        u,v,w[,T]).  Returns the 15 quantities in the order of the reference's output file."""
        k, R, mu, rho0, Vol, F = self.kappa, self.R, self.mu0, self.rho0, self.Vol, self.IntFactor
        UA = _interp3(self.V, U)
        G = np.stack([_interp3(self.V, g[0:3]) for g in (gradUx, gradUy, gradUz)], axis=1)  # G[i, j] = du_i/dx_j
        rho = UA[0]
        vel = UA[1:4] / rho
        p = (k - 1.0) * (UA[4] - 0.5 * np.sum(UA[1:4] * vel, axis=0))
        T = p / (rho * R)
        q2 = np.sum(vel * vel, axis=0)
        divU = G[0, 0] + G[1, 1] + G[2, 2]
        S = 0.5 * (G + np.swapaxes(G, 0, 1))
        Sd = S.copy()
        for i in range(3):
            Sd[i, i] = Sd[i, i] - 1.0 / 3.0 * divU
        vort = np.stack([G[2, 1] - G[1, 2], G[0, 2] - G[2, 0], G[1, 0] - G[0, 1]])
        vort2 = np.sum(vort * vort, axis=0)
        T_mean = np.sum(F * T) / Vol
        Entropy = np.sum(F * (-1.0 / (k - 1.0)) * rho * (np.log(p) - k * np.log(rho))) / Vol
        Ekin = np.sum(F * 0.5 * q2) / Vol
        Ekin_comp = np.sum(F * 0.5 * rho * q2) / Vol / rho0
        Enstr = np.sum(F * 0.5 * rho * vort2) / (rho0 * Vol)
        DR_u = np.sum(F * np.sum(G * G, axis=(0, 1))) * mu / Vol
        DR_S = np.sum(F * np.sum(S * S, axis=(0, 1))) * 2.0 * mu / (rho0 * Vol)
        DR_Sd = np.sum(F * np.sum(Sd * Sd, axis=(0, 1))) * 2.0 * mu / (rho0 * Vol)
        DR_p = -np.sum(F * p * divU) / (rho0 * Vol)
        ED_S = np.sum(F * mu * vort2) / Vol
        ED_D = np.sum(F * mu * divU ** 2) * 4.0 / 3.0 / Vol
        vmax = float(np.sqrt(vort2.max()))
        uPrime = np.sqrt(2.0 / 3.0 * Ekin)
        return np.array([DR_S, DR_Sd + DR_p, Ekin, Ekin_comp, Enstr, DR_u, DR_S, DR_Sd, DR_p, vmax, T_mean, uPrime,
                         Entropy, ED_S, ED_D])


def run(op, U0, tend, analysis, nAnalyzeTestCase=10, analyze_dt=None, max_iter=None):
    """TimeDisc on a DGOperator from U0 at t=0 to tend.  Returns an array of rows (Time + 15 quantities) at the
    times the reference writes them: t=0, every nAnalyzeTestCase steps and at analyze times."""
    analyze_dt = tend if analyze_dt is None else analyze_dt
    t = 0.0
    op.set_state(U0)
    rows = [np.concatenate([[t], analysis.analyze(U0, *op.get_gradients())])]
    tAnalyze = min(t + analyze_dt, tend)
    doAnalyze = doFinalize = False
    it = 0

    def timestep_limits(dt_cfl):
        nonlocal doAnalyze, doFinalize
        dmin = [dt_cfl, tAnalyze - t, tend - t]
        dt = min(dmin)
        if dt == dmin[1]:
            doAnalyze = True
        if dt == dmin[2]:
            doAnalyze = doFinalize = True
        dt = min(v for v in dmin if v > 0)
        return dt, dmin

    def calc_dt():
        dt_cfl, err = op.CalcTimeStep()
        if err != 0:
            raise RuntimeError(f"CALCTIMESTEP errType {err} at t={t}")
        return dt_cfl

    # InitTimeStep
    dt, _ = timestep_limits(calc_dt())
    while True:
        # UpdateTimeStep (nCalcTimeStepMax = 1: dt from CALCTIMESTEP every step)
        dt, dmin = timestep_limits(calc_dt())
        if dmin[1] - dt < dt / 100.0 and dmin[1] > 0:
            dt, doAnalyze = dmin[1], True
        if dmin[2] - dt < dt / 100.0 and dmin[2] > 0:
            dt, doAnalyze, doFinalize = dmin[2], True, True
        testcase_now = (it + 1) % nAnalyzeTestCase == 0 or doAnalyze
        grads = None
        for stage in range(1, 6):
            if stage == 5 and testcase_now and not doAnalyze:
                grads = op.get_gradients()  # the operator's last lifting of this step: stage-5 input state
            op.lserk_stage(t, dt, stage)
        it += 1
        t = t + dt
        if doAnalyze:
            grads = op.get_gradients()  # DGTimeDerivative_weakForm(t) on the new state before the analysis
        if testcase_now:
            rows.append(np.concatenate([[t], analysis.analyze(op.get_state(), *grads)]))
        if doAnalyze:
            tAnalyze = min(tAnalyze + analyze_dt, tend)
            doAnalyze = False
        if doFinalize or (max_iter is not None and it >= max_iter):
            break
    return np.array(rows)
