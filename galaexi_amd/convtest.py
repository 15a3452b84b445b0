"""Convergence runs of the synthetic waves with a manufactured source (the reference's convtest/h_3D and p_3D
regression checks) on the galx engine.

  ErrorNorms   CalcErrorNorms (src/analyze/analyze.f90:383-479): L2 and Linf errors against an exact function,
               evaluated on NAnalyze Gauss-Lobatto points (InitAnalyzeBasis, analyze.f90:235-273;
               NAnalyze defaults to 2(N+1), analyze.f90:125; Vol from analyze.f90:158-167)
  run          TimeDisc (src/timedisc/timedisc.f90:100-200) with one analyze time at tend: dt from CALCTIMESTEP
               every step (nCalcTimeStepMax = 1), the last step shortened to end on tend, and a step that would
               leave less than 1 % of dt before tend stretched to reach it (UpdateTimeStep, timedisc_func.f90:246-296)
  orders       the convergence order between consecutive meshes, log(e_i/e_i+1) / log(h_i/h_i+1)
"""
from __future__ import annotations

import numpy as np

from . import basis
from . import testcases


def _interp3(V, f):
    """f(..., i, j, k, e) -> (..., a, b, c, e) with V(a, i) along each of the three node axes."""
    return np.einsum("ai,bj,ck,...ijke->...abce", V, V, V, f, optimize=True)


class ErrorNorms:
    def __init__(self, mesh, NAnalyze=None):
        N = mesh.N
        self.NAnalyze = 2 * (N + 1) if NAnalyze is None else NAnalyze
        _x, wGP, _wb = basis.nodes_and_weights(N, basis.GAUSS_LOBATTO)
        _xa, wA, _wba = basis.nodes_and_weights(self.NAnalyze, basis.GAUSS_LOBATTO)
        self.V = basis.vandermonde(N, basis.GAUSS_LOBATTO, self.NAnalyze, basis.GAUSS_LOBATTO)
        Wv = wGP[:, None, None] * wGP[None, :, None] * wGP[None, None, :]
        self.Vol = float(np.sum(Wv[..., None] / mesh.sJ))
        WA = wA[:, None, None] * wA[None, :, None] * wA[None, None, :]
        # the Jacobian itself (1/sJ) is interpolated, as the reference does
        self.IntWeight = WA[..., None] * _interp3(self.V, 1.0 / mesh.sJ)
        self.xA = _interp3(self.V, mesh.Elem_xGP)

    def __call__(self, U, exact):
        """U(5,0:N,0:N,0:N,nElems); exact(xA) -> U_exact(5, ...) on the analyze points.  Returns (L2(5), Linf(5))."""
        err = _interp3(self.V, np.asarray(U)) - exact(self.xA)
        L2 = np.sqrt(np.sum(err * err * self.IntWeight, axis=(1, 2, 3, 4)) / self.Vol)
        Linf = np.max(np.abs(err), axis=(1, 2, 3, 4))
        return L2, Linf

    def wave(self, U, case, t, AdvVel):
        """Errors against the synthetic wave `case` (ExactFunc 4, 41, 42, 43) at time t."""
        return self(U, lambda x: testcases.synthetic_wave(case, x, t, AdvVel))


class TimeStepError(RuntimeError):
    """CALCTIMESTEP reported an invalid state (errType: 1 non-positive density, 2/3 NaN convective/viscous
    eigenvalue) before step `step` at time t."""

    def __init__(self, errType, t, step):
        super().__init__(f"CALCTIMESTEP errType {errType} at t={t} (step {step})")
        self.errType, self.t, self.step = errType, t, step


def run(op, U0, tend, t0=0.0, max_steps=1_000_000):
    """Advance a DGOperator from U0 at t0 to tend.  Returns (final state, number of steps)."""
    op.set_state(U0)
    t, it = t0, 0
    while True:
        dt_cfl, err = op.CalcTimeStep()
        if err != 0:
            raise TimeStepError(err, t, it)
        left = tend - t
        dt = min(dt_cfl, left)
        final = left - dt < dt / 100.0
        if final:
            dt = left
        op.TimeStepByLSERKW2(t, dt, 1)
        it += 1
        t = tend if final else t + dt
        if final:
            return op.get_state(), it
        if it >= max_steps:
            raise RuntimeError(f"no arrival at tend={tend} within {max_steps} steps (t={t})")


def orders(h, e):
    """Convergence orders between consecutive entries of the mesh sizes h and errors e (rows: meshes)."""
    h = np.asarray(h, dtype=np.float64)
    e = np.asarray(e, dtype=np.float64)
    return np.log(e[:-1] / e[1:]) / np.log(h[:-1, None] / h[1:, None]) if e.ndim == 2 else \
        np.log(e[:-1] / e[1:]) / np.log(h[:-1] / h[1:])
