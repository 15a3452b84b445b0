"""Manufactured-source runs on the GPU: the source against the oracle, and the reference's convergence checks
(regressioncheck/checks/convtest/h_3D and p_3D) replayed on the reference's own meshes.

Setup of both checks (their parameter.ini): ExactFunc 4 (oblique wave rho = rho*u = rho*v = rho*w = 2 + 0.1 sin(pi(x+y+z)
- 2 pi 0.3 t), E = rho^2) with its source, AdvVel (0.3,0.3,0.3), mu0 1e-3, Pr 0.72, kappa 1.4, Carpenter RK4-5 to
tend = 1.0 with dt from CALCTIMESTEP every step; Riemann RoeEntropyFix (the reference's build default,
src/CMakeLists.txt:98; the parameter is ignored, riemann.f90:161-192), Gauss-Lobatto nodes, weak and split (PI) form
(convtest/builds.ini builds FLEXI_SPLIT_DG ON and OFF).
  h_3D: N=3 on CART_HEX_PERIODIC_0{02,04,08,16}_3D_mesh.h5 (two HOPR zones on [-1,1]^3), CFL = DFL = 0.7,
        NAnalyze 10.  analyze.ini: L2 < 0.5, convergence order vs N+1 within tolerance 0.15, rate 0.8.
  p_3D: N=2..13 on CART_HEX_PERIODIC_002_3D_mesh.h5, CFL = DFL = 0.25, NAnalyze 2(N+1).  analyze.ini: L2 < 5,
        p-convergence rate 0.8, percentage 0.75.
The analyze.ini keys are read by reggie2.0, the reference's external regression tool, which is not part of the
reference repository.  The criteria here are this file's reading of them: h_3D passes when, for at least 80 % of the
variables, the mean of the orders between consecutive meshes lies within 15 % of N+1; p_3D passes when, for every
variable, at least 75 % of the steps N -> N+1 cut the L2 error to 0.8 of its value or less.  The thresholds are the
reference's; how reggie combines them is unpinned.
Each GPU run is also compared with the FP64 oracle's run of the same case (same time loop): relative L2 of the final
state <= 1e-10 after the whole run (dt from the two CALCTIMESTEP implementations agrees to rounding).
p_3D runs every degree of the reference's list, N = 2..13 (N >= 10 on the high-order kernels, whose elements have
more volume nodes than a workgroup has threads).  Weak form at N=2 on the 2^3 mesh is not run: the state (Mach 3.3 at
this exact solution, p ~ 0.4) loses positivity there near t = 0.97 in both the oracle and the GPU path, the
aliasing instability of the collocated weak form at three nodes per element; test_weak_n2_diverges_like_the_oracle
pins that both stop at the same step.
"""
import os

import numpy as np
import pytest

import oracle as O
from galaexi_amd import convtest as CT
from galaexi_amd import meshfile
from galaexi_amd import testcases as T

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ADV = (0.3, 0.3, 0.3)
GAS = dict(kappa=1.4, R=287.058, Pr=0.72, mu0=1e-3)
RIE = {"lf": O.RIEMANN_LF, "roe": O.RIEMANN_ROE, "roeentropyfix": O.RIEMANN_ROEEF}


def say(*a):
    print(*a, flush=True)


def rel_l2(a, b):
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300)


def conv_mesh(ne, N):
    return meshfile.read_mesh(os.path.join(GOLD, f"CART_HEX_PERIODIC_{ne:03d}_3D_mesh.h5"), N)


def gpu_op(m, form, cfl, riemann="roeentropyfix", precision="fp64", case=4, mu0=GAS["mu0"]):
    from galaexi_amd import dg
    gas = dict(GAS, mu0=mu0)
    op = dg.DGOperator(m.N, form=form, riemann=riemann, precision=precision, CFLScale=cfl, DFLScale=cfl, **gas)
    op.set_mesh(m)
    op.set_source(case, ADV, m.Elem_xGP)
    return op


class OracleRun:
    """The oracle behind the operator interface that convtest.run drives."""

    def __init__(self, m, form, cfl, riemann="roeentropyfix", case=4, mu0=GAS["mu0"]):
        self.op = O.Operator(m, split=form == "split", riemann=RIE[riemann], **dict(GAS, mu0=mu0))
        self.op.set_source(case, ADV, m.Elem_xGP)
        self.cfl = cfl

    def set_state(self, U):
        self.U = np.asfortranarray(np.array(U, dtype=np.float64))
        self.K = np.zeros_like(self.U, order="F")

    def get_state(self):
        return self.U.copy(order="F")

    def CalcTimeStep(self):
        dt, err, _, _ = self.op.calc_dt(self.U, self.cfl, self.cfl)
        return dt, err

    def TimeStepByLSERKW2(self, t, dt, nSteps=1):
        for s in range(nSteps):
            self.op.lserk_step(self.U, self.K, t + s * dt, dt)


@pytest.mark.parametrize("form,riemann,precision,N,bar", [
    ("weak", "roeentropyfix", "fp64", 3, 1e-12),
    ("split", "roeentropyfix", "fp64", 4, 1e-12),
    ("split", "lf", "fp64", 5, 1e-12),
    ("split", "lf", "mixed", 3, 1e-4),
])
def test_rhs_with_source_vs_oracle(form, riemann, precision, N, bar):
    """DGTimeDerivative_weakForm(t) with each source on the reference's 4^3 two-zone mesh, a perturbed wave, t=0.37."""
    say(f"case: rhs with source, {form} {riemann} {precision} N={N}, 4^3 convtest mesh")
    m = conv_mesh(4, N)
    rng = np.random.default_rng(N)
    t = 0.37
    ref = O.Operator(m, split=form == "split", riemann=RIE[riemann], **GAS)
    op = gpu_op(m, form, 0.7, riemann, precision)
    for case in T.SOURCE_FUNCS:
        U = T.synthetic_wave(case, m.Elem_xGP, t, ADV)
        U = np.asfortranarray(U * (1.0 + 1e-3 * rng.standard_normal(U.shape)) + 1e-3 * (U == 0))
        op.set_source(case, ADV, m.Elem_xGP)
        op.set_state(U)
        Ut = op.DGTimeDerivative_weakForm(t)
        ref.set_source(case, ADV, m.Elem_xGP)
        Ut_ref = ref.rhs(U, t)
        errs = [rel_l2(Ut[v], Ut_ref[v]) for v in range(5)]
        S = ref.source(t)
        say(f"  ExactFunc {case}: per-variable rel L2 {['%.1e' % e for e in errs]}; |S|/|Ut| "
            f"{np.linalg.norm(S) / np.linalg.norm(Ut_ref):.2f}")
        assert max(errs) <= bar
    # any other exact function switches the source off (CASE DEFAULT)
    op.set_source(7, ADV, m.Elem_xGP)
    ref.set_source(7, ADV, m.Elem_xGP)
    Ut = op.DGTimeDerivative_weakForm(t)
    assert max(rel_l2(Ut[v], ref.rhs(U, t)[v]) for v in range(5)) <= bar
    op.close()


def test_lserk_with_source_vs_oracle():
    """20 LSERK steps from t0 = 0.25 (the source at every stage time), the steps entry and the stage entry."""
    say("case: lserk with source, weak roeentropyfix N=3, 4^3 convtest mesh, 20 steps from t=0.25")
    m = conv_mesh(4, 3)
    U0 = T.synthetic_wave(4, m.Elem_xGP, 0.25, ADV)
    op = gpu_op(m, "weak", 0.7)
    op.set_state(U0)
    dt, err = op.CalcTimeStep()
    assert err == 0
    t0, n = 0.25, 20
    op.TimeStepByLSERKW2(t0, dt, n)
    Ug = op.get_state()
    ref = OracleRun(m, "weak", 0.7)
    ref.set_state(U0)
    dt_ref, _ = ref.CalcTimeStep()
    assert abs(dt - dt_ref) <= 1e-14 * dt_ref
    ref.TimeStepByLSERKW2(t0, dt, n)
    rel = rel_l2(Ug, ref.U)
    say(f"  dt {dt:.4e}: GPU vs oracle rel L2 {rel:.2e}")
    assert rel <= 1e-12
    # the stage entry: five stage calls make the same step as one steps call, bit for bit
    a, b = gpu_op(m, "weak", 0.7), gpu_op(m, "weak", 0.7)
    a.set_state(U0)
    b.set_state(U0)
    for s in range(3):
        a.TimeStepByLSERKW2(t0 + s * dt, dt, 1)
        for stage in range(1, 6):
            b.lserk_stage(t0 + s * dt, dt, stage)
    assert np.array_equal(a.get_state(), b.get_state())
    for o in (op, a, b):
        o.close()


def test_adjoint_with_source_vs_fd():
    """The reverse sweep recomputes the forward trajectory with the source at the same stage times: dJ/dU0 . v of
    J = <w, U(t0 + 12 dt)> against central differences of the GPU time integration (split LF, the adjoint's form)."""
    say("case: lserk adjoint with source, split lf N=3, 2^3 convtest mesh, 12 steps from t=0.3, ckpt 5")
    m = conv_mesh(2, 3)
    rng = np.random.default_rng(5)
    U0 = T.synthetic_wave(4, m.Elem_xGP, 0.3, ADV)
    U0 = np.asfortranarray(U0 * (1.0 + 1e-2 * rng.standard_normal(U0.shape)))
    w = np.asfortranarray(rng.standard_normal(U0.shape))
    op = gpu_op(m, "split", 0.7, "lf")
    op.set_state(U0)
    dt, _ = op.CalcTimeStep()
    t0, n = 0.3, 12
    op.forward_ckpt(t0, dt, n, 5)
    g = op.lserk_adjoint(w)
    fwd = gpu_op(m, "split", 0.7, "lf")

    def J(X):
        fwd.set_state(np.asfortranarray(X))
        fwd.TimeStepByLSERKW2(t0, dt, n)
        return float(np.sum(w * fwd.get_state()))

    nosrc = gpu_op(m, "split", 0.7, "lf", case=0)
    nosrc.set_state(U0)
    nosrc.forward_ckpt(t0, dt, n, 5)
    g0 = nosrc.lserk_adjoint(w)
    say(f"  the source changes the gradient by {rel_l2(g, g0):.2e} (relative)")
    assert rel_l2(g, g0) > 1e-4
    for it in range(2):
        v = np.asfortranarray(0.01 * np.abs(U0) * rng.standard_normal(U0.shape))
        exact = float(np.sum(g * v))
        errs = []
        for eps in (1e-3, 1e-4, 1e-5, 1e-6):  # central differences cross the LF max() switch: best over eps
            fd = (J(U0 + eps * v) - J(U0 - eps * v)) / (2 * eps)
            errs.append(abs(fd - exact) / abs(exact))
        say(f"  dir {it}: g.v={exact:.6e} fd rel errs {['%.1e' % e for e in errs]}")
        assert min(errs) <= 1e-6
    for o in (op, fwd, nosrc):
        o.close()


@pytest.mark.parametrize("form,mu0", [("weak", 1e-3), ("split", 1e-3), ("weak", 0.0), ("split", 0.0)])
def test_convtest_h_3d(form, mu0):
    """mu0 = 0 is the Euler build (FLEXI_PARABOLIC=OFF in convtest/builds.ini): no viscous terms, no viscous time
    step limit, and the source without its heat-flux term."""
    N, cells = 3, (2, 4, 8, 16)
    say(f"case: convtest h_3D, {form} form, mu0={mu0}, N={N}, meshes {cells}")
    errs = []
    for ne in cells:
        m = conv_mesh(ne, N)
        op = gpu_op(m, form, 0.7, mu0=mu0)
        U0 = T.synthetic_wave(4, m.Elem_xGP, 0.0, ADV)
        U, steps = CT.run(op, U0, 1.0)
        op.close()
        L2, Linf = CT.ErrorNorms(m, 10).wave(U, 4, 1.0, ADV)
        errs.append(L2)
        line = f"  {ne:2d}^3: {steps} steps, L2 {['%.3e' % e for e in L2]}"
        if ne <= 8:
            Ur, steps_r = CT.run(OracleRun(m, form, 0.7, mu0=mu0), U0, 1.0)
            rel = rel_l2(U, Ur)
            line += f"; vs oracle run: rel L2 {rel:.1e}, {steps_r} steps"
            assert steps == steps_r and rel <= 1e-10
        say(line)
        assert np.all(L2 < 0.5)
    orders = CT.orders(2.0 / np.array(cells), np.array(errs))
    mean = orders.mean(axis=0)
    say(f"  orders {np.array2string(orders, precision=2)}; mean {np.array2string(mean, precision=2)}")
    passed = np.abs(mean / (N + 1) - 1.0) <= 0.15
    assert passed.mean() >= 0.8, mean


@pytest.mark.parametrize("form,Ns", [("split", tuple(range(2, 14))), ("weak", tuple(range(3, 14)))])
def test_convtest_p_3d(form, Ns):
    say(f"case: convtest p_3D, {form} form, N={Ns[0]}..{Ns[-1]}, 2^3 mesh")
    errs = []
    for N in Ns:
        m = conv_mesh(2, N)
        op = gpu_op(m, form, 0.25)
        U0 = T.synthetic_wave(4, m.Elem_xGP, 0.0, ADV)
        U, steps = CT.run(op, U0, 1.0)
        op.close()
        L2, _ = CT.ErrorNorms(m).wave(U, 4, 1.0, ADV)
        Ur, steps_r = CT.run(OracleRun(m, form, 0.25), U0, 1.0)
        rel = rel_l2(U, Ur)
        say(f"  N={N}: {steps} steps, L2 {['%.3e' % e for e in L2]}; vs oracle run: rel L2 {rel:.1e}")
        assert steps == steps_r and rel <= 1e-10
        assert np.all(L2 < 5.0)
        errs.append(L2)
    ratio = np.array(errs[1:]) / np.array(errs[:-1])
    frac = (ratio <= 0.8).mean(axis=0)
    say(f"  error ratios N+1/N per variable: {np.array2string(ratio[:, [0, 4]], precision=2)} (rho, E)")
    assert np.all(frac >= 0.75), frac


def test_weak_n2_diverges_like_the_oracle():
    say("case: convtest p_3D weak form N=2 (positivity lost near t=0.97), GPU and oracle")
    m = conv_mesh(2, 2)
    U0 = T.synthetic_wave(4, m.Elem_xGP, 0.0, ADV)
    with pytest.raises(CT.TimeStepError) as eg:
        CT.run(gpu_op(m, "weak", 0.25), U0, 1.0)
    with pytest.raises(CT.TimeStepError) as er:
        CT.run(OracleRun(m, "weak", 0.25), U0, 1.0)
    say(f"  GPU: {eg.value}; oracle: {er.value}")
    assert eg.value.step == er.value.step and abs(eg.value.t - er.value.t) <= 1e-12


def test_set_source_call_order():
    """galx_set_source needs the mesh (GALX_E_STATE before it) and switches off for any other exact function."""
    from galaexi_amd import dg
    from galaexi_amd._lib import GalxError

    say("case: set_source call order")
    op = dg.DGOperator(3, form="weak", riemann="roeentropyfix", **GAS)
    with pytest.raises(GalxError):
        op.set_source(4, ADV)  # no mesh yet
    m = conv_mesh(2, 3)
    op.set_mesh(m)
    op.set_source(0, ADV)  # off: no coordinates needed
    U = T.synthetic_wave(4, m.Elem_xGP, 0.0, ADV)
    op.set_state(U)
    Ut0 = op.DGTimeDerivative_weakForm(0.2)
    op.set_source(4, ADV, m.Elem_xGP)
    Ut4 = op.DGTimeDerivative_weakForm(0.2)
    ref = O.Operator(m, split=False, riemann=O.RIEMANN_ROEEF, **GAS)
    ref.set_source(4, ADV, m.Elem_xGP)
    S = ref.source(0.2)
    assert rel_l2(Ut4 - Ut0, S) <= 1e-12  # the source is additive in physical space
    op.close()
