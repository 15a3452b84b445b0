"""GPU tests of the host-facing pieces the TGV testcase uses: the lifted gradients (galx_get_gradients) against the
oracle's BR1 gradients, and the per-stage LSERK entry point against the full step."""
import numpy as np
import pytest

import oracle as O
from galaexi_amd import mesh as M
from galaexi_amd import testcases as T

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def dgmod():
    from galaexi_amd import dg
    return dg


@pytest.mark.parametrize("precision,tol,N,nvar", [("fp64", 1e-12, 5, 4), ("mixed", 1e-5, 5, 4), ("fp64", 1e-12, 5, 5),
                                                  ("mixed", 1e-5, 5, 5), ("fp64", 1e-12, 11, 5)])
def test_gradients_match_oracle(precision, tol, N, nvar):
    """BR1 gradients on a curved, rotated mesh against the oracle's gradUx/y/z, which lifts the reference's default
    variables (rho,u,v,w,T): nVarLifting 4 = its rows 2..5 (PP_OPTLIFT), 5 = all rows.  Relative L2 per direction
    and variable within 1e-12 (FP64) / 1e-5 (mixed); N = 11 runs the high-order kernels."""
    say(f"case: gradients {precision}, nVarLifting {nvar}, (3,2,2) N={N} curved rotated")
    m = M.box_mesh((3, 2, 2), N, lo=(-1, -1, -1), hi=(1, 1, 1), deform_amp=0.05, deform_phase=(0.1, 0.4, 0.2),
                   rotations="random", seed=3)
    U = T.smooth_random(m.Elem_xGP, (-1, -1, -1), (1, 1, 1), seed=5)
    op = dgmod().DGOperator(N, form="split", riemann="lf", precision=precision, kappa=1.4, R=1.0, Pr=0.72,
                            mu0=1e-2)
    op.set_mesh(m)
    op.set_state(U)
    g = op.get_gradients(nvar)
    ref = O.Operator(m, split=True, riemann=O.RIEMANN_LF, kappa=1.4, R=1.0, Pr=0.72, mu0=1e-2)
    ref.rhs(U)
    n = m.N + 1
    for d, w in enumerate((3, 4, 5)):
        r = ref.array(w).reshape((5, n, n, n, m.nElems), order="F")[5 - nvar:]
        for v in range(nvar):
            err = np.linalg.norm((g[d][v] - r[v]).ravel()) / np.linalg.norm(r[v].ravel())
            assert err <= tol, (d, v, err)
    if nvar == 5:
        assert np.array_equal(g[0][1:], op.get_gradients(4)[0])  # the same (u,v,w,T) rows either way
    op.close()


def test_five_stages_equal_one_step():
    say("case: lserk_stage x5 vs lserk_step, (2,2,2) N=4")
    m = M.box_mesh((2, 2, 2), 4, lo=(-1, -1, -1), hi=(1, 1, 1), deform_amp=0.05, deform_phase=(0.3, 0.2, 0.1))
    U = T.smooth_random(m.Elem_xGP, (-1, -1, -1), (1, 1, 1), seed=9)
    a = dgmod().DGOperator(4, form="split", riemann="roeentropyfix", mu0=1e-2, R=1.0)
    b = dgmod().DGOperator(4, form="split", riemann="roeentropyfix", mu0=1e-2, R=1.0)
    for op in (a, b):
        op.set_mesh(m)
        op.set_state(U)
    dt, _ = a.CalcTimeStep()
    for _ in range(3):
        a.TimeStepByLSERKW2(0.0, dt)
        for s in range(1, 6):
            b.lserk_stage(0.0, dt, s)
    assert np.array_equal(a.get_state(), b.get_state())
    a.close()
    b.close()
