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


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 1e-5)])
def test_gradients_match_oracle(precision, tol):
    """BR1 gradients of (u,v,w,T) on a curved, rotated mesh (N=5) against the oracle's gradUx/y/z rows 2..5:
    relative L2 per direction within 1e-12 (FP64) / 1e-5 (mixed)."""
    say(f"case: gradients {precision}, (3,2,2) N=5 curved rotated")
    m = M.box_mesh((3, 2, 2), 5, lo=(-1, -1, -1), hi=(1, 1, 1), deform_amp=0.05, deform_phase=(0.1, 0.4, 0.2),
                   rotations="random", seed=3)
    U = T.smooth_random(m.Elem_xGP, (-1, -1, -1), (1, 1, 1), seed=5)
    op = dgmod().DGOperator(5, form="split", riemann="lf", precision=precision, kappa=1.4, R=1.0, Pr=0.72, mu0=1e-2)
    op.set_mesh(m)
    op.set_state(U)
    g = op.get_gradients()
    ref = O.Operator(m, split=True, riemann=O.RIEMANN_LF, kappa=1.4, R=1.0, Pr=0.72, mu0=1e-2)
    ref.rhs(U)
    n = m.N + 1
    for d, w in enumerate((3, 4, 5)):
        r = ref.array(w).reshape((5, n, n, n, m.nElems), order="F")[1:5]
        err = np.linalg.norm((g[d] - r).ravel()) / np.linalg.norm(r.ravel())
        assert err <= tol, (d, err)
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
