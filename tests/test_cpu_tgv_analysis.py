"""The TGV testcase analysis (AnalyzeTestcase, src/testcase/taylorgreenvortex/testcase.f90:283-515) against the
reference's regression fixture (regressioncheck/checks/tgv/split, converted to tests/golden).

At t=0 the reference evaluates the analysis on the initial state and the BR1 gradients of the first operator call
(src/timedisc/timedisc.f90:124-125,150); the oracle's gradients of the same state through galaexi_amd.tgv must
reproduce that row, which pins the analysis code, the initial condition and the oracle's lifted gradients on
the reference's own case.  A planted mistake (y and z gradients exchanged) must fail."""
import os

import numpy as np
import pytest

import oracle as O
from galaexi_amd import mesh as M
from galaexi_amd import testcases as T
from galaexi_amd import tgv


@pytest.fixture(scope="module")
def ref(golden_dir):
    return np.load(os.path.join(golden_dir, "tgv_re1600_split_reference.npz"))


@pytest.fixture(scope="module")
def case():
    m = M.box_mesh((8, 8, 8), 7, lo=(0, 0, 0), hi=(2 * np.pi,) * 3)
    U = T.taylor_green(m.Elem_xGP, Ma=0.1)
    op = O.Operator(m, split=True, riemann=O.RIEMANN_ROEEF, kappa=1.4, R=71.42857, Pr=0.72, mu0=6.25e-4)
    op.rhs(U)
    n = m.N + 1
    g = [op.array(w).reshape((5, n, n, n, m.nElems), order="F")[1:4] for w in (3, 4, 5)]
    return m, U, g


def test_fixture_layout(ref):
    assert tuple(ref["columns"]) == tgv.COLUMNS
    d = ref["data"]
    assert d.shape == (554, 16) and d[0, 0] == 0.0 and d[-1, 0] == 13.0
    assert np.all(np.diff(d[:, 0]) > 0)


def test_initial_row_matches_reference(ref, case):
    m, U, g = case
    A = tgv.TGVAnalysis(m, kappa=1.4, R=71.42857, mu0=6.25e-4)
    row = A.analyze(U, *g)
    want = ref["data"][0, 1:]
    scale = np.maximum(np.abs(want), 1e-12 * np.abs(want).max())
    rel = np.abs(row - want) / scale
    # the reference prints 14 significant digits; DR_p and ED_D are round-off-level at t=0
    for i, name in enumerate(tgv.COLUMNS[1:]):
        if name in ("DR_p", "ED_D"):
            assert abs(row[i]) < 1e-12, (name, row[i])
        else:
            assert rel[i] < 1e-11, (name, row[i], want[i])
    # planted mistake: y and z gradients exchanged (strain-rate and dissipation terms change)
    bad = A.analyze(U, g[0], g[2], g[1])
    assert np.max((np.abs(bad - want) / scale)[[0, 1, 6, 7]]) > 1e-2
