"""State files on the GPU path: a run that writes its state (WriteState, src/io_hdf5/hdf5_output.f90:84-216) and is
restarted from that file into a fresh context (Restart, src/restart/restart.f90:304-560, RestartMode 1) continues
exactly as the uninterrupted run.

Bar: FP64, bit-identical final state (the state file holds U exactly, the first LSERK stage does not read the
register, A_1 = 0, and the traces the restarted context forms from U are the records the uninterrupted kernels form
from the same values).  Mixed form: the restart recomputes the pressure offset of the FP32 arithmetic from the
restart state (the mean pressure has moved since t=0), so the continuation is not bitwise the uninterrupted one.  It
must agree with it at least as closely as the uninterrupted mixed run agrees with the FP64 run over the same steps:
the restart adds no error beyond the form's own.
"""
import os

import numpy as np
import pytest

from galaexi_amd import h5
from galaexi_amd import mesh as M
from galaexi_amd import statefile as S
from galaexi_amd import testcases as T

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not h5.available(), reason="no HDF5 C library on this machine")]

GAS = dict(kappa=1.4, R=71.42857, Pr=0.72, mu0=1.0 / 1600.0, CFLScale=0.5, DFLScale=0.5)


def say(*a):
    print(*a, flush=True)


def op_for(m, precision):
    from galaexi_amd import dg
    op = dg.DGOperator(m.N, form="split", riemann="lf", precision=precision, **GAS)
    op.set_mesh(m)
    return op


def run_uninterrupted(m, U0, precision, dt, steps):
    op = op_for(m, precision)
    op.set_state(U0)
    op.TimeStepByLSERKW2(0.0, dt, steps)
    U = op.get_state()
    op.close()
    return U


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_write_restart_continue(tmp_path, precision):
    say(f"case: write state / restart / continue, curved TGV 3^3 N=7, {precision}")
    m = M.box_mesh((3, 3, 3), 7, lo=(0.0, 0.0, 0.0), hi=(2 * np.pi,) * 3, deform_amp=0.05,
                   deform_phase=(0.31, 0.17, 0.83))
    U0 = T.taylor_green(m.Elem_xGP, Ma=0.1, noise=1e-3, seed=1600)
    ref = op_for(m, precision)
    ref.set_state(U0)
    dt, err = ref.CalcTimeStep()
    assert err == 0
    k1, k2 = 7, 6
    t1 = ref.TimeStepByLSERKW2(0.0, dt, k1)
    path = S.write_state(str(tmp_path), "tgv", ref.get_state(), m.N, "box.h5", t1)
    ref.TimeStepByLSERKW2(t1, dt, k2)
    U_ref = ref.get_state()
    ref.close()
    U_r, t_r = S.restart_state(path, m.N, m.nElems)
    assert t_r == t1 and os.path.basename(path) == S.state_file_name("tgv", t1)
    op = op_for(m, precision)
    op.set_state(U_r)
    op.TimeStepByLSERKW2(t_r, dt, k2)
    U = op.get_state()
    op.close()
    rel = np.linalg.norm((U - U_ref).ravel()) / np.linalg.norm(U_ref.ravel())
    say(f"  restarted vs uninterrupted after {k1}+{k2} steps: rel L2 {rel:.2e}")
    if precision == "fp64":
        assert np.array_equal(U, U_ref)
    else:
        U64 = run_uninterrupted(m, U0, "fp64", dt, k1 + k2)
        form = np.linalg.norm((U_ref - U64).ravel()) / np.linalg.norm(U64.ravel())
        say(f"  uninterrupted mixed vs FP64 after {k1 + k2} steps: rel L2 {form:.2e}")
        assert 0 < rel <= form
