"""galx_surf_int_noalloc (no device allocation: staging in the library's file-scope device scratch) against
galx_surf_int (device allocations per call), through the C ABI: bitwise equal (the same kernel on the same staged
values), on two calls in a row whose second input is smaller, so the second call runs on scratch that still holds
the first call's larger contents; and a request beyond the scratch size is refused with GALX_E_ARG."""
import ctypes as C

import numpy as np
import pytest

from galaexi_amd import mesh as M

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def call(fn, m, nVar, Fm, Fs, Ut):
    from galaexi_amd import _lib
    L = _lib.lib()
    f = getattr(L, fn)
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))  # noqa: E731
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    s2e = np.asfortranarray(m.SideToElem, dtype=np.int32)
    s2v2 = np.asfortranarray(m.S2V2, dtype=np.int32)
    out = np.asfortranarray(Ut.copy())
    rc = f(0, nVar, m.N, m.nElems, m.nSides, ip(s2e), ip(s2v2), dp(Fm), dp(Fs), C.c_double(4.5), C.c_double(4.5),
           dp(out))
    return rc, out


@pytest.mark.parametrize("sizes", [(((4, 3, 3), 5), ((2, 2, 2), 3))])
def test_surf_int_noalloc_matches_allocating_entry(sizes):
    rng = np.random.default_rng(9)
    for nel, N in sizes:
        say(f"case: surf_int_noalloc vs surf_int, mesh {nel} N={N}")
        m = M.box_mesh(nel, N, lo=(-1, -1, -1), hi=(1, 1, 1), deform_amp=0.05, deform_phase=(0.1, 0.2, 0.3),
                       rotations="random", seed=N)
        n, nVar = N + 1, 5
        Fm = np.asfortranarray(rng.standard_normal((nVar, n, n, m.nSides)))
        Fs = np.asfortranarray(rng.standard_normal((nVar, n, n, m.nSides)))
        Ut = np.asfortranarray(rng.standard_normal((nVar, n, n, n, m.nElems)))
        rc_a, ref = call("galx_surf_int", m, nVar, Fm, Fs, Ut)
        rc_b, out = call("galx_surf_int_noalloc", m, nVar, Fm, Fs, Ut)
        assert rc_a == 0 and rc_b == 0
        assert np.array_equal(out, ref)
        assert not np.array_equal(out, Ut)


def test_surf_int_noalloc_refuses_oversize():
    from galaexi_amd import _lib
    say("case: surf_int_noalloc oversize request")
    m = M.box_mesh((16, 16, 16), 7, lo=(-1, -1, -1), hi=(1, 1, 1))
    n, nVar = 8, 5
    Fm = np.zeros((nVar, n, n, m.nSides), order="F")
    Ut = np.zeros((nVar, n, n, n, m.nElems), order="F")
    rc, _ = call("galx_surf_int_noalloc", m, nVar, Fm, Fm, Ut)
    assert rc == _lib.GALX_E_ARG
    assert b"scratch" in _lib.lib().galx_last_error()
