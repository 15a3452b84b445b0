"""The oracle against the reference's own unit-test fixtures (unitTests/*.bin, converted to .npz by
tests/golden/convert_unittest_bins.py).  Tolerance as in the reference tests: ALMOSTEQUALABSORREL with
100*PP_RealTolerance (src/flexi.h:66-69, unitTests/ProlongToFace.f90:132-137)."""
import os

import numpy as np
import pytest

import oracle as O

EPS = np.finfo(np.float64).eps


def almost_equal_abs_or_rel(a, b, tol=100 * EPS):
    d = np.abs(a - b)
    return np.all((d <= tol) | (d <= np.maximum(np.abs(a), np.abs(b)) * tol))


@pytest.fixture(scope="module")
def elem(golden_dir):
    return np.load(os.path.join(golden_dir, "unittest_reference_element.npz"))


def test_prolong_to_face_gl3d(elem, golden_dir):
    """unitTests/ProlongToFace.f90: Uvol on the curved NRef=9 element -> Uface_master."""
    g = np.load(os.path.join(golden_dir, "unittest_prolongtoface_gl3d.npz"))
    N = int(elem["NRef"])
    Uv = np.asfortranarray(np.repeat(g["Uvol"][None], 5, axis=0))
    Um, _ = O.prolong_to_face(N, 1, 6, elem["SideToElem"], elem["S2V2"], Uv, 5)
    for v in range(5):
        assert almost_equal_abs_or_rel(Um[v], g["Uface_master_dg"])
    # closed form: GL trace = boundary node values (pure gather, bit-exact)
    assert np.array_equal(Um[0], g["Uface_master_dg"])


def test_surfint_gl3d(elem, golden_dir):
    """unitTests/SurfInt.f90: SurfIntCons of a random face flux, GL non-EXACT_MM."""
    g = np.load(os.path.join(golden_dir, "unittest_surfint_gl3d.npz"))
    N = int(elem["NRef"])
    Fl = np.asfortranarray(np.repeat(g["Flux"][None], 5, axis=0))
    Ut = O.surfint(N, 1, 6, elem["SideToElem"], elem["S2V2"], Fl, Fl, float(elem["L_HatMinus"][0]),
                   float(elem["L_HatPlus"][N]), 5)
    for v in range(5):
        assert almost_equal_abs_or_rel(Ut[v][..., 0], g["Ut_dg"][..., 0])
    # plausible mistakes must fail: the interior L_Hat entry, or a transposed face orientation
    bad = O.surfint(N, 1, 6, elem["SideToElem"], elem["S2V2"], Fl, Fl, float(elem["L_HatMinus"][1]),
                    float(elem["L_HatPlus"][N]), 5)
    assert not almost_equal_abs_or_rel(bad[0][..., 0], g["Ut_dg"][..., 0])
    s2v2_t = np.asfortranarray(elem["S2V2"][::-1])
    bad = O.surfint(N, 1, 6, elem["SideToElem"], s2v2_t, Fl, Fl, float(elem["L_HatMinus"][0]),
                    float(elem["L_HatPlus"][N]), 5)
    assert not almost_equal_abs_or_rel(bad[0][..., 0], g["Ut_dg"][..., 0])


@pytest.mark.parametrize("it,ntype", [(0, O.NODE_G), (1, O.NODE_GL), (2, O.NODE_CL), (3, O.NODE_VISU)])
def test_nodes_weights_derivative(golden_dir, it, ntype):
    """unitTests/NodesAndWeights.f90 and unitTests/DerivativeMatrix.f90 for N=1..10 and 4 node types."""
    b = np.load(os.path.join(golden_dir, "unittest_basis.npz"))
    for N in range(1, 11):
        x, w, wb = O.nodes_and_weights(N, ntype)
        assert almost_equal_abs_or_rel(x, b["xi"][: N + 1, N - 1, it])
        assert almost_equal_abs_or_rel(w, b["w"][: N + 1, N - 1, it])
        assert almost_equal_abs_or_rel(wb, b["wBary"][: N + 1, N - 1, it])
        D = O.derivative_matrix(N, ntype)
        assert almost_equal_abs_or_rel(D, b["D"][: N + 1, : N + 1, it * 10 + N - 1])


def test_dg_basis_closed_forms():
    """D exact on polynomials of degree N; GL weights integrate degree 2N-1 exactly; DVolSurf corner
    entries vanish for GL (src/dg/dg.f90:225-237); D_Hat = -M^-1 D^T M."""
    for N in (1, 3, 5, 7):
        b = O.dg_basis(N)
        x, w, D = b["xGP"], b["wGP"], b["D"]
        for p in range(N + 1):
            assert np.allclose(D @ x ** p, p * x ** max(p - 1, 0) if p else 0 * x, atol=1e-12)
        for p in range(2 * N):
            exact = (1 - (-1) ** (p + 1)) / (p + 1)
            assert abs(np.dot(w, x ** p) - exact) < 1e-13
        assert abs(b["DVolSurf"][0, 0]) < 1e-12 and abs(b["DVolSurf"][N, N]) < 1e-12
        Dh = -np.diag(1 / w) @ D.T @ np.diag(w)
        assert np.allclose(b["D_Hat_T"], Dh.T, rtol=1e-15, atol=1e-15)
        # SBP property: M D + D^T M = B
        B = np.zeros((N + 1, N + 1)); B[0, 0] = -1; B[N, N] = 1
        assert np.allclose(np.diag(w) @ D + D.T @ np.diag(w), B, atol=1e-13)
