"""Host-side geometry producer (galaexi_amd.mesh/basis) against the reference fixtures, the reference's own
mapping self-checks, closed forms and the FP64 oracle (free-stream preservation)."""
import os

import numpy as np
import pytest

import oracle as O
from galaexi_amd import basis as B
from galaexi_amd import mesh as M
from galaexi_amd import testcases as T

EPS = np.finfo(np.float64).eps


def close(a, b, tol=100 * EPS):
    d = np.abs(a - b)
    return np.all((d <= tol) | (d <= tol * np.maximum(np.abs(a), np.abs(b))))


@pytest.mark.parametrize("it,nt", [(0, B.GAUSS), (1, B.GAUSS_LOBATTO), (2, B.CHEBYSHEV_LOBATTO), (3, B.VISU)])
def test_numpy_basis_vs_fixture(golden_dir, it, nt):
    g = np.load(os.path.join(golden_dir, "unittest_basis.npz"))
    for N in range(1, 11):
        x, w, wb = B.nodes_and_weights(N, nt)
        assert close(x, g["xi"][: N + 1, N - 1, it])
        assert close(w, g["w"][: N + 1, N - 1, it])
        assert close(wb, g["wBary"][: N + 1, N - 1, it])
        assert close(B.derivative_matrix(x), g["D"][: N + 1, : N + 1, it * 10 + N - 1])


def test_s2v2_matches_fixture_and_self_checks(golden_dir):
    """S2V2 equals the reference element's table (UnittestElementData3D.bin) and passes the sanity checks of
    buildMappings (src/mesh/mappings.f90:188-226): SideToVol2 consistent with VolToSide for all flips/sides."""
    el = np.load(os.path.join(golden_dir, "unittest_reference_element.npz"))
    N = int(el["NRef"])
    s2v2 = M.build_s2v2(N)
    assert np.array_equal(s2v2, el["S2V2"])

    def cgns_vol_to_side(i, j, k, ls):
        return {5: (k, j, i), 3: (j, k, N - i), 2: (i, k, j), 4: (N - i, k, N - j), 1: (j, i, k), 6: (i, j, N - k)}[ls]

    for ls in range(1, 7):
        for f in range(5):
            for k in range(0, N + 1, 3):
                for j in range(N + 1):
                    for i in range(N + 1):
                        p, q, _l = cgns_vol_to_side(i, j, k, ls)
                        pq = M.flip_s2m(N, p, q, f)
                        a, b = s2v2[:, pq[0], pq[1], f, ls - 1]
                        want = {5: (j, k), 3: (j, k), 2: (i, k), 4: (i, k), 1: (i, j), 6: (i, j)}[ls]
                        assert (a, b) == want


@pytest.mark.parametrize("rot", [None, "random"])
def test_connectivity_consistent(rot):
    m = M.box_mesh((3, 2, 2), 3, rotations=rot, seed=9)
    assert m.nSides == 3 * m.nElems
    s2e, e2s = m.SideToElem, m.ElemToSide
    seen = np.zeros((m.nSides, 2), dtype=int)
    for e in range(m.nElems):
        for ls in range(1, 7):
            sid, flip, ism = e2s[:, ls - 1, e]
            assert (flip == 0) == (ism == 1)
            if flip == 0:
                assert s2e[0, sid - 1] == e + 1 and s2e[2, sid - 1] == ls
            else:
                assert s2e[1, sid - 1] == e + 1 and s2e[3, sid - 1] == ls and s2e[4, sid - 1] == flip
            seen[sid - 1, 0 if flip == 0 else 1] += 1
    assert np.all(seen == 1)
    if rot:
        assert len(set(s2e[4].tolist())) >= 3


def test_affine_metrics_closed_form():
    """Affine box: Ja^d = (h_e h_f / 4) e_d, sJ = 8/(hx hy hz), |n| = 1 along the axes, SurfElem = h_e h_f / 4."""
    lo, hi, ne, N = np.array([0.0, -1.0, 2.0]), np.array([3.0, 1.0, 2.5]), (3, 4, 2), 4
    m = M.box_mesh(ne, N, lo=tuple(lo), hi=tuple(hi))
    h = (hi - lo) / np.array(ne)
    ja = [np.array([h[1] * h[2], 0, 0]), np.array([0, h[0] * h[2], 0]), np.array([0, 0, h[0] * h[1]])]
    for d, arr in enumerate((m.Metrics_fTilde, m.Metrics_gTilde, m.Metrics_hTilde)):
        assert np.allclose(arr, (ja[d] / 4.0)[:, None, None, None, None], rtol=1e-13, atol=1e-14)
    assert np.allclose(m.sJ, 8.0 / np.prod(h), rtol=1e-13)
    assert np.allclose(np.linalg.norm(m.NormVec, axis=0), 1.0, rtol=1e-14)
    for s in range(m.nSides):
        ax = int(np.argmax(np.abs(m.NormVec[:, 0, 0, s])))
        others = [a for a in range(3) if a != ax]
        assert np.allclose(m.SurfElem[:, :, s], h[others[0]] * h[others[1]] / 4.0, rtol=1e-13)
        # normal, tangents orthonormal and right-handed
        n_, t1, t2 = m.NormVec[:, 0, 0, s], m.TangVec1[:, 0, 0, s], m.TangVec2[:, 0, 0, s]
        assert abs(np.dot(n_, t1)) < 1e-14 and abs(np.dot(n_, t2)) < 1e-14
        assert np.allclose(np.cross(n_, t1), t2)


def test_curved_metrics_vs_analytic_jacobian():
    """Curl-form metrics of the deformed mapping approach the analytic cofactors J*grad(xi) at high N."""
    amp, ph = 0.05, (0.1, 0.2, 0.3)
    m = M.box_mesh((2, 2, 2), 9, lo=(-1, -1, -1), hi=(1, 1, 1), deform_amp=amp, deform_phase=ph)
    # analytic mapping x(X) = X + amp*prod sin(pi(X+ph)) (1,1,1); element frames are axis aligned, so
    # d x/d xi = dx/dX * (h/2) with h = 1
    x = m.Elem_xGP
    # invert the mapping for the reference point X by fixed-point iteration
    X = x.copy()
    for _ in range(60):
        s = np.prod([np.sin(np.pi * (X[d] + ph[d])) for d in range(3)], axis=0)
        X = x - amp * s[None]
    sn = [np.sin(np.pi * (X[d] + ph[d])) for d in range(3)]
    cs = [np.cos(np.pi * (X[d] + ph[d])) for d in range(3)]
    grad = np.stack([amp * np.pi * cs[0] * sn[1] * sn[2], amp * np.pi * sn[0] * cs[1] * sn[2],
                     amp * np.pi * sn[0] * sn[1] * cs[2]])
    Jm = np.zeros((3, 3) + x.shape[1:])
    for a in range(3):
        for b in range(3):
            Jm[a, b] = (1.0 if a == b else 0.0) + grad[b]  # dx_a/dX_b
    Jm *= 0.5  # d/dxi = (h/2) d/dX
    # Ja^d_c = cofactor: J * d xi_d/d x_c
    cof = np.zeros_like(Jm)
    for d in range(3):
        d1, d2 = (d + 1) % 3, (d + 2) % 3
        cof[d] = np.cross(Jm[:, d1], Jm[:, d2], axis=0)
    for d, arr in enumerate((m.Metrics_fTilde, m.Metrics_gTilde, m.Metrics_hTilde)):
        assert np.abs(arr - cof[d]).max() < 1e-4 * np.abs(cof[d]).max()
    detJ = np.einsum("aijke,aijke->ijke", Jm[:, 0], np.cross(Jm[:, 1], Jm[:, 2], axis=0))
    assert np.abs(1.0 / m.sJ - detJ).max() < 1e-4 * np.abs(detJ).max()


@pytest.mark.parametrize("split,riemann", [(True, O.RIEMANN_LF), (False, O.RIEMANN_ROE)])
def test_freestream_preserved_curved_rotated(split, riemann):
    """Curl metrics keep a constant state exactly steady on the curved mesh with rotated element frames."""
    m = M.box_mesh((3, 2, 2), 4, lo=(-1, -1, -1), hi=(1, 1, 1), deform_amp=0.05, deform_phase=(0.3, 0.1, 0.7),
                   rotations="random", seed=4)
    U = T.freestream(m.Elem_xGP)
    Ut = O.Operator(m, split=split, riemann=riemann, mu0=1e-3, R=1.0).rhs(U)
    scale = np.abs(m.sJ).max() * np.abs(U).max() * 25
    assert np.abs(Ut).max() < 1e-13 * scale
    # a plausible mistake in the metrics (cross-product form at this geometry degree is not free-stream
    # preserving once perturbed) must show: perturb Ja by 1e-6 and the residual grows far above round-off
    m.Metrics_fTilde = m.Metrics_fTilde * (1 + 1e-6 * np.sin(np.arange(m.Metrics_fTilde.size)).reshape(
        m.Metrics_fTilde.shape, order="F"))
    Ut2 = O.Operator(m, split=split, riemann=riemann, mu0=1e-3, R=1.0).rhs(U)
    assert np.abs(Ut2).max() > 1e3 * np.abs(Ut).max()
