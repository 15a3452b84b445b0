"""HDF5 mesh ingest and state I/O against the reference's own files.

Fixtures (tests/golden/copy_hdf5_fixtures.py): the TGV regression mesh CART_HEX_PERIODIC_008_mesh.h5 (HOPR,
8^3 periodic [0,2pi]^3, NGeo=1), the freestream check's mesh cartbox3D_mesh.h5 (non-periodic boundaries) and
its state file cartbox_State_0000000.000001000.h5 (N=3, Gauss nodes, RefState rho=u=v=w=p=1).

What pins the ingest:
  * the faces close geometrically: every face node seen from the master side (S2V2 flip 0) and from the slave
    side (S2V2 with the slave's flip) lands on the same point up to the box period; a wrong flip, S2V2 or
    side pairing fails this;
  * the metrics equal those of the analytic 8^3 box (elements matched by centroid), and the oracle's Ut on the
    file mesh equals its Ut on the analytic box up to rounding;
  * NGeo=2 (curved) node ordering: a file written with the analytic deformation applied to NGeo=2 nodes
    reproduces the analytic curved box's metrics.
"""
import os

import numpy as np
import pytest

from galaexi_amd import basis, h5
from galaexi_amd import mesh as M

pytestmark = pytest.mark.skipif(not h5.available(), reason="no HDF5 C library on this machine")


@pytest.fixture(scope="module")
def tgv_mesh_path(golden_dir):
    return os.path.join(golden_dir, "CART_HEX_PERIODIC_008_mesh.h5")


def _read(path, N):
    from galaexi_amd import meshfile
    return meshfile.read_mesh(path, N)


def _centroid_perm(a, b):
    """perm with a.elem[e] == b.elem[perm[e]] by centroid"""
    from scipy.spatial import cKDTree
    ca = a.Elem_xGP.mean(axis=(1, 2, 3)).T
    cb = b.Elem_xGP.mean(axis=(1, 2, 3)).T
    d, perm = cKDTree(cb).query(ca)
    assert d.max() < 1e-10 and len(set(perm.tolist())) == a.nElems
    return perm


def test_h5_roundtrip(tmp_path):
    p = str(tmp_path / "t.h5")
    a = np.arange(24, dtype=np.float64).reshape(2, 3, 4)
    i = np.arange(6, dtype=np.int32).reshape(3, 2)
    with h5.File(p, "w") as f:
        f.write("A", a)
        f.write("I", i)
        f.write_attr("n", 7)
        f.write_attr("x", 2.5)
        f.write_attr("s", "hello")
        f.write_attr("names", ["ab", "cde"], strlen=255)
    with h5.File(p) as f:
        assert np.array_equal(f.read("A"), a) and f.read("A").dtype == np.float64
        assert np.array_equal(f.read("I"), i) and f.read("I").dtype == np.int32
        assert f.attr("n") == 7 and f.attr("x") == 2.5 and f.attr("s") == "hello"
        assert list(f.attr("names")) == ["ab", "cde"]
        assert f.has("A") and not f.has("B") and f.has_attr("n") and not f.has_attr("m")
        with pytest.raises(h5.H5Error):
            f.read("B")
    with pytest.raises(h5.H5Error):
        h5.File(str(tmp_path / "missing.h5"))


def test_file_mesh_topology(tgv_mesh_path):
    m = _read(tgv_mesh_path, 3)
    assert (m.nElems, m.nSides) == (512, 1536)
    e2s, s2e = m.ElemToSide, m.SideToElem
    assert sorted(set(e2s[0].ravel().tolist())) == list(range(1, m.nSides + 1))
    assert np.all((e2s[1] >= 0) & (e2s[1] <= 4))
    assert np.array_equal(e2s[2], (e2s[1] == 0).astype(e2s.dtype))
    for s in range(m.nSides):
        em, en, lm, ln, fl = s2e[:, s]
        assert e2s[0, lm - 1, em - 1] == s + 1 and e2s[1, lm - 1, em - 1] == 0
        assert e2s[0, ln - 1, en - 1] == s + 1 and e2s[1, ln - 1, en - 1] == fl and fl > 0


def test_file_mesh_faces_coincide(tgv_mesh_path):
    """Face node positions from both sides agree modulo the 2pi period (pins flips, S2V2 and pairing)."""
    N = 3
    m = _read(tgv_mesh_path, N)
    L = 2 * np.pi
    x = m.Elem_xGP
    worst = 0.0
    for s in range(m.nSides):
        em, en, lm, ln, fl = (int(v) for v in m.SideToElem[:, s])
        for q in range(N + 1):
            for p in range(N + 1):
                am, bm = m.S2V2[:, p, q, 0, lm - 1]
                an, bn = m.S2V2[:, p, q, fl, ln - 1]
                xm = x[(slice(None),) + M.face_to_vol(lm, am, bm, N) + (em - 1,)]
                xn = x[(slice(None),) + M.face_to_vol(ln, an, bn, N) + (en - 1,)]
                d = (xm - xn) / L
                worst = max(worst, float(np.abs(d - np.round(d)).max()))
    assert worst < 1e-12, worst


def test_file_mesh_matches_analytic_box(tgv_mesh_path):
    N = 7
    m = _read(tgv_mesh_path, N)
    b = M.box_mesh((8, 8, 8), N, lo=(0, 0, 0), hi=(2 * np.pi,) * 3)
    perm = _centroid_perm(m, b)
    for k in ("Metrics_fTilde", "Metrics_gTilde", "Metrics_hTilde", "sJ", "Elem_xGP"):
        A, B = getattr(m, k), getattr(b, k)[..., perm]
        assert np.abs(A - B).max() <= 1e-12 * np.abs(B).max(), k
    # every face has unit normals and the analytic face area (h/2)^2 per reference area
    assert np.allclose(np.sum(m.NormVec ** 2, axis=0), 1.0, atol=1e-14)
    assert np.allclose(m.SurfElem, (np.pi / 8) ** 2, rtol=1e-12)


def test_oracle_rhs_file_mesh_equals_box(tgv_mesh_path):
    """Oracle Ut (split PI + LF, BR1, TGV state with noise) on the ingested mesh = on the analytic box, elements
    matched.  The file's corners carry 15 digits, each element's metrics are evaluated at its own coordinates,
    and each face's metrics come from its master element (the file's master/slave choice differs from the
    box's), so the geometry differs in the last bits; the energy flux (~320 at p0 = 71) times L_Hat*sJ turns
    that into ~1e-10 absolute at face nodes: relative L2 < 5e-12.  A wrong flip, S2V2 entry or pairing gives
    differences of order 1e-3 and more."""
    import oracle as O
    from galaexi_amd import testcases as T
    N = 3
    m = _read(tgv_mesh_path, N)
    b = M.box_mesh((8, 8, 8), N, lo=(0, 0, 0), hi=(2 * np.pi,) * 3)
    perm = _centroid_perm(m, b)
    kw = dict(split=True, riemann=O.RIEMANN_LF, kappa=1.4, R=71.42857, Pr=0.72, mu0=6.25e-4)
    Um = T.taylor_green(m.Elem_xGP, Ma=0.1, noise=1e-3, seed=3)
    Ub = np.empty_like(Um)
    Ub[..., perm] = Um
    Ut_m = O.Operator(m, **kw).rhs(Um)
    Ut_b = O.Operator(b, **kw).rhs(Ub)
    err = np.linalg.norm((Ut_m - Ut_b[..., perm]).ravel()) / np.linalg.norm(Ut_b.ravel())
    assert err < 5e-12, err


def test_file_mesh_curved_ngeo2(tgv_mesh_path, tmp_path):
    """NGeo=2 nodes (i fastest) written with the analytic deformation reproduce the curved analytic box."""
    N, NGeo, amp, ph = 4, 2, 0.05, (0.1, 0.4, 0.2)
    with h5.File(tgv_mesh_path) as f:
        attrs = {k: f.attr(k) for k in ("Ngeo", "nElems", "nSides", "nNodes", "nUniqueSides", "nUniqueNodes",
                                        "nBCs", "Version")}
        ei, si, nc = f.read("ElemInfo"), f.read("SideInfo"), f.read("NodeCoords")
        bct, bcn = f.read("BCType"), list(f.read("BCNames"))
    nE, nn = ei.shape[0], NGeo + 1
    corners = nc[ei[:, 4][:, None] + np.arange(8)[None, :]].reshape(nE, 2, 2, 2, 3)  # (e, k, j, i, comp)
    t = np.linspace(0.0, 1.0, nn)
    w = np.stack([1 - t, t], axis=1)  # (a, corner)
    X = np.einsum("ka,jb,ic,eabcd->ekjid", w, w, w, corners)
    span, lo = 2 * np.pi, 0.0
    xh = 2.0 * (X - lo) / span - 1.0
    disp = amp * np.prod(np.sin(np.pi * (xh + np.asarray(ph))), axis=-1)
    X = X + 0.5 * span * disp[..., None]
    ei2 = ei.copy()
    ei2[:, 0] = 208
    ei2[:, 4] = np.arange(nE) * nn ** 3
    ei2[:, 5] = ei2[:, 4] + nn ** 3
    p = str(tmp_path / "curved_mesh.h5")
    attrs.update(Ngeo=NGeo, nNodes=nE * nn ** 3)
    with h5.File(p, "w") as f:
        for k, v in attrs.items():
            f.write_attr(k, v)
        f.write("ElemInfo", ei2)
        f.write("SideInfo", si)
        f.write("NodeCoords", X.reshape(-1, 3))
        f.write("BCType", bct)
    m = _read(p, N)
    b = M.box_mesh((8, 8, 8), N, lo=(0, 0, 0), hi=(2 * np.pi,) * 3, NGeo=NGeo, deform_amp=amp, deform_phase=ph)
    perm = _centroid_perm(m, b)
    for k in ("Metrics_fTilde", "Metrics_gTilde", "Metrics_hTilde", "sJ", "Elem_xGP"):
        A, B = getattr(m, k), getattr(b, k)[..., perm]
        assert np.abs(A - B).max() <= 1e-12 * np.abs(B).max(), k
    assert bcn  # names travel with the file; the written copy omits them


def test_rejects_non_periodic_boundaries(golden_dir):
    from galaexi_amd import meshfile
    with pytest.raises(meshfile.MeshFileError, match="non-periodic"):
        meshfile.read_mesh(os.path.join(golden_dir, "cartbox3D_mesh.h5"), 3)


def test_reference_state_file(golden_dir):
    from galaexi_amd import statefile as S
    name = "cartbox_State_0000000.000001000.h5"
    assert S.state_file_name("cartbox", 1e-6) == name  # TIMESTAMP formatting
    s = S.read_state(os.path.join(golden_dir, name))
    assert s.U.shape == (5, 4, 4, 4, 8) and s.N == 3 and s.NodeType == "GAUSS" and s.Time == 1e-6
    assert s.MeshFile == "cartbox3D_mesh.h5" and s.Project_Name == "cartbox" and s.attrs["File_Type"] == "State"
    assert list(s.VarNames) == list(S.VAR_NAMES)
    # RefState (rho, u, v, w, p) = (1, 1, 1, 1, 1), kappa = 1.4 -> E = 1/0.4 + 1.5 = 4
    assert np.allclose(s.U, np.array([1, 1, 1, 1, 4.0])[:, None, None, None, None], rtol=0, atol=1e-12)
    # restart onto Gauss-Lobatto nodes of the same degree: the constant state is kept
    U, t = S.restart_state(os.path.join(golden_dir, name), 3, 8)
    assert t == 1e-6 and np.abs(U - s.U).max() < 1e-13


def test_state_roundtrip_and_layout(golden_dir, tmp_path):
    from galaexi_amd import statefile as S
    rng = np.random.default_rng(5)
    U = np.asfortranarray(rng.standard_normal((5, 6, 6, 6, 9)))
    p = S.write_state(str(tmp_path), "tgv", U, 5, "CART_HEX_PERIODIC_008_mesh.h5", 2.5, next_time=3.0)
    assert os.path.basename(p) == "tgv_State_0000002.500000000.h5"
    s = S.read_state(p)
    assert np.array_equal(s.U, U) and s.N == 5 and s.NodeType == basis.GAUSS_LOBATTO and s.Time == 2.5
    assert s.attrs["NextFile"] == "tgv_State_0000003.000000000.h5"
    with h5.File(p) as f:
        assert f.read("DG_Solution").shape == (9, 6, 6, 6, 5)  # the reference's on-disk dims
        mine = {k for k in ("Program", "File_Type", "Project_Name", "File_Version", "N", "Dimension", "Time",
                            "MeshFile", "NextFile", "NodeType", "VarNames", "TIME") if f.has_attr(k)}
    with h5.File(os.path.join(golden_dir, "cartbox_State_0000000.000001000.h5")) as f:
        ref = {k for k in mine if f.has_attr(k)}
    assert mine == ref == {"Program", "File_Type", "Project_Name", "File_Version", "N", "Dimension", "Time",
                           "MeshFile", "NextFile", "NodeType", "VarNames", "TIME"}


def test_restart_interpolation(tmp_path):
    """Restart from a Gauss N=3 file onto Gauss-Lobatto N=5: a degree-3 polynomial state is reproduced."""
    from galaexi_amd import statefile as S
    nR, N, nE = 3, 5, 2
    xg, _w, _wb = basis.nodes_and_weights(nR, basis.GAUSS)
    xl, _w, _wb = basis.nodes_and_weights(N, basis.GAUSS_LOBATTO)

    def poly(x, e):
        X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
        return np.stack([1 + 0.1 * X ** 3 * Y + e, 0.2 * Y * Z ** 2, Z ** 3 - X, 0.3 * X * Y * Z, 2 + X ** 2 * Z])

    U = np.asfortranarray(np.stack([poly(xg, e) for e in range(nE)], axis=-1))
    p = S.write_state(str(tmp_path), "p", U, nR, "m.h5", 0.0, node_type=basis.GAUSS)
    V, _t = S.restart_state(p, N, nE)
    ref = np.stack([poly(xl, e) for e in range(nE)], axis=-1)
    assert np.abs(V - ref).max() < 1e-13
    with pytest.raises(NotImplementedError):
        S.restart_state(p, 2, nE)
    with pytest.raises(ValueError):
        S.restart_state(p, N, nE + 1)
