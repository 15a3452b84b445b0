"""Host-side geometry producer for periodic hexahedral box meshes.

The reference host reads an HDF5 mesh and builds connectivity and metrics (src/mesh/mesh.f90:91-451,
src/mesh/prepare_mesh.f90:688-760, src/mesh/mappings.f90:75-495, src/mesh/metrics.f90:171-793).  This module
produces the same arrays, in the same Fortran layouts, for structured periodic boxes with an optional
smooth periodic deformation and optional per-element rotations of the local frame:

  ElemToSide(3,6,nElems)   [SideID, flip, isMaster], 1-based IDs        (src/flexi.h:113-115)
  SideToElem(5,nSides)     [Elem, nbElem, locSide, nbLocSide, flip]      (src/flexi.h:106-110)
  S2V2(2,0:N,0:N,0:4,6)                                                  (src/mesh/mappings.f90:75-235)
  Metrics_{f,g,h}Tilde(3,0:N,0:N,0:N,nElems), sJ(0:N,0:N,0:N,nElems)      (curl form, metrics.f90:171-610)
  NormVec/TangVec1/TangVec2(3,0:N,0:N,nSides), SurfElem(0:N,0:N,nSides)   (metrics.f90:611-793)
  Elem_xGP(3,0:N,0:N,0:N,nElems)

All arrays are numpy arrays in Fortran order indexed exactly like the reference arrays (0-based).
"""
from __future__ import annotations

import itertools
import os
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import basis

XI_MINUS, XI_PLUS, ETA_MINUS, ETA_PLUS, ZETA_MINUS, ZETA_PLUS = 5, 3, 2, 4, 1, 6
# local outward direction (axis, sign) for locSide 1..6
_LOCSIDE_DIR = {ZETA_MINUS: (2, -1), ETA_MINUS: (1, -1), XI_PLUS: (0, 1), ETA_PLUS: (1, 1),
                XI_MINUS: (0, -1), ZETA_PLUS: (2, 1)}
NORMAL_DIRS = (3, 2, 1, 2, 1, 3)      # src/mesh/mesh_vars.f90:112
TANG_DIRS = (1, 3, 2, 3, 2, 1)        # src/mesh/mesh_vars.f90:114
NORMAL_SIGNS = (-1.0, -1.0, 1.0, 1.0, -1.0, 1.0)  # src/mesh/mesh_vars.f90:118


def flip_s2m(N, p, q, flip):
    """Flip_S2M (src/mesh/mappings.f90:246-266)."""
    return [(p, q), (q, p), (N - p, q), (N - q, N - p), (p, N - q)][flip]


def cgns_side_to_vol2(N, p, q, loc_side):
    """CGNS_SideToVol2 (src/mesh/mappings.f90:383-408)."""
    return {XI_MINUS: (q, p), XI_PLUS: (p, q), ETA_MINUS: (p, q), ETA_PLUS: (N - p, q),
            ZETA_MINUS: (q, p), ZETA_PLUS: (p, q)}[loc_side]


def build_s2v2(N):
    """S2V2(2,0:N,0:N,0:4,6) = CGNS_SideToVol2(Flip_M2S(p,q,flip)) (src/mesh/mappings.f90:144-150,472-490)."""
    n = N + 1
    s2v2 = np.zeros((2, n, n, 5, 6), dtype=np.int32, order="F")
    for ls in range(1, 7):
        for f in range(5):
            for q in range(n):
                for p in range(n):
                    a, b = cgns_side_to_vol2(N, *flip_s2m(N, p, q, f), ls)
                    s2v2[0, p, q, f, ls - 1] = a
                    s2v2[1, p, q, f, ls - 1] = b
    return s2v2


def face_to_vol(loc_side, a, b, N):
    """Volume node of face node (a,b) given in the volume frame of loc_side (GL trace)."""
    return {XI_MINUS: (0, a, b), XI_PLUS: (N, a, b), ETA_MINUS: (a, 0, b), ETA_PLUS: (a, N, b),
            ZETA_MINUS: (a, b, 0), ZETA_PLUS: (a, b, N)}[loc_side]


def proper_rotations():
    """The 24 signed permutation matrices with det=+1 (element frame orientations)."""
    mats = []
    for perm in itertools.permutations(range(3)):
        for signs in itertools.product((1, -1), repeat=3):
            m = np.zeros((3, 3), dtype=np.int64)
            for c in range(3):
                m[perm[c], c] = signs[c]
            if round(np.linalg.det(m)) == 1:
                mats.append(m)
    return mats


_ROTS = proper_rotations()


def _find_flip(N, s2v2, ls_m, rot_m, ls_s, rot_s):
    """Orientation of the slave face relative to the master face, found by matching node positions
    on the integer node lattice (the reference reads it from the mesh file)."""
    n = N + 1
    axis, sgn = _LOCSIDE_DIR[ls_m]
    loc = np.zeros(3, dtype=np.int64)
    loc[axis] = sgn
    gnorm = _ROTS[rot_m] @ loc
    cell_s = gnorm  # slave element sits in the neighbouring cell

    def lattice(cell, rot, ls, f):
        pts = np.zeros((n, n, 3), dtype=np.int64)
        for q in range(n):
            for p in range(n):
                a, b = s2v2[0, p, q, f, ls - 1], s2v2[1, p, q, f, ls - 1]
                ijk = np.array(face_to_vol(ls, a, b, N), dtype=np.int64)
                pts[p, q] = 2 * N * cell + _ROTS[rot] @ (2 * ijk - N)
        return pts

    ref = lattice(np.zeros(3, dtype=np.int64), rot_m, ls_m, 0)
    for f in range(1, 5):
        if np.array_equal(ref, lattice(cell_s, rot_s, ls_s, f)):
            return f
    raise RuntimeError("no matching flip between faces")


@dataclass
class Mesh:
    N: int
    nElems: int
    nSides: int
    ElemToSide: np.ndarray
    SideToElem: np.ndarray
    S2V2: np.ndarray
    Metrics_fTilde: np.ndarray
    Metrics_gTilde: np.ndarray
    Metrics_hTilde: np.ndarray
    sJ: np.ndarray
    NormVec: np.ndarray
    TangVec1: np.ndarray
    TangVec2: np.ndarray
    SurfElem: np.ndarray
    Elem_xGP: np.ndarray
    lo: tuple = (0.0, 0.0, 0.0)
    hi: tuple = (1.0, 1.0, 1.0)
    info: dict = field(default_factory=dict)

    @property
    def nDOF(self):
        return self.nElems * (self.N + 1) ** 3


def _along_i(V, X):
    """Apply the 1D operator V(a,i) along the third-last axis of X(..., i, j, k)."""
    ni, nj, nk = X.shape[-3:]
    return np.matmul(V, X.reshape(X.shape[:-3] + (ni, nj * nk))).reshape(X.shape[:-3] + (V.shape[0], nj, nk))


def _along_j(V, X):
    return np.matmul(V, X)


def _along_k(V, X):
    return np.matmul(X, V.T)


def _tensor3(V1, V2, V3, X):
    """Apply 1D operators along the last three axes of X(..., i, j, k)."""
    return _along_i(V1, _along_j(V2, _along_k(V3, X)))


def _surf_metrics_from_ja(normal_dir, tang_dir, normal_sign, Ja):
    """SurfMetricsFromJa (src/mesh/metrics.f90:760-793); Ja(dd,comp,p,q,...)"""
    jn = Ja[normal_dir - 1]
    se = np.sqrt(np.sum(jn * jn, axis=0))
    nv = normal_sign * jn / se
    jt = Ja[tang_dir - 1]
    t1 = jt - np.sum(jt * nv, axis=0) * nv
    t1 = t1 / np.sqrt(np.sum(t1 * t1, axis=0))
    t2 = np.cross(nv, t1, axis=0)
    return nv, t1, t2, se


class _Geometry:
    """Volume and surface metrics from the equidistant NGeo element nodes (metrics.f90:171-793): the nodes are
    interpolated to Chebyshev-Lobatto points, the Jacobian determinant is evaluated on GL points of degree
    3*NGeo and projected to N, the metric terms use the curl form on the CL points of degree N and are
    interpolated to the GL solution points."""

    def __init__(self, N, NGeo):
        self.N, self.NGeo = N, NGeo
        self.xEQ, _w, _wb = basis.nodes_and_weights(NGeo, basis.VISU)
        xCLg, _w, _wb = basis.nodes_and_weights(NGeo, basis.CHEBYSHEV_LOBATTO)
        xCLn, _w, _wb = basis.nodes_and_weights(N, basis.CHEBYSHEV_LOBATTO)
        self.xGL, _w, _wb = basis.nodes_and_weights(N, basis.GAUSS_LOBATTO)
        NGeoRef = 3 * NGeo
        self.V_EQ_CL = basis.vandermonde(NGeo, basis.VISU, NGeo, basis.CHEBYSHEV_LOBATTO)
        self.V_CLg_CLn = basis.vandermonde(NGeo, basis.CHEBYSHEV_LOBATTO, N, basis.CHEBYSHEV_LOBATTO)
        self.V_CLg_Ref = basis.vandermonde(NGeo, basis.CHEBYSHEV_LOBATTO, NGeoRef, basis.GAUSS_LOBATTO)
        self.V_Ref_N = basis.vandermonde(NGeoRef, basis.GAUSS_LOBATTO, N, basis.GAUSS_LOBATTO, modal=True)
        self.V_CLn_N = basis.vandermonde(N, basis.CHEBYSHEV_LOBATTO, N, basis.GAUSS_LOBATTO)
        self.V_CLg_N = basis.vandermonde(NGeo, basis.CHEBYSHEV_LOBATTO, N, basis.GAUSS_LOBATTO)
        self.DCLg = basis.derivative_matrix(xCLg)
        self.DCLn = basis.derivative_matrix(xCLn)
        self.s2v2 = build_s2v2(N)

    def volume(self, X):
        """X(e, comp, i, j, k) on the equidistant NGeo nodes -> Ja(e, dir, comp, i, j, k), sJ(e, i, j, k) on the GL
        points, JaCL(e, dir, comp, i, j, k) on the CL points of degree N, xGP(e, comp, i, j, k)."""
        N, NGeo = self.N, self.NGeo
        DCLg, DCLn = self.DCLg, self.DCLn
        XCLg = _tensor3(self.V_EQ_CL, self.V_EQ_CL, self.V_EQ_CL, X)
        XCLn = _tensor3(self.V_CLg_CLn, self.V_CLg_CLn, self.V_CLg_CLn, XCLg)
        dXg = np.stack([_along_i(DCLg, XCLg), _along_j(DCLg, XCLg), _along_k(DCLg, XCLg)], axis=1)  # (e,dir,comp,..)
        a = _tensor3(self.V_CLg_Ref, self.V_CLg_Ref, self.V_CLg_Ref, dXg)
        det = (a[:, 0, 0] * (a[:, 1, 1] * a[:, 2, 2] - a[:, 2, 1] * a[:, 1, 2]) +
               a[:, 1, 0] * (a[:, 2, 1] * a[:, 0, 2] - a[:, 0, 1] * a[:, 2, 2]) +
               a[:, 2, 0] * (a[:, 0, 1] * a[:, 1, 2] - a[:, 1, 1] * a[:, 0, 2]))
        detN = _tensor3(self.V_Ref_N, self.V_Ref_N, self.V_Ref_N, det)
        sJe = 1.0 / detN
        if N >= NGeo:
            dXn = _tensor3(self.V_CLg_CLn, self.V_CLg_CLn, self.V_CLg_CLn, dXg)
        else:
            dXn = np.stack([_along_i(DCLn, XCLn), _along_j(DCLn, XCLn), _along_k(DCLn, XCLn)], axis=1)
        # curl form (metrics.f90:450-484): Rc(e, dir, comp)
        Xc = XCLn[:, None]
        Rc = np.empty_like(dXn)
        Rc[:, :, 0] = 0.5 * (Xc[:, :, 2] * dXn[:, :, 1] - Xc[:, :, 1] * dXn[:, :, 2])
        Rc[:, :, 1] = 0.5 * (Xc[:, :, 0] * dXn[:, :, 2] - Xc[:, :, 2] * dXn[:, :, 0])
        Rc[:, :, 2] = 0.5 * (Xc[:, :, 1] * dXn[:, :, 0] - Xc[:, :, 0] * dXn[:, :, 1])
        JaCL = np.empty_like(dXn)
        JaCL[:, 0] = -_along_j(DCLn, Rc[:, 2]) + _along_k(DCLn, Rc[:, 1])
        JaCL[:, 1] = -_along_k(DCLn, Rc[:, 0]) + _along_i(DCLn, Rc[:, 2])
        JaCL[:, 2] = -_along_i(DCLn, Rc[:, 1]) + _along_j(DCLn, Rc[:, 0])
        Ja = _tensor3(self.V_CLn_N, self.V_CLn_N, self.V_CLn_N, JaCL)
        xg = _tensor3(self.V_CLg_N, self.V_CLg_N, self.V_CLg_N, XCLg)
        return Ja, sJe, JaCL, xg

    def face(self, JaCL_m, ls):
        """Ja on the master faces of local side ls in side (p,q) order (metrics.f90:700-740).
        JaCL_m(m, dir, comp, i, j, k) -> (dir, comp, p, q, m)"""
        N = self.N
        if ls == XI_MINUS:
            F = JaCL_m[:, :, :, 0, :, :]
        elif ls == XI_PLUS:
            F = JaCL_m[:, :, :, N, :, :]
        elif ls == ETA_MINUS:
            F = JaCL_m[:, :, :, :, 0, :]
        elif ls == ETA_PLUS:
            F = JaCL_m[:, :, :, :, N, :]
        elif ls == ZETA_MINUS:
            F = JaCL_m[:, :, :, :, :, 0]
        else:
            F = JaCL_m[:, :, :, :, :, N]
        F = np.matmul(np.matmul(self.V_CLn_N, F), self.V_CLn_N.T)  # (m, dir, comp, a, b)
        A, B = self.s2v2[0, :, :, 0, ls - 1], self.s2v2[1, :, :, 0, ls - 1]
        return np.moveaxis(F[..., A, B], 0, -1)


class _GeometryArrays:
    """The geometry arrays in the reference layouts (Fortran order)."""

    def __init__(self, N, nE, nS):
        n = N + 1
        self.mf = np.zeros((3, n, n, n, nE), order="F")
        self.mg = np.zeros((3, n, n, n, nE), order="F")
        self.mh = np.zeros((3, n, n, n, nE), order="F")
        self.sJ = np.zeros((n, n, n, nE), order="F")
        self.xGP = np.zeros((3, n, n, n, nE), order="F")
        self.nv = np.zeros((3, n, n, nS), order="F")
        self.t1 = np.zeros((3, n, n, nS), order="F")
        self.t2 = np.zeros((3, n, n, nS), order="F")
        self.se = np.zeros((n, n, nS), order="F")


def fill_geometry(geo, nE, s2e, nodes, chunk=256, threads=None):
    """Volume metrics of every element and surface metrics of every side, the latter from the side's master
    element (CalcSurfMetrics metrics.f90:611-740 skips flip /= 0).  nodes(elem_ids) returns the equidistant
    NGeo nodes X(e, comp, i, j, k) of a chunk of elements; chunks run on `threads` worker threads."""
    nS = s2e.shape[1]
    G = _GeometryArrays(geo.N, nE, nS)
    master_elem = s2e[0, :] - 1
    ls_of_side = s2e[2, :]

    def fill(c0):
        ids = np.arange(c0, min(nE, c0 + chunk))
        Ja, sJc, JaCL, xg = geo.volume(nodes(ids))
        G.mf[..., ids] = np.moveaxis(Ja[:, 0], 0, -1)
        G.mg[..., ids] = np.moveaxis(Ja[:, 1], 0, -1)
        G.mh[..., ids] = np.moveaxis(Ja[:, 2], 0, -1)
        G.sJ[..., ids] = np.moveaxis(sJc, 0, -1)
        G.xGP[..., ids] = np.moveaxis(xg, 0, -1)
        in_chunk = (master_elem >= c0) & (master_elem < c0 + len(ids))
        for ls in range(1, 7):
            sides = np.nonzero(in_chunk & (ls_of_side == ls))[0]
            if len(sides) == 0:
                continue
            JaF = geo.face(JaCL[master_elem[sides] - c0], ls)
            a, b, c, d = _surf_metrics_from_ja(NORMAL_DIRS[ls - 1], TANG_DIRS[ls - 1], NORMAL_SIGNS[ls - 1], JaF)
            G.nv[:, :, :, sides] = a
            G.t1[:, :, :, sides] = b
            G.t2[:, :, :, sides] = c
            G.se[:, :, sides] = d

    # chunks are independent and write disjoint slices; numpy releases the GIL in matmul and ufuncs
    workers = max(1, min(threads or os.cpu_count() or 1, 16, (nE + chunk - 1) // chunk))
    if workers == 1:
        for c0 in range(0, nE, chunk):
            fill(c0)
    else:
        with ThreadPoolExecutor(workers) as ex:
            list(ex.map(fill, range(0, nE, chunk)))
    return G


def box_mesh(nelems, N, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), deform_amp=0.0, deform_phase=(0.0, 0.0, 0.0),
             NGeo=None, rotations=None, seed=None, chunk=256, threads=None):
    """Periodic box of nelems=(nx,ny,nz) hexahedra with polynomial degree N.

    deform_amp: node displacement x_d += (hi-lo)_d/2 * amp * prod_e sin(pi*(xhat_e + phase_e)), xhat in [-1,1],
        applied at the equidistant NGeo geometry nodes (2-periodic in xhat, so periodicity is kept).
    NGeo: geometry degree (default 1 for affine boxes, N for deformed ones).
    rotations: None (all elements in the global frame), 'random' (seeded), or an int array of rotation indices.
    chunk, threads: elements per geometry batch and worker threads (default: all cores, at most 16).
    """
    nx, ny, nz = (int(v) for v in nelems)
    nE = nx * ny * nz
    n = N + 1
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    if NGeo is None:
        NGeo = N if deform_amp != 0.0 else 1
    rng = np.random.default_rng(seed)
    if rotations is None:
        rot = np.zeros(nE, dtype=np.int64)
    elif isinstance(rotations, str) and rotations == "random":
        rot = rng.integers(0, 24, size=nE)
    else:
        rot = np.asarray(rotations, dtype=np.int64)
    assert np.array_equal(_ROTS[0], np.eye(3, dtype=np.int64))  # index 0 is the global frame

    cells = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), axis=-1)
    cells = cells.transpose(2, 1, 0, 3).reshape(-1, 3)  # element ordering: x fastest
    ncell = np.array([nx, ny, nz])
    s2v2 = build_s2v2(N)

    # ---- connectivity (prepare_mesh.f90:688-760 semantics: first visitor of a face is its master) ----
    e2s = np.zeros((3, 6, nE), dtype=np.int32, order="F")
    side_key = {}
    s2e_list = []
    flip_cache = {}
    key_arr = np.zeros((nE, 6, 4), dtype=np.int64)
    for ls in range(1, 7):
        axis, sgn = _LOCSIDE_DIR[ls]
        loc = np.zeros(3, dtype=np.int64)
        loc[axis] = sgn
        g = np.stack([_ROTS[r] @ loc for r in range(24)])[rot]  # global outward dir per element
        gax = np.argmax(np.abs(g), axis=1)
        gsg = g[np.arange(nE), gax]
        c = cells.copy()
        shift = np.zeros_like(c)
        shift[np.arange(nE), gax] = (gsg < 0).astype(np.int64)
        c = (c - shift) % ncell
        key_arr[:, ls - 1, 0] = gax
        key_arr[:, ls - 1, 1:] = c
    for e in range(nE):
        for ls in range(1, 7):
            k = tuple(key_arr[e, ls - 1])
            if k not in side_key:
                sid = len(s2e_list) + 1
                side_key[k] = sid
                s2e_list.append([e + 1, -1, ls, -1, -1])
                e2s[:, ls - 1, e] = (sid, 0, 1)
            else:
                sid = side_key[k]
                rec = s2e_list[sid - 1]
                em, lsm = rec[0] - 1, rec[2]
                ck = (lsm, int(rot[em]), ls, int(rot[e]))
                if ck not in flip_cache:
                    flip_cache[ck] = _find_flip(N, s2v2, lsm, int(rot[em]), ls, int(rot[e]))
                f = flip_cache[ck]
                rec[1], rec[3], rec[4] = e + 1, ls, f
                e2s[:, ls - 1, e] = (sid, f, 0)
    nS = len(s2e_list)
    s2e = np.asfortranarray(np.array(s2e_list, dtype=np.int32).T)

    # ---- geometry (metrics.f90:171-610) ----
    geo = _Geometry(N, NGeo)
    h = (hi - lo) / ncell
    uniform = deform_amp == 0.0 and np.all(rot == 0)
    xEQ = geo.xEQ

    def nodes(elem_ids, centre_override=None):
        """Equidistant NGeo nodes X(e, comp, i, j, k) of a chunk of elements."""
        ne = len(elem_ids)
        xi = np.stack(np.meshgrid(xEQ, xEQ, xEQ, indexing="ij"), axis=0)  # (3, i, j, k)
        R = np.stack(_ROTS)[rot[elem_ids]].astype(np.float64)  # (ne,3,3)
        loc = np.einsum("eab,bijk->eaijk", R, xi)
        if centre_override is not None:
            centre = np.broadcast_to(np.asarray(centre_override, dtype=np.float64), (ne, 3))
        else:
            centre = lo[None, :] + (cells[elem_ids] + 0.5) * h[None, :]
        X = centre[:, :, None, None, None] + 0.5 * h[None, :, None, None, None] * loc  # (e, 3, i, j, k)
        if deform_amp != 0.0:
            span = (hi - lo)[None, :, None, None, None]
            xh = 2.0 * (X - lo[None, :, None, None, None]) / span - 1.0
            ph = np.asarray(deform_phase, dtype=np.float64)[None, :, None, None, None]
            disp = deform_amp * np.prod(np.sin(np.pi * (xh + ph)), axis=1)
            X = X + 0.5 * span * disp[:, None]
        return X

    if uniform:
        # every element has the metrics of element 0: compute one and broadcast
        G = _GeometryArrays(N, nE, nS)
        Ja1, sJ1, JaCL1, _x = geo.volume(nodes(np.array([0]), centre_override=np.zeros(3)))
        G.mf[:] = Ja1[0, 0][..., None]
        G.mg[:] = Ja1[0, 1][..., None]
        G.mh[:] = Ja1[0, 2][..., None]
        G.sJ[:] = sJ1[0][..., None]
        xGL = geo.xGL
        xi3 = np.stack(np.meshgrid(xGL, xGL, xGL, indexing="ij"), axis=0)
        centre = (lo[:, None] + (cells.T + 0.5) * h[:, None])
        G.xGP[:] = centre[:, None, None, None, :] + 0.5 * h[:, None, None, None, None] * xi3[..., None]
        ls_of_side = s2e[2, :]
        for ls in range(1, 7):
            ids = np.nonzero(ls_of_side == ls)[0]
            if len(ids) == 0:
                continue
            a, b, c, d = _surf_metrics_from_ja(NORMAL_DIRS[ls - 1], TANG_DIRS[ls - 1], NORMAL_SIGNS[ls - 1],
                                               geo.face(JaCL1, ls))
            G.nv[:, :, :, ids] = a
            G.t1[:, :, :, ids] = b
            G.t2[:, :, :, ids] = c
            G.se[:, :, ids] = d
    else:
        G = fill_geometry(geo, nE, s2e, nodes, chunk=chunk, threads=threads)
    mf, mg, mh, sJ, xGP, nv, t1, t2, se = G.mf, G.mg, G.mh, G.sJ, G.xGP, G.nv, G.t1, G.t2, G.se
    return Mesh(N=N, nElems=nE, nSides=nS, ElemToSide=e2s, SideToElem=s2e, S2V2=s2v2,
                Metrics_fTilde=mf, Metrics_gTilde=mg, Metrics_hTilde=mh, sJ=sJ, NormVec=nv, TangVec1=t1,
                TangVec2=t2, SurfElem=se, Elem_xGP=xGP, lo=tuple(lo), hi=tuple(hi),
                info=dict(NGeo=NGeo, deform_amp=deform_amp, deform_phase=tuple(deform_phase),
                          rotations=rot.copy(), uniform=uniform))
