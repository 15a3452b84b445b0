"""Mesh ingest from the reference's HDF5 mesh format (HOPR output).

Reads ElemInfo, SideInfo, NodeCoords and BCType (ReadMesh src/mesh/mesh_readin.f90:196-520) and produces the
arrays of galaexi_amd.mesh.Mesh in the reference layouts:

  * sides: a side is oriented (flip 0, the master candidate) when its SideInfo ID is positive, else its flip is
    MOD(SideInfo flip entry, 10) (mesh_readin.f90:361-371); the two sides of a face are paired by |ID|
    (mesh_readin.f90:396-458); periodic (BC type 1) and inner-dummy (type 100) boundary sides are connected like
    inner sides;
  * SideIDs are handed out in element / local-side order, a face getting the next ID on its first visit
    (prepare_mesh.f90:187-237, no BC, mortar or MPI sides here); ElemToSide, SideToElem and E2S_IS_MASTER follow
    fillMeshInfo (prepare_mesh.f90:688-740) and mesh.f90:415-431;
  * geometry from the equidistant NGeo element nodes with the same metric code as the analytic meshes
    (CalcMetrics / CalcSurfMetrics, src/mesh/metrics.f90:171-793).

The hot path handles conforming hexahedral meshes whose faces all connect two elements: files with non-periodic
boundary sides, mortar (non-conforming) sides or other element types are rejected with a clear error.
"""
from __future__ import annotations

import os

import numpy as np

from . import h5
from .mesh import Mesh, _Geometry, build_s2v2, fill_geometry

# ElemInfo / SideInfo columns (mesh_readin.f90:33-50), 0-based here
ELEM_TYPE, ELEM_ZONE, ELEM_FIRST_SIDE, ELEM_LAST_SIDE, ELEM_FIRST_NODE, ELEM_LAST_NODE = range(6)
SIDE_TYPE, SIDE_ID, SIDE_NB_ELEM, SIDE_FLIP, SIDE_BCID = range(5)
HEXA_TYPES = (108, 118, 208)
BC_PERIODIC, BC_INNER_DUMMY = 1, 100


class MeshFileError(ValueError):
    pass


def _connectivity(elem_info, side_info, bc_type):
    """ElemToSide(3,6,nElems), SideToElem(5,nSides) (1-based IDs, Fortran order) from ElemInfo/SideInfo."""
    nE = elem_info.shape[0]
    bad = [t for t in np.unique(elem_info[:, ELEM_TYPE]) if t not in HEXA_TYPES]
    if bad:
        raise MeshFileError(f"element types {bad} are not hexahedra; the DG operator here handles hexahedra only")
    nloc = elem_info[:, ELEM_LAST_SIDE] - elem_info[:, ELEM_FIRST_SIDE]
    if np.any(nloc != 6):
        raise MeshFileError("elements with more than 6 SideInfo rows (mortar sides) are not supported")
    rows = elem_info[:, ELEM_FIRST_SIDE][:, None] + np.arange(6)[None, :]  # (e, ls-1) -> SideInfo row
    S = side_info[rows]  # (e, 6, 5)
    if np.any(S[..., SIDE_TYPE] < 0) or np.any(S[..., SIDE_NB_ELEM] < 0):
        raise MeshFileError("mortar (non-conforming) sides are not supported")
    bcid = S[..., SIDE_BCID]
    if np.any(bcid != 0):
        types = np.zeros(len(bc_type) + 1, dtype=np.int64)
        types[1:] = bc_type[:, 0]
        t = types[bcid]
        if np.any((bcid != 0) & (t != BC_PERIODIC) & (t != BC_INNER_DUMMY)):
            raise MeshFileError("the mesh has non-periodic boundary sides; the operator here covers meshes whose "
                                "faces all connect two elements (periodic or inner)")
    if np.any(S[..., SIDE_NB_ELEM] == 0):
        raise MeshFileError("a side without neighbour element that is not a boundary side")
    ind = np.abs(S[..., SIDE_ID]).astype(np.int64)
    oriented = S[..., SIDE_ID] > 0
    flip = np.where(oriented, 0, S[..., SIDE_FLIP] % 10)
    if np.any((flip < 0) | (flip > 4)):
        raise MeshFileError("side flip outside 0..4")
    nb = S[..., SIDE_NB_ELEM] - 1
    # the partner of (e, ls): the local side of the neighbour element with the same |ID|
    partner = np.full((nE, 6), -1, dtype=np.int64)
    for ls in range(6):
        hit = ind[nb[:, ls]] == ind[:, ls][:, None]  # (e, 6): sides of nb(e, ls) with the same |ID|
        # a face connecting an element to itself (periodic, one element across) matches twice: take the other
        self_pair = nb[:, ls] == np.arange(nE)
        if np.any(self_pair):
            hit[self_pair, ls] = False
        if np.any(hit.sum(axis=1) != 1):
            raise MeshFileError("sides with the same index do not pair up")
        partner[:, ls] = np.argmax(hit, axis=1)
    e_idx = np.arange(nE)
    if np.any(partner[nb, partner] != np.arange(6)[None, :]) or np.any(nb[nb, partner] != e_idx[:, None]):
        raise MeshFileError("side pairing is not symmetric")
    if np.any((flip == 0) == (flip[nb, partner] == 0)):
        raise MeshFileError("a face without exactly one oriented (flip 0) side")
    # SideIDs in element / local-side order (first visit assigns)
    side_id = np.zeros((nE, 6), dtype=np.int64)
    nS = 0
    for e in range(nE):
        for ls in range(6):
            if side_id[e, ls] == 0:
                nS += 1
                side_id[e, ls] = nS
                side_id[nb[e, ls], partner[e, ls]] = nS
    e2s = np.zeros((3, 6, nE), dtype=np.int32, order="F")
    s2e = np.full((5, nS), -1, dtype=np.int32, order="F")
    e2s[0] = side_id.T
    e2s[1] = flip.T
    e2s[2] = (flip == 0).T
    sid = side_id - 1
    m = flip == 0
    s2e[0, sid[m]] = (np.nonzero(m)[0] + 1)
    s2e[2, sid[m]] = (np.nonzero(m)[1] + 1)
    s = ~m
    s2e[1, sid[s]] = (np.nonzero(s)[0] + 1)
    s2e[3, sid[s]] = (np.nonzero(s)[1] + 1)
    s2e[4, sid[s]] = flip[s]
    if np.any(s2e[:4] < 0):
        raise MeshFileError("a face without exactly one master (flip 0) and one slave side")
    return e2s, s2e


def read_mesh(path, N, chunk=256, threads=None):
    """Mesh (galaexi_amd.mesh.Mesh) of polynomial degree N from a HOPR HDF5 mesh file."""
    with h5.File(path, "r") as f:
        NGeo = int(f.attr("Ngeo"))
        nE = int(f.attr("nElems"))
        elem_info = np.asarray(f.read("ElemInfo"), dtype=np.int64)
        side_info = np.asarray(f.read("SideInfo"), dtype=np.int64)
        node_coords = np.asarray(f.read("NodeCoords"), dtype=np.float64)
        bc_type = np.asarray(f.read("BCType"), dtype=np.int64).reshape(-1, 4) if f.has("BCType") \
            else np.zeros((0, 4), dtype=np.int64)
        bc_names = list(f.read("BCNames")) if f.has("BCNames") else []
    if elem_info.shape != (nE, 6) or side_info.ndim != 2 or side_info.shape[1] != 5 or node_coords.shape[1] != 3:
        raise MeshFileError("unexpected ElemInfo/SideInfo/NodeCoords shapes")
    nn = NGeo + 1
    if np.any(elem_info[:, ELEM_LAST_NODE] - elem_info[:, ELEM_FIRST_NODE] != nn ** 3):
        raise MeshFileError(f"elements do not carry (NGeo+1)^3 = {nn ** 3} nodes")
    e2s, s2e = _connectivity(elem_info, side_info, bc_type)
    first = elem_info[:, ELEM_FIRST_NODE]

    def nodes(elem_ids):
        # NodeCoords(3, 0:NGeo, 0:NGeo, 0:NGeo, nElems) with i fastest (mesh_readin.f90:470-473)
        r = first[elem_ids][:, None] + np.arange(nn ** 3)[None, :]
        X = node_coords[r].reshape(len(elem_ids), nn, nn, nn, 3)  # (e, k, j, i, comp)
        return np.ascontiguousarray(X.transpose(0, 4, 3, 2, 1))  # (e, comp, i, j, k)

    geo = _Geometry(N, NGeo)
    G = fill_geometry(geo, nE, s2e, nodes, chunk=chunk, threads=threads)
    lo = tuple(float(v) for v in node_coords.min(axis=0))
    hi = tuple(float(v) for v in node_coords.max(axis=0))
    return Mesh(N=N, nElems=nE, nSides=s2e.shape[1], ElemToSide=e2s, SideToElem=s2e, S2V2=build_s2v2(N),
                Metrics_fTilde=G.mf, Metrics_gTilde=G.mg, Metrics_hTilde=G.mh, sJ=G.sJ, NormVec=G.nv,
                TangVec1=G.t1, TangVec2=G.t2, SurfElem=G.se, Elem_xGP=G.xGP, lo=lo, hi=hi,
                info=dict(NGeo=NGeo, source=os.path.basename(str(path)), BCNames=bc_names,
                          ElemBarycentersGP=G.xGP.mean(axis=(1, 2, 3)).T.copy()))
