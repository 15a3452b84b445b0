"""State files in the reference's HDF5 format: write (WriteState, src/io_hdf5/hdf5_output.f90:84-216 with
GenerateFileSkeleton :743-849, WriteHeader :939-961, MarkWriteSuccessfull :856-871) and restart (Restart,
src/restart/restart.f90:304-560, RestartMode 1: conservative variables from DG_Solution).

On disk: dataset DG_Solution with Fortran dims (nVar, N+1, N+1, N+1, nElems) (C-order dims reversed), root
attributes Program, File_Type, Project_Name, File_Version, N, Dimension, Time, MeshFile, NextFile, NodeType,
VarNames and TIME (the date of the write, as DATE_AND_TIME values).  File name
<ProjectName>_State_<Time as F17.9, blanks as 0>.h5 (TIMESTAMP, src/globals/globals.f90:455-482).

U arrays are in the reference layout U(5, 0:N, 0:N, 0:N, nElems) (Fortran order), as galx_set_state/get_state
take them.
"""
from __future__ import annotations

import datetime
import os

import numpy as np

from . import basis, h5

VAR_NAMES = ("Density", "MomentumX", "MomentumY", "MomentumZ", "EnergyStagnationDensity")  # StrVarNames
PROGRAM = "Flexi"
FILE_VERSION = 0.1


def timestamp(name, time):
    """TIMESTAMP (globals.f90:455-482): name_ + time written as F17.9 with the leading blanks as zeros."""
    return f"{name}_{f'{time:17.9f}'.replace(' ', '0')}"


def state_file_name(project, time):
    return timestamp(f"{project}_State", time) + ".h5"


def _date_values():
    """DATE_AND_TIME(VALUES=...): year, month, day, UTC offset in minutes, hour, minute, second, millisecond."""
    now = datetime.datetime.now().astimezone()
    off = now.utcoffset()
    return [now.year, now.month, now.day, int(off.total_seconds() // 60) if off else 0, now.hour, now.minute,
            now.second, now.microsecond // 1000]


def write_state(directory, project, U, N, mesh_file, time, next_time=None, node_type=basis.GAUSS_LOBATTO,
                var_names=VAR_NAMES):
    """Write U(5,0:N,0:N,0:N,nElems) as <directory>/<project>_State_<time>.h5; returns the path."""
    U = np.asarray(U, dtype=np.float64)
    n = N + 1
    if U.ndim != 5 or U.shape[1:4] != (n, n, n):
        raise ValueError(f"U must have shape (nVar, {n}, {n}, {n}, nElems), got {U.shape}")
    if len(var_names) != U.shape[0]:
        raise ValueError("one variable name per variable")
    path = os.path.join(directory, state_file_name(project, time))
    with h5.File(path, "w") as f:
        f.write("DG_Solution", np.asfortranarray(U).T)  # C-order dims (nElems, k, j, i, var)
        f.write_attr("VarNames", list(var_names), strlen=255)
        f.write_attr("Program", PROGRAM)
        f.write_attr("File_Type", "State")
        f.write_attr("Project_Name", project)
        f.write_attr("File_Version", float(FILE_VERSION))
        f.write_attr("N", int(N))
        f.write_attr("Dimension", 3)
        f.write_attr("Time", float(time))
        f.write_attr("MeshFile", str(mesh_file))
        f.write_attr("NextFile", state_file_name(project, time if next_time is None else next_time), strlen=255)
        f.write_attr("NodeType", node_type, strlen=255)
        f.write_attr("TIME", np.array(_date_values(), dtype=np.int32))
    return path


class StateFile:
    """Contents of a state file: U (Fortran layout), N, NodeType, Time, MeshFile, Project_Name, VarNames and the
    remaining root attributes in `attrs`."""

    def __init__(self, U, N, node_type, time, mesh_file, project, var_names, attrs):
        self.U, self.N, self.NodeType, self.Time = U, N, node_type, time
        self.MeshFile, self.Project_Name, self.VarNames, self.attrs = mesh_file, project, var_names, attrs


def read_state(path):
    with h5.File(path, "r") as f:
        if not f.has("DG_Solution"):
            raise h5.H5Error(f"{path}: no DG_Solution dataset")
        raw = f.read("DG_Solution")
        attrs = {}
        for name in ("N", "Time", "NodeType", "MeshFile", "Project_Name", "File_Type", "Program", "File_Version",
                     "Dimension", "NextFile"):
            if f.has_attr(name):
                attrs[name] = f.attr(name)
        var_names = [str(v) for v in np.atleast_1d(f.attr("VarNames"))] if f.has_attr("VarNames") else []
    U = np.asfortranarray(raw.T)  # (nVar, i, j, k, nElems)
    N = int(attrs.get("N", U.shape[1] - 1))
    return StateFile(U, N, str(attrs.get("NodeType", basis.GAUSS_LOBATTO)), float(attrs.get("Time", 0.0)),
                     str(attrs.get("MeshFile", "")), str(attrs.get("Project_Name", "")), var_names, attrs)


def restart_state(path, N, nElems, node_type=basis.GAUSS_LOBATTO):
    """U(5,0:N,0:N,0:N,nElems) on the computational grid from a state file and its time (Restart,
    restart.f90:304-560, RestartMode 1).  Same N and node type: the file's values.  Otherwise
    (InterpolateSolution) a restart degree N_Restart <= N is interpolated with the modal Vandermonde
    Vdm(N_Restart, NodeType_Restart -> N, NodeType) per element (restart.f90:529-537); N_Restart > N needs
    the Jacobian-weighted projection with the reference-element Jacobian of the mesh (restart.f90:514-528)
    and is not provided here."""
    s = read_state(path)
    U = s.U
    if U.shape[0] != 5:
        raise ValueError(f"{path}: {U.shape[0]} variables, the Navier-Stokes state has 5")
    if U.shape[4] != nElems:
        raise ValueError(f"{path}: {U.shape[4]} elements, the mesh has {nElems}")
    nR = U.shape[1] - 1
    if U.shape[2] - 1 != nR or U.shape[3] - 1 != nR:
        raise ValueError(f"{path}: DG_Solution dims {U.shape} are not a 3D tensor-product state")
    if nR == N and s.NodeType == node_type:
        return np.asfortranarray(U.copy()), s.Time
    if nR > N:
        raise NotImplementedError(f"restart from N={nR} to N={N} needs the Jacobian-weighted projection")
    V = basis.vandermonde(nR, s.NodeType, N, node_type, modal=True)
    out = np.einsum("ai,bj,ck,vijke->vabce", V, V, V, U, optimize=True)
    return np.asfortranarray(out), s.Time
