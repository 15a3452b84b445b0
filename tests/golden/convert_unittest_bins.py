"""Convert the reference's unit-test golden files (Fortran unformatted sequential, little endian, 4-byte
record markers) into numpy fixtures committed next to this script.

Source: unitTests/*.bin of flexi-framework/galaexi (GPL-3.0, see the reference's LICENSE.md).  The files hold
only numbers (inputs and expected outputs of the reference's unit tests), no code.  Record layouts follow
unitTests/unittest.f90:58-100 (ReadInReferenceElementData), unitTests/ProlongToFace.f90,
unitTests/SurfInt.f90, unitTests/DerivativeMatrix.f90 and unitTests/NodesAndWeights.f90.

Run:  python tests/golden/convert_unittest_bins.py [/root/reference]
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def records(path):
    raw = open(path, "rb").read()
    out, pos = [], 0
    while pos < len(raw):
        n = int(np.frombuffer(raw, dtype="<i4", count=1, offset=pos)[0])
        out.append(raw[pos + 4: pos + 4 + n])
        n2 = int(np.frombuffer(raw, dtype="<i4", count=1, offset=pos + 4 + n)[0])
        assert n == n2, path
        pos += 8 + n
    return out


def farr(buf, dtype, shape, offset=0):
    cnt = int(np.prod(shape))
    a = np.frombuffer(buf, dtype=dtype, count=cnt, offset=offset)
    return np.asfortranarray(a.reshape(shape, order="F")), offset + cnt * np.dtype(dtype).itemsize


def main(ref):
    ut = os.path.join(ref, "unitTests")
    NRef, n = 9, 10
    # UnittestElementData3D.bin: nElems,SideToElem(5,6),firstMPISide_YOUR,lastMPISide_MINE,nSides,
    #   S2V2(2,0:9,0:9,0:4,6),L_Minus,L_Plus,L_HatPlus,L_HatMinus (10 each),sJ(0:9,0:9,0:9,0:1,1)
    (r,) = records(os.path.join(ut, "UnittestElementData3D.bin"))
    o = 0
    nElems, o = farr(r, "<i4", (1,), o)
    s2e, o = farr(r, "<i4", (5, 6), o)
    fy, o = farr(r, "<i4", (1,), o)
    lm, o = farr(r, "<i4", (1,), o)
    nSides, o = farr(r, "<i4", (1,), o)
    s2v2, o = farr(r, "<i4", (2, n, n, 5, 6), o)
    L_Minus, o = farr(r, "<f8", (n,), o)
    L_Plus, o = farr(r, "<f8", (n,), o)
    L_HatPlus, o = farr(r, "<f8", (n,), o)
    L_HatMinus, o = farr(r, "<f8", (n,), o)
    sJ, o = farr(r, "<f8", (n, n, n, 2, 1), o)
    assert o == len(r)
    (uvol,) = records(os.path.join(ut, "ProlongToFaceUvol.bin"))
    Uvol, _ = farr(uvol, "<f8", (n, n, n, 1))
    ptf = records(os.path.join(ut, "ProlongToFace_GL3D.bin"))
    ptf_dg, _ = farr(ptf[0], "<f8", (n, n, 6))
    ptf_fv, _ = farr(ptf[1], "<f8", (n, n, 6))
    (fl,) = records(os.path.join(ut, "SurfIntFlux.bin"))
    Flux, _ = farr(fl, "<f8", (n, n, 6))
    si = records(os.path.join(ut, "SurfInt_GL3D.bin"))
    si_dg, _ = farr(si[0], "<f8", (n, n, n, 1))
    si_emm = records(os.path.join(ut, "SurfInt_GL3D_EMM.bin"))
    si_emm_dg, _ = farr(si_emm[0], "<f8", (n, n, n, 1))
    (dm,) = records(os.path.join(ut, "DerivativeMatrix.bin"))
    D, _ = farr(dm, "<f8", (11, 11, 40))
    (nw,) = records(os.path.join(ut, "NodesAndWeights.bin"))
    o = 0
    xi, o = farr(nw, "<f8", (11, 10, 4), o)
    w, o = farr(nw, "<f8", (11, 10, 4), o)
    wb, o = farr(nw, "<f8", (11, 10, 4), o)
    np.savez_compressed(
        os.path.join(HERE, "unittest_reference_element.npz"),
        NRef=NRef, nElems=nElems, SideToElem=s2e, firstMPISide_YOUR=fy, lastMPISide_MINE=lm, nSides=nSides,
        S2V2=s2v2, L_Minus=L_Minus, L_Plus=L_Plus, L_HatPlus=L_HatPlus, L_HatMinus=L_HatMinus, sJ=sJ)
    np.savez_compressed(os.path.join(HERE, "unittest_prolongtoface_gl3d.npz"), Uvol=Uvol, Uface_master_dg=ptf_dg,
                        Uface_master_fv=ptf_fv)
    np.savez_compressed(os.path.join(HERE, "unittest_surfint_gl3d.npz"), Flux=Flux, Ut_dg=si_dg,
                        Ut_emm_dg=si_emm_dg)
    np.savez_compressed(os.path.join(HERE, "unittest_basis.npz"), D=D, xi=xi, w=w, wBary=wb)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
