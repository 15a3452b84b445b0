"""Copy the reference's HDF5 mesh and state fixtures into tests/golden/ (data files of the reference's
own regression checks, GPL-3.0 reference repository), and print their SHA-256.

  CART_HEX_PERIODIC_008_mesh.h5        regressioncheck/checks/tgv/split/ (the TGV regression case's mesh:
                                       8^3 periodic box [0,2pi]^3, NGeo=1, HOPR output)
  cartbox3D_mesh.h5                    regressioncheck/checks/timedisc/freestream_3D/ (2^3 box with
                                       non-periodic boundaries, BC type 2)
  cartbox_State_0000000.000001000.h5   regressioncheck/checks/timedisc/freestream_3D/ (a DG_Solution state file,
                                       N=3, written by the reference at t=1e-6 for that mesh); its 4 KB user
                                       block (the writer's parameter file and build record, which no test reads)
                                       is removed, the HDF5 data from the superblock on are kept byte for byte
  CART_HEX_PERIODIC_0{02,04,08,16}_3D_mesh.h5
                                       regressioncheck/checks/convtest/h_3D/ (the h- and p-convergence meshes:
                                       periodic [-1,1]^3 in two HOPR zones of 1x2x2 .. 8x16x16 elements, the second
                                       zone built from mirrored corners; p_3D's 2^3 mesh is the same file)

Usage: python tests/golden/copy_hdf5_fixtures.py [/path/to/reference]
"""
import hashlib
import os
import shutil
import sys

FILES = ("regressioncheck/checks/tgv/split/CART_HEX_PERIODIC_008_mesh.h5",
         "regressioncheck/checks/timedisc/freestream_3D/cartbox3D_mesh.h5",
         "regressioncheck/checks/timedisc/freestream_3D/cartbox_State_0000000.000001000.h5",
         "regressioncheck/checks/convtest/h_3D/CART_HEX_PERIODIC_002_3D_mesh.h5",
         "regressioncheck/checks/convtest/h_3D/CART_HEX_PERIODIC_004_3D_mesh.h5",
         "regressioncheck/checks/convtest/h_3D/CART_HEX_PERIODIC_008_3D_mesh.h5",
         "regressioncheck/checks/convtest/h_3D/CART_HEX_PERIODIC_016_3D_mesh.h5")


SIGNATURE = b"\x89HDF\r\n\x1a\n"


def strip_user_block(path):
    """Remove an HDF5 user block: the superblock starts at the user block's size (a power of two from 512) and
    everything from there on is kept byte for byte, as HDF5's h5unjam writes it."""
    with open(path, "rb") as f:
        data = f.read()
    if data[:8] == SIGNATURE:
        return
    off = 512
    while off < len(data) and data[off:off + 8] != SIGNATURE:
        off *= 2
    assert off < len(data), "no HDF5 superblock found"
    with open(path, "wb") as f:
        f.write(data[off:])


def main(ref_root):
    here = os.path.dirname(os.path.abspath(__file__))
    for rel in FILES:
        dst = os.path.join(here, os.path.basename(rel))
        shutil.copyfile(os.path.join(ref_root, rel), dst)
        if "_State_" in rel:
            strip_user_block(dst)
        with open(dst, "rb") as f:
            print(os.path.basename(rel), hashlib.sha256(f.read()).hexdigest())


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
