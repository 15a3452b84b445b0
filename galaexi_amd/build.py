"""Build libgalx.so (gfx950) in-tree with hipcc.

python -m galaexi_amd.build [--degrees 3,5,7] [--mixed-degrees 3,5,7]
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgalx.so")

# FP64: N <= 9 one volume node per thread, N = 10..13 the high-order kernels (the degrees of the reference's
# p-convergence check, regressioncheck/checks/convtest/p_3D); the discrete adjoint for N <= 8
DEFAULT_FP64 = tuple(range(1, 14))
DEFAULT_MIXED = (3, 5, 7)
DEFAULT_ADJ = tuple(range(1, 9))

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # no packed FP32 instructions (v_pk_*_f32) anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def sources():
    return [os.path.join(CSRC, "galx_api.hip")]


def deps():
    return [os.path.join(CSRC, f) for f in os.listdir(CSRC)] + [os.path.join(HERE, "..", "include", "galx.h")]


def build(degrees=DEFAULT_FP64, mixed_degrees=DEFAULT_MIXED, out=LIB, verbose=False, resource_report=None,
          adj_degrees=DEFAULT_ADJ):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    xs = " ".join(f"X({d})" for d in degrees)
    xm = " ".join(f"X({d})" for d in mixed_degrees)
    xa = " ".join(f"X({d})" for d in adj_degrees if d in degrees)
    cmd = [hipcc] + HIPCC_FLAGS + [f"-DGALX_DEGREES_FP64={xs}", f"-DGALX_DEGREES_MIXED={xm}",
                                   f"-DGALX_DEGREES_ADJ={xa}"]
    if resource_report:
        cmd.append("-Rpass-analysis=kernel-resource-usage")
    tmp = out + ".tmp"
    cmd += ["-o", tmp] + sources()
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if resource_report:
        with open(resource_report, "w") as f:
            f.write(res.stderr)
    if res.returncode != 0:
        sys.stderr.write(res.stderr[-20000:])
        raise RuntimeError("hipcc failed")
    os.replace(tmp, out)
    return out


FORTRAN_HOST_SRC = os.path.join(HERE, "..", "tests", "fortran", "galx_host.f90")
FORTRAN_MODULE = os.path.join(HERE, "..", "include", "galx_mod.f90")
FORTRAN_HOST = os.path.join(HERE, "..", "tests", "fortran", "galx_host")


def build_fortran_host(out=FORTRAN_HOST, verbose=False):
    """Compile the Fortran host test program against include/galx_mod.f90 and libgalx.so (rpath relative to the
    binary, so the tree can move).  Returns the binary's path, or None when no Fortran compiler is installed."""
    fc = os.environ.get("GALX_FC", "/opt/rocm/bin/amdflang")
    if not os.path.exists(fc):
        print(f"no Fortran compiler at {fc}: the Fortran host test program is not built", flush=True)
        return None
    import tempfile

    with tempfile.TemporaryDirectory() as mod_dir:
        cmd = [fc, "-O2", "-module-dir", mod_dir, "-o", out, FORTRAN_MODULE, FORTRAN_HOST_SRC,
               "-L" + HERE, "-lgalx", "-Wl,-rpath,$ORIGIN/../../galaexi_amd"]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout[-5000:] + res.stderr[-5000:])
        raise RuntimeError("Fortran host build failed")
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--degrees", default=",".join(map(str, DEFAULT_FP64)))
    ap.add_argument("--mixed-degrees", default=",".join(map(str, DEFAULT_MIXED)))
    ap.add_argument("--report", default=None, help="write the kernel resource-usage remarks to this file")
    a = ap.parse_args(argv)
    degs = tuple(int(x) for x in a.degrees.split(",") if x)
    mdegs = tuple(int(x) for x in a.mixed_degrees.split(",") if x)
    print(build(degs, mdegs, verbose=True, resource_report=a.report))


if __name__ == "__main__":
    main()
