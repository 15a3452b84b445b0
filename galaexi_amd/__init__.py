"""galaexi_amd -- MI355X (gfx950) DGSEM right-hand-side engine for GALAEXI's hot path.

libgalx.so (HIP, C ABI in include/galx.h) holds the kernels; galaexi_amd.dg is the host-side mirror of the
reference interface over that ABI; mesh.py/basis.py produce periodic box geometry and testcases.py initial states.
"""
from . import basis, mesh, testcases  # noqa: F401
