"""Convert the reference's TGV regression fixture into tests/golden/tgv_re1600_split_reference.npz.

Source: regressioncheck/checks/tgv/split/TGV_Re1600_Split_TGVAnalysis_Reference.csv of the reference repository
(GPL-3.0), written by AnalyzeTestcase (src/testcase/taylorgreenvortex/testcase.f90:283-515) for the case of
regressioncheck/checks/tgv/split/parameter.ini: 8^3 periodic box [0,2pi]^3, N=7 Gauss-Lobatto, split form PI,
RoeEntropyFix, BR1, mu0=6.25e-4, R=71.42857, Ma 0.1, CFLScale=DFLScale=0.8, tend=13, compared at relative 1e-4
(analyze.ini).  Usage: python tests/golden/convert_tgv_csv.py /path/to/reference
"""
import os
import sys

import numpy as np

REL = "regressioncheck/checks/tgv/split/TGV_Re1600_Split_TGVAnalysis_Reference.csv"


def main(ref_root):
    src = os.path.join(ref_root, REL)
    with open(src) as f:
        header = [h.strip() for h in f.readline().split(",")]
        rows = [[float(v.strip()) for v in line.split(",")] for line in f if line.strip()]
    data = np.array(rows, dtype=np.float64)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tgv_re1600_split_reference.npz")
    np.savez_compressed(out, columns=np.array(header), data=data, source=np.array(REL),
                        params=np.array("N=7 GL, 8^3 [0,2pi]^3, split PI, RoeEntropyFix, BR1, mu0=6.25e-4, "
                                        "R=71.42857, kappa=1.4, Pr=0.72, Ma=0.1, CFL=DFL=0.8, tend=13, "
                                        "nAnalyzeTestCase=10, tolerance relative 1e-4"))
    print(out, data.shape)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
