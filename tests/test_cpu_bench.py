"""The benchmark's host-side bookkeeping, on the CPU: the algorithmic bytes its rooflines divide by, the counter
traffic it reports, and its refusal of more than one GPU.

The algorithmic bytes are recounted here from an explicit list of the arrays each kernel reads and writes for one
N=7 element (512 volume nodes, 296 surface nodes, 6 x 64 face nodes), so that a dropped or doubled array in the
closed form of bench.alg_bytes fails.
"""
import json

import pytest

import bench

N = 7
NODES = (N + 1) ** 3                      # 512
SURF = NODES - (N - 1) ** 3               # 296 surface nodes
FACE = 6 * (N + 1) ** 2                   # 384 face nodes (edges and corners counted once per face)


def per_dof(arrays):
    """bytes per DOF of a list of (count per element, values per entry, bytes per value)"""
    return sum(c * v * b for c, v, b in arrays) / NODES


def test_alg_bytes_fp64_recounted():
    lift = per_dof([
        (NODES, 5, 8),   # U
        (SURF, 10, 8),   # metrics and sJ at the surface nodes (surface-node copy)
        (FACE, 4, 8),    # face normal and SurfElem
        (FACE, 5, 8),    # neighbour U trace records
        (FACE, 4, 8),    # own viscous normal-flux traces (written)
    ])
    rhs = per_dof([
        (NODES, 5, 8),   # U
        (NODES, 10, 8),  # metrics and sJ
        (FACE, 4, 8),    # face normal and SurfElem (split LF in Cartesian components)
        (FACE, 5, 8),    # neighbour U trace records
        (FACE, 4, 8),    # neighbour viscous normal-flux traces
        (NODES, 5, 8),   # LSERK register read
        (NODES, 5, 8),   # U written
        (NODES, 5, 8),   # LSERK register written
        (FACE, 5, 8),    # own next U trace records (written)
    ])
    assert (lift, rhs) == (164.25, 348.0)
    assert bench.alg_bytes(N, 8) == pytest.approx((lift, rhs), rel=0, abs=1e-12)


def test_alg_bytes_mixed_recounted():
    lift = per_dof([(NODES, 5, 8), (SURF, 10, 4), (FACE, 4, 4), (FACE, 5, 4), (FACE, 4, 4)])
    rhs = per_dof([
        (NODES, 5, 8), (NODES, 10, 4), (NODES, 1, 8),   # U; FP32 metrics and sJ; the FP64 sJ of the update
        (FACE, 4, 4), (FACE, 5, 4), (FACE, 4, 4),       # face geometry, neighbour records, neighbour F_v.n
        (NODES, 5, 8), (NODES, 5, 8), (NODES, 5, 8),    # register read, U and register written
        (FACE, 5, 4),                                    # own next records
    ])
    assert (lift, rhs) == (102.125, 262.0)
    assert bench.alg_bytes(N, 4, 4) == pytest.approx((lift, rhs), rel=0, abs=1e-12)


def test_roofline_fields():
    r = bench.roofline("k", 348.0, 1000, 1e-6, 4.2e5)
    assert r["achieved"] == pytest.approx(348.0 * 1000 / 1e-6 / 1e9)
    assert r["frac"] == pytest.approx(r["achieved"] / 8000.0)
    assert r["traffic_over_alg"] == pytest.approx(4.2e5 / 348e3)
    assert bench.roofline("k", 348.0, 1000, 1e-6, None)["traffic_over_alg"] is None


def test_traffic_file_matches_its_own_scaling():
    """bench's traffic per launch is the committed counter summary's FETCH/WRITE readings times the copy
    calibration's scales, and those scales follow from the calibration kernel's known bytes."""
    d = json.load(open(bench.TRAFFIC_FILE))
    cal = d["copy_calibration"]
    assert cal["fetch_scale"] == pytest.approx(cal["bytes_read"] / 1024 / cal["FETCH_SIZE_KiB"], rel=1e-4)
    assert cal["write_scale"] == pytest.approx(cal["bytes_written"] / 1024 / cal["WRITE_SIZE_KiB"], rel=1e-4)
    nDOF = 48 ** 3 * NODES
    for key, k in d["kernels"].items():
        t = bench.load_traffic(key, 48, N)
        # the summary's bytes per launch used the unrounded scales; the file keeps them to four decimals
        assert t == pytest.approx(k["bytes_per_launch"], rel=5e-5)
        assert t / nDOF == pytest.approx(k["bytes_per_dof"], abs=0.06)
        # every kernel moves at least its algorithmic bytes
        b_t = 8 if key.endswith("fp64") else 4
        lift, rhs = bench.alg_bytes(N, b_t, b_t)
        assert t / nDOF >= (rhs if key.startswith("k_rhs") else lift) - 0.5
    assert bench.load_traffic("k_rhs_fp64", 32, N) is None  # other meshes carry no counter summary


def test_more_than_one_gpu_is_refused(capsys):
    assert bench.main(["--gpus", "2"]) == 2
    assert "one GPU" in capsys.readouterr().err
