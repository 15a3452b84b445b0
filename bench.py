"""Benchmark of the DGSEM RHS + LSERK hot path on one MI355X.

Headline workload: Taylor-Green vortex Re=1600, Ma=0.1 on [0,2pi]^3, 32^3 elements, N=7 Gauss-Lobatto (16.8M
DOF), split-form Pirozzoli volume flux, Lax-Friedrichs surface flux, BR1 lifting, Carpenter 5-stage LSERK,
FP64, synthetic initial state with seeded 1e-3 velocity noise, dt from CALCTIMESTEP at t=0 held fixed.  One
"step" = one LSERK step (5 RHS stages) over the whole mesh.  value = nDOF * 5 * steps / wall seconds
(DOF-updates/s), inputs resident in HBM before the timed region.

Also reported (extra keys): PID, the same configuration at 48^3 in FP64 and in the mixed form with their ratio
and per-variable error, the discrete adjoint on the curved 32^3 mesh (gradient of the kinetic energy after 400
steps, checkpoints every 20: adjoint/forward cost ratio and a finite-difference check), per-kernel times, the
HBM roofline of the dominant kernel and the oracle's CPU throughput.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "PID & DOF-updates/s, FP64 vs mixed at stated L2; adjoint/forward cost ratio"
HBM_PEAK_GBS = 8000.0
NSTAGES = 5
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_rhs_k_lift_tgv_n7_32cubed_fp64_traffic.json")


def alg_bytes(N, b_t, face_geo=4):
    """Algorithmic HBM bytes per DOF for one stage's two kernels (b_t: bytes per trace value, 8 in FP64, 4 in the
    mixed form; phi = 6 n^2 / n^3 face nodes per volume node; face_geo: face geometry values k_rhs reads per face
    node, 4 (normal, SurfElem) for the split form with Lax-Friedrichs, which is formed in Cartesian components,
    10 (normal, both tangents, SurfElem) for the solvers that rotate into the normal frame):
      k_lift: read U (5x8), metrics+sJ at the surface nodes (sigma 10x8, sigma = (n^3 - (n-2)^3) / n^3), face
              normal+SurfElem (phi 4x8), neighbour U traces (phi 5 b_t); write own viscous normal-flux traces
              (phi 4 b_t)
      k_rhs : read U (5x8), metrics+sJ (10x8), face geometry (phi face_geo x8), neighbour U and viscous-flux traces
              (phi 9 b_t), LSERK register (5x8); write U (5x8), register (5x8), own next U traces (phi 5 b_t)"""
    n = N + 1
    phi = 6.0 / n
    sigma = (n ** 3 - max(n - 2, 0) ** 3) / n ** 3
    lift = 5 * 8 + sigma * 10 * 8 + phi * (4 * 8 + 5 * b_t + 4 * b_t)
    rhs = 5 * 8 + 10 * 8 + phi * face_geo * 8 + phi * 9 * b_t + 5 * 8 + 5 * 8 + 5 * 8 + phi * 5 * b_t
    return lift, rhs


def load_traffic(kernel, ne, N):
    """HBM bytes per launch of a kernel of the headline workload from the committed rocprofv3 counter summary
    (FETCH_SIZE and WRITE_SIZE scaled by a plain copy kernel's readings), or None for other workloads."""
    if ne != 32 or N != 7 or not os.path.exists(TRAFFIC_FILE):
        return None
    d = json.load(open(TRAFFIC_FILE))
    k = d["kernels"].get(kernel)
    if k is None:
        return None
    cal = d["copy_calibration"]
    return 1024.0 * (k["FETCH_SIZE_KiB"] * cal["fetch_scale"] + k["WRITE_SIZE_KiB"] * cal["write_scale"])


def tgv_case(ne, N=7, noise=1e-3, curved=False):
    from galaexi_amd import mesh as M
    from galaexi_amd import testcases as T

    kw = dict(deform_amp=0.05, deform_phase=(0.31, 0.17, 0.83)) if curved else {}
    m = M.box_mesh((ne, ne, ne), N, lo=(0.0, 0.0, 0.0), hi=(2 * np.pi,) * 3, **kw)
    U = T.taylor_green(m.Elem_xGP, Ma=0.1, noise=noise, seed=1600)
    return m, U


# CFL 0.5: with 1e-3 i.i.d. nodal velocity noise the split-PI/LF scheme is unstable at the tutorial CFL 0.9 (the
# FP64 oracle blows up the same way); 0.5 keeps 1000+ steps stable.
TGV_GAS = dict(kappa=1.4, R=71.42857, Pr=0.72, mu0=1.0 / 1600.0, CFLScale=0.5, DFLScale=0.5)


def run_gpu(m, U, precision, steps, warmup, dt=None):
    from galaexi_amd import dg

    op = dg.DGOperator(m.N, form="split", riemann="lf", precision=precision, **TGV_GAS)
    op.set_mesh(m)
    op.set_state(U)
    if dt is None:
        dt, err = op.CalcTimeStep()
        assert err == 0, err
    t = op.TimeStepByLSERKW2(0.0, dt, nSteps=warmup)
    op.timers(True)
    op.timers_reset()
    t0 = time.perf_counter()
    op.TimeStepByLSERKW2(t, dt, nSteps=steps)  # one call: no host synchronisation between steps
    t1 = time.perf_counter()
    ker = {k: op.timer(k) for k in ("lift", "rhs")}
    op.timers(False)
    Uend = op.get_state()
    mem = op.memory_bytes()
    op.close()
    return dict(wall=t1 - t0, kernels=ker, dt=dt, U=Uend, mem=mem)


def mass_weights(m):
    from galaexi_amd import basis

    _x, w, _wb = basis.nodes_and_weights(m.N, basis.GAUSS_LOBATTO)
    W = w[:, None, None, None] * w[None, :, None, None] * w[None, None, :, None]
    return W / m.sJ


def kinetic_energy(W, U):
    return float(np.sum(W * 0.5 * (U[1] ** 2 + U[2] ** 2 + U[3] ** 2) / U[0]))


def kinetic_energy_grad(W, U):
    g = np.zeros_like(U, order="F")
    g[0] = -0.5 * W * (U[1] ** 2 + U[2] ** 2 + U[3] ** 2) / U[0] ** 2
    for d in (1, 2, 3):
        g[d] = W * U[d] / U[0]
    return g


def _adjoint_leg(m, U0, W, dt, steps, every, precision):
    """Forward wall, checkpointed forward wall, reverse-sweep wall, dJ/dU0 and kernel timers of one context."""
    from galaexi_amd import dg

    op = dg.DGOperator(m.N, form="split", riemann="lf", precision=precision, **TGV_GAS)
    op.set_mesh(m)
    op.set_state(U0)
    op.TimeStepByLSERKW2(0.0, dt, 2)  # warm-up
    op.set_state(U0)
    t0 = time.perf_counter()
    op.TimeStepByLSERKW2(0.0, dt, steps)
    t_fwd = time.perf_counter() - t0
    op.set_state(U0)
    t0 = time.perf_counter()
    op.forward_ckpt(0.0, dt, steps, every)
    t_rec = time.perf_counter() - t0
    UN = op.get_state()
    J = kinetic_energy(W, UN)
    op.timers(True)
    op.timers_reset()
    t0 = time.perf_counter()
    g = op.lserk_adjoint(kinetic_energy_grad(W, UN))
    t_adj = time.perf_counter() - t0
    ker = {k: op.timer(k) for k in ("trace", "lift", "rhs", "adjoint")}
    op.timers(False)
    return op, dict(t_fwd=t_fwd, t_rec=t_rec, t_adj=t_adj, g=g, J=J, ker=ker, mem=op.memory_bytes())


def run_adjoint(ne, steps, every, fd_eps, mixed_leg=True):
    """J = int 1/2 rho |u|^2 dV after `steps` LSERK steps of the curved TGV case; dJ/dU0 by the checkpointed
    discrete adjoint.  Cost ratio = wall of the reverse sweep (segment recompute + stage VJPs) over the wall of
    the plain forward integration.  FP64 leg: a central finite difference along one seeded direction checks
    the gradient.  Mixed leg: the forward (and the segment recompute) in the mixed form, the adjoint accumulated
    in FP64 along that trajectory; its gradient is compared with the FP64 leg's and with the same FD value."""
    N = 7
    m, U0 = tgv_case(ne, N, curved=True)
    W = mass_weights(m)
    from galaexi_amd import dg

    probe = dg.DGOperator(N, form="split", riemann="lf", precision="fp64", **TGV_GAS)
    probe.set_mesh(m)
    probe.set_state(U0)
    dt, err = probe.CalcTimeStep()
    probe.close()
    assert err == 0, err
    op, r = _adjoint_leg(m, U0, W, dt, steps, every, "fp64")
    g0 = r["g"]
    rng = np.random.default_rng(400)
    s = np.abs(U0).reshape(5, -1).mean(axis=1).reshape(5, 1, 1, 1, 1)
    v = np.asfortranarray(0.1 * s * rng.standard_normal(U0.shape))
    exact = float(np.sum(g0 * v))
    fd, cd = {}, {}
    for eps in fd_eps:
        Js = []
        for sg in (1.0, -1.0):
            op.set_state(np.asfortranarray(U0 + sg * eps * v))
            op.TimeStepByLSERKW2(0.0, dt, steps)
            Js.append(kinetic_energy(W, op.get_state()))
        cd[eps] = (Js[0] - Js[1]) / (2 * eps)
        fd[f"{eps:g}"] = abs(cd[eps] - exact) / abs(exact)
    # Richardson extrapolation of the two largest steps (2e, e) removes the O(e^2) truncation term
    e2, e1 = sorted(cd)[-1], sorted(cd)[-2]
    fd_best = cd[e1]
    if abs(e2 - 2 * e1) < 1e-12 * e2:
        fd_best = (4 * cd[e1] - cd[e2]) / 3
        fd[f"richardson({e1:g},{e2:g})"] = abs(fd_best - exact) / abs(exact)
    op.close()
    ker = r["ker"]
    out = {
        "workload": f"TGV Re1600 Ma0.1 [0,2pi]^3 {ne}^3 elems N={N}, curved (0.05 sine mapping), split-PI + LF, "
                    f"BR1, FP64; J = int 1/2 rho|u|^2 dV after {steps} LSERK steps, checkpoints every {every}",
        "nDOF": m.nDOF, "dt": dt, "J": r["J"],
        "forward_s": r["t_fwd"], "forward_with_checkpoints_s": r["t_rec"], "adjoint_sweep_s": r["t_adj"],
        "adjoint_over_forward": r["t_adj"] / r["t_fwd"],
        "gradient_total_over_forward": (r["t_rec"] + r["t_adj"]) / r["t_fwd"],
        "adjoint_kernels_ms_avg": {k: (v_[0] / v_[1] if v_[1] else 0.0) for k, v_ in ker.items()},
        "adjoint_kernels_ms_total": {k: v_[0] for k, v_ in ker.items()},
        "device_memory_GB": r["mem"] / 1e9,
        "fd_check": {"direction": "seeded N(0,1) x 0.1 mean|U_v| per variable", "dJ_dot_v": exact,
                     "central_fd_rel_err_by_eps": fd, "best": min(fd.values())},
    }
    if mixed_leg:
        op, rm = _adjoint_leg(m, U0, W, dt, steps, every, "mixed")
        op.close()
        gm = rm["g"]
        kerm = rm["ker"]
        out["mixed_forward"] = {
            "what": "forward and segment recompute in the mixed form (FP32 fluxes), adjoint accumulated in FP64 "
                    "along that trajectory",
            "forward_s": rm["t_fwd"], "forward_with_checkpoints_s": rm["t_rec"], "adjoint_sweep_s": rm["t_adj"],
            "adjoint_over_forward": rm["t_adj"] / rm["t_fwd"],
            "adjoint_over_fp64_forward": rm["t_adj"] / r["t_fwd"],
            "gradient_total_over_forward": (rm["t_rec"] + rm["t_adj"]) / rm["t_fwd"],
            "J_rel_diff_vs_fp64": abs(rm["J"] - r["J"]) / abs(r["J"]),
            "gradient_rel_l2_vs_fp64_adjoint": float(np.linalg.norm((gm - g0).ravel()) / np.linalg.norm(g0.ravel())),
            "dJ_dot_v_rel_err_vs_fp64_fd": abs(float(np.sum(gm * v)) - fd_best) / abs(fd_best),
            "adjoint_kernels_ms_avg": {k: (v_[0] / v_[1] if v_[1] else 0.0) for k, v_ in kerm.items()},
            "device_memory_GB": rm["mem"] / 1e9,
        }
    return out


def cpu_baseline(N=7, ne=16, min_wall=1.0, max_steps=20):
    """The FP64 oracle (C/OpenMP restatement of the reference operator) on the host cores (at most 16, the CPU
    share of one GPU), LSERK steps of the same TGV configuration on a 16^3 mesh (same per-DOF work as the
    headline mesh) until at least min_wall seconds have passed (about 16 CPU-seconds on 16 threads)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    m, U = tgv_case(ne, N)
    cores = max(1, min(os.cpu_count() or 1, 16))
    O.set_threads(cores)
    op = O.Operator(m, split=True, riemann=O.RIEMANN_LF, kappa=1.4, R=71.42857, Pr=0.72, mu0=1.0 / 1600.0)
    dt, _e, _c, _v = op.calc_dt(U, TGV_GAS["CFLScale"], TGV_GAS["DFLScale"])
    K = np.zeros_like(U, order="F")
    Uc = np.asfortranarray(U.copy())
    op.rhs(Uc)  # first touch
    steps = 0
    t0 = time.perf_counter()
    while steps < max_steps and (steps == 0 or time.perf_counter() - t0 < min_wall):
        op.lserk_step(Uc, K, steps * dt, dt)
        steps += 1
    wall = time.perf_counter() - t0
    return dict(value=m.nDOF * NSTAGES * steps / wall, unit="DOF-updates/s", cores=cores, kind="port",
                sample=f"oracle FP64, {steps} LSERK steps (5 stages each) of TGV N={N} {ne}^3 ({m.nDOF} DOF), "
                       f"{wall:.2f} s wall on {cores} threads",
                cpu=platform.processor() or platform.machine())


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--ne", type=int, default=32, help="elements per direction of the headline workload")
    ap.add_argument("--compare-ne", type=int, default=48, help="mesh for the FP64-vs-mixed comparison (0: skip)")
    ap.add_argument("--compare-steps", type=int, default=1000)
    ap.add_argument("--adjoint-ne", type=int, default=32, help="mesh of the adjoint measurement (0: skip)")
    ap.add_argument("--adjoint-steps", type=int, default=400)
    ap.add_argument("--adjoint-ckpt", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    a = ap.parse_args(argv)
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (one rank's element share per GPU; "
              "run with --gpus 1)", file=sys.stderr)
        return 2
    N = 7
    m, U = tgv_case(a.ne, N)
    print(f"# headline: TGV N={N} {a.ne}^3 ({m.nDOF} DOF) FP64, {a.steps} steps", file=sys.stderr, flush=True)
    r = run_gpu(m, U, "fp64", a.steps, a.warmup)
    nDOF = m.nDOF
    value = nDOF * NSTAGES * a.steps / r["wall"]
    ms_step = 1e3 * r["wall"] / a.steps
    lift_b, rhs_b = alg_bytes(N, 8)
    ms_rhs, cnt_rhs = r["kernels"]["rhs"]
    ms_lift, cnt_lift = r["kernels"]["lift"]
    rhs_avg_s = ms_rhs / cnt_rhs / 1e3
    lift_avg_s = ms_lift / cnt_lift / 1e3
    achieved = rhs_b * nDOF / rhs_avg_s / 1e9
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (TGV initial state + seeded 1e-3 velocity noise, analytic periodic box mesh)",
        "config": {"workload": f"TGV Re1600 Ma0.1 [0,2pi]^3 {a.ne}^3 elems N={N} GL, split-PI + LF, BR1, "
                               f"Carpenter LSERK4-5, FP64", "nDOF": nDOF, "N": N, "elements": m.nElems,
                   "dt": r["dt"]},
        "pid_s": 1.0 / value,
        "kernels_ms_avg": {"lift": 1e3 * lift_avg_s, "rhs": 1e3 * rhs_avg_s},
        "roofline": {"bound": "hbm", "kernel": "k_rhs<7,double,double,split,LF>", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": load_traffic("k_rhs", a.ne, N), "alg_bytes_per_dof": rhs_b,
                     "alg_bytes_per_launch": rhs_b * nDOF},
        "stage_roofline": {"alg_bytes_per_dof_stage": lift_b + rhs_b,
                           "achieved_GBs": (lift_b + rhs_b) * nDOF / (lift_avg_s + rhs_avg_s) / 1e9,
                           "k_lift_achieved_GBs": lift_b * nDOF / lift_avg_s / 1e9,
                           "k_lift_traffic": load_traffic("k_lift", a.ne, N)},
        "device_memory_GB": r["mem"] / 1e9,
    }
    del m, U, r
    if a.compare_ne:
        m2, U2 = tgv_case(a.compare_ne, N)
        print(f"# compare: TGV N={N} {a.compare_ne}^3 ({m2.nDOF} DOF) FP64 vs mixed, {a.compare_steps} steps",
              file=sys.stderr, flush=True)
        r64 = run_gpu(m2, U2, "fp64", a.compare_steps, 1)
        r32 = run_gpu(m2, U2, "mixed", a.compare_steps, 1, dt=r64["dt"])
        v64 = m2.nDOF * NSTAGES * a.compare_steps / r64["wall"]
        v32 = m2.nDOF * NSTAGES * a.compare_steps / r32["wall"]
        Ua, Ub = r64["U"], r32["U"]
        l2 = [float(np.linalg.norm((Ub[v] - Ua[v]).ravel()) / np.linalg.norm(Ua[v].ravel())) for v in range(5)]
        out["compare"] = {"workload": f"TGV N={N} {a.compare_ne}^3 ({m2.nDOF} DOF), {a.compare_steps}+1 steps",
                          "fp64_dof_updates_per_s": v64, "mixed_dof_updates_per_s": v32,
                          "mixed_over_fp64": v32 / v64,
                          "mixed_rel_l2_per_var_after_steps": l2,
                          "mixed_alg_bytes_per_dof_stage": sum(alg_bytes(N, 4)),
                          "fp64_kernels_ms_avg": {k: r64["kernels"][k][0] / r64["kernels"][k][1]
                                                  for k in ("lift", "rhs")},
                          "mixed_kernels_ms_avg": {k: r32["kernels"][k][0] / r32["kernels"][k][1]
                                                   for k in ("lift", "rhs")}}
        del m2, U2, r64, r32
    if a.adjoint_ne:
        print(f"# adjoint: curved TGV N={N} {a.adjoint_ne}^3, {a.adjoint_steps} steps, checkpoints every "
              f"{a.adjoint_ckpt}", file=sys.stderr, flush=True)
        out["adjoint"] = run_adjoint(a.adjoint_ne, a.adjoint_steps, a.adjoint_ckpt, (2e-4, 1e-4, 1e-5))
    if not a.no_cpu_baseline:
        print("# cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
