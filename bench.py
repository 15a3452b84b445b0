"""Benchmark of the DGSEM RHS + LSERK hot path on one MI355X.

Headline workload: Taylor-Green vortex Re=1600, Ma=0.1 on [0,2pi]^3, 48^3 elements, N=7 Gauss-Lobatto (56.6M
DOF), split-form Pirozzoli volume flux, Lax-Friedrichs surface flux, BR1 lifting, Carpenter 5-stage LSERK,
FP64, synthetic initial state with seeded 1e-3 velocity noise, dt from CALCTIMESTEP at t=0 held fixed.  One
"step" = one LSERK step (5 RHS stages) over the whole mesh.  value = nDOF * 5 * steps / wall seconds
(DOF-updates/s), inputs resident in HBM before the timed region.

Also reported (extra keys): PID; on the same mesh the mixed form against FP64 over 1000 steps (throughput ratio,
per-variable relative L2) with the mixed kernels' own roofline; the discrete adjoint on the curved 32^3 mesh
(gradient of the kinetic energy after 400 steps, checkpoints every 20: cost ratios, memory, a central
finite-difference check at every step size); per-kernel times, the HBM roofline of the dominant kernel and the
oracle's CPU throughput on the headline mesh.
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
# HBM traffic per launch from rocprofv3 counter passes over the headline workload, per kernel and form
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_rhs_k_lift_tgv_n7_48cubed_traffic.json")
TRAFFIC_NE = 48


def alg_bytes(N, b_t, b_g=8, face_geo=4):
    """Algorithmic HBM bytes per DOF for one stage's two kernels (b_t: bytes per trace value, 8 in FP64, 4 in the
    mixed form; b_g: bytes per metric and face-geometry value the kernels compute with, 8 in FP64, 4 in the mixed
    form, which reads the FP32 roundings of the FP64 geometry; phi = 6 n^2 / n^3 face nodes per volume node;
    face_geo: face geometry values k_rhs reads per face node, 4 (normal, SurfElem) for the split form with
    Lax-Friedrichs, which is formed in Cartesian components, 10 (normal, both tangents, SurfElem) for the solvers
    that rotate into the normal frame):
      k_lift: read U (5x8), metrics+sJ at the surface nodes (sigma 10 b_g, sigma = (n^3 - (n-2)^3) / n^3), face
              normal+SurfElem (phi 4 b_g), neighbour U traces (phi 5 b_t); write own viscous normal-flux traces
              (phi 4 b_t)
      k_rhs : read U (5x8), metrics+sJ (10 b_g; the mixed form also reads the FP64 sJ for the update, +8), face
              geometry (phi face_geo b_g), neighbour U and viscous-flux traces (phi 9 b_t), LSERK register (5x8);
              write U (5x8), register (5x8), own next U traces (phi 5 b_t)"""
    n = N + 1
    phi = 6.0 / n
    sigma = (n ** 3 - max(n - 2, 0) ** 3) / n ** 3
    lift = 5 * 8 + sigma * 10 * b_g + phi * (4 * b_g + 5 * b_t + 4 * b_t)
    rhs = (5 * 8 + 10 * b_g + (8 if b_g < 8 else 0) + phi * face_geo * b_g + phi * 9 * b_t + 5 * 8 + 5 * 8 + 5 * 8
           + phi * 5 * b_t)
    return lift, rhs


def load_traffic(kernel, ne, N):
    """HBM bytes per launch of a kernel of the headline workload (kernel: "k_rhs_fp64", "k_lift_fp64",
    "k_rhs_mixed", "k_lift_mixed") from the committed rocprofv3 counter summary (FETCH_SIZE and WRITE_SIZE scaled by
    a plain copy kernel's readings), or None for other workloads."""
    if ne != TRAFFIC_NE or N != 7 or not os.path.exists(TRAFFIC_FILE):
        return None
    d = json.load(open(TRAFFIC_FILE))
    k = d["kernels"].get(kernel)
    if k is None:
        return None
    cal = d["copy_calibration"]
    return 1024.0 * (k["FETCH_SIZE_KiB"] * cal["fetch_scale"] + k["WRITE_SIZE_KiB"] * cal["write_scale"])


def roofline(kernel_name, alg_per_dof, nDOF, avg_s, traffic):
    achieved = alg_per_dof * nDOF / avg_s / 1e9
    return {"bound": "hbm", "kernel": kernel_name, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "alg_bytes_per_dof": alg_per_dof,
            "alg_bytes_per_launch": alg_per_dof * nDOF,
            "traffic_over_alg": (traffic / (alg_per_dof * nDOF)) if traffic else None}


def stream_rates(nbytes, reps=10):
    """STREAM-like rates on this GPU (GB/s of bytes read + written, HIP events on torch's stream, torch's own
    kernels on float64 arrays of nbytes each): copy b = a, scale b = s a, add c = a + b, triad c = a + s b.  The best
    of them is the achievable streaming rate to read the kernels' rooflines against, beside the 8 TB/s peak."""
    import torch

    n = nbytes // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.full_like(a, 2.0)
    c = torch.empty_like(a)
    ops = {"copy": (lambda: b.copy_(a), 2), "scale": (lambda: torch.mul(a, 3.0, out=b), 2),
           "add": (lambda: torch.add(a, b, out=c), 3), "triad": (lambda: torch.add(a, b, alpha=3.0, out=c), 3)}
    rates = {}
    for name, (f, arrays) in ops.items():
        f()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            f()
        e1.record()
        torch.cuda.synchronize()
        rates[name] = arrays * n * 8 * reps / (e0.elapsed_time(e1) / 1e3) / 1e9
    del a, b, c
    torch.cuda.empty_cache()
    return rates


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


def _adjoint_leg(m, U0, W, dt, steps, every, precision, storage="stages"):
    """Forward wall, checkpointed forward wall, reverse-sweep wall, dJ/dU0 and kernel timers of one context."""
    from galaexi_amd import dg

    op = dg.DGOperator(m.N, form="split", riemann="lf", precision=precision, **TGV_GAS)
    op.set_mesh(m)
    op.set_adjoint_storage(storage)
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
    # the seed dJ/dU(final) goes to device memory before the clock starts and the gradient stays there until it
    # stops: the sweep is timed on resident data, as the forward is (the host side of the seed, its copy in and the
    # gradient's copy out are 0.67 GB each at 32^3 and not part of either clock)
    import torch

    seed = np.asfortranarray(kinetic_energy_grad(W, UN))
    dseed = torch.from_numpy(seed.ravel(order="K")).to("cuda")
    torch.cuda.synchronize()
    op.timers(True)
    op.timers_reset()
    t0 = time.perf_counter()
    op.lserk_adjoint_device(dseed.data_ptr())  # returns after its stream has finished
    t_adj = time.perf_counter() - t0
    g = dseed.cpu().numpy().reshape(seed.shape, order="F")
    del dseed
    ker = {k: op.timer(k) for k in ("trace", "lift", "rhs", "adjoint")}
    op.timers(False)
    return op, dict(t_fwd=t_fwd, t_rec=t_rec, t_adj=t_adj, g=g, J=J, ker=ker, mem=op.memory_bytes())


def run_adjoint(ne, steps, every, fd_eps, mixed_leg=True):
    """J = int 1/2 rho |u|^2 dV after `steps` LSERK steps of the curved TGV case; dJ/dU0 by the checkpointed
    discrete adjoint.  Cost ratios over the wall of the plain forward integration: the whole gradient (forward
    with checkpoints + reverse sweep) and the reverse sweep alone (segment recompute + stage VJPs).  FP64 leg: a
    central finite difference along a well-conditioned direction (the normalised gradient plus a seeded
    perturbation, so that dJ.v is a sizeable part of J) checks the gradient at every step size.  Mixed leg: the
    forward (and the segment recompute) in the mixed form, the adjoint accumulated in FP64 along that trajectory;
    its gradient is compared with the FP64 leg's and with the FP64 finite differences.  Both legs keep each
    segment's stage inputs (GALX_ADJ_STAGES); a third leg repeats the mixed one keeping only each segment's step
    inputs and recomputing the stages of each step (GALX_ADJ_STEPS): less memory, more recompute, the same
    gradient bit for bit."""
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
    rv = rng.standard_normal(U0.shape)
    v = np.asfortranarray(2.0 * (g0 / np.linalg.norm(g0.ravel()) + 0.3 * rv / np.linalg.norm(rv.ravel())))
    del rv
    exact = float(np.sum(g0 * v))
    fd_rel, cd = {}, {}
    for eps in fd_eps:
        Js = []
        for sg in (1.0, -1.0):
            op.set_state(np.asfortranarray(U0 + sg * eps * v))
            op.TimeStepByLSERKW2(0.0, dt, steps)
            Js.append(kinetic_energy(W, op.get_state()))
        cd[eps] = (Js[0] - Js[1]) / (2 * eps)
        fd_rel[f"{eps:g}"] = abs(cd[eps] - exact) / abs(exact)
    op.close()
    ker = r["ker"]
    out = {
        "workload": f"TGV Re1600 Ma0.1 [0,2pi]^3 {ne}^3 elems N={N}, curved (0.05 sine mapping), split-PI + LF, "
                    f"BR1, FP64; J = int 1/2 rho|u|^2 dV after {steps} LSERK steps, checkpoints every {every}",
        "nDOF": m.nDOF, "dt": dt, "J": r["J"],
        "forward_s": r["t_fwd"], "forward_with_checkpoints_s": r["t_rec"], "adjoint_sweep_s": r["t_adj"],
        "gradient_over_forward": (r["t_rec"] + r["t_adj"]) / r["t_fwd"],
        "sweep_over_forward": r["t_adj"] / r["t_fwd"],
        "adjoint_kernels_ms_avg": {k: (v_[0] / v_[1] if v_[1] else 0.0) for k, v_ in ker.items()},
        "adjoint_kernels_ms_total": {k: v_[0] for k, v_ in ker.items()},
        "device_memory_GB": r["mem"] / 1e9,
        "fd_check": {"direction": "2 (g/|g| + 0.3 r/|r|), r seeded N(0,1)", "dJ_dot_v": exact,
                     "dJ_dot_v_over_J": exact / r["J"], "central_fd_rel_err_by_eps": fd_rel},
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
            "gradient_over_forward": (rm["t_rec"] + rm["t_adj"]) / rm["t_fwd"],
            "sweep_over_forward": rm["t_adj"] / rm["t_fwd"],
            "gradient_over_fp64_forward": (rm["t_rec"] + rm["t_adj"]) / r["t_fwd"],
            "J_rel_diff_vs_fp64": abs(rm["J"] - r["J"]) / abs(r["J"]),
            "gradient_rel_l2_vs_fp64_adjoint": float(np.linalg.norm((gm - g0).ravel()) / np.linalg.norm(g0.ravel())),
            "dJ_dot_v_rel_err_vs_fp64_fd_by_eps": {f"{e:g}": abs(float(np.sum(gm * v)) - cd[e]) / abs(cd[e])
                                                   for e in fd_eps},
            "adjoint_kernels_ms_avg": {k: (v_[0] / v_[1] if v_[1] else 0.0) for k, v_ in kerm.items()},
            "device_memory_GB": rm["mem"] / 1e9,
        }
        op, rs = _adjoint_leg(m, U0, W, dt, steps, every, "mixed", storage="steps")
        op.close()
        out["mixed_forward_step_storage"] = {
            "what": "the mixed leg with GALX_ADJ_STEPS: each segment's step inputs kept, each step's stages and the "
                    "FP64 trace records recomputed before its VJPs",
            "forward_s": rs["t_fwd"], "forward_with_checkpoints_s": rs["t_rec"], "adjoint_sweep_s": rs["t_adj"],
            "gradient_over_forward": (rs["t_rec"] + rs["t_adj"]) / rs["t_fwd"],
            "sweep_over_forward": rs["t_adj"] / rs["t_fwd"],
            "gradient_bitwise_equal_to_stage_storage": bool(np.array_equal(rs["g"], gm)),
            "device_memory_GB": rs["mem"] / 1e9,
        }
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_threads():
    """The host threads a run may use: the OMP_NUM_THREADS the machine sets for one GPU's CPU share, else the CPUs
    this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.strip().isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_baseline(m, U, steps=1):
    """The FP64 oracle (C/OpenMP restatement of the reference operator) on the host threads of this GPU's CPU share:
    `steps` LSERK steps of the headline workload itself (same mesh, state and dt rule), timed after one RHS
    evaluation that first-touches the oracle's work arrays."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    threads = cpu_threads()
    O.set_threads(threads)
    op = O.Operator(m, split=True, riemann=O.RIEMANN_LF, kappa=1.4, R=71.42857, Pr=0.72, mu0=1.0 / 1600.0)
    dt, _e, _c, _v = op.calc_dt(U, TGV_GAS["CFLScale"], TGV_GAS["DFLScale"])
    K = np.zeros_like(U, order="F")
    Uc = np.asfortranarray(U.copy())
    op.rhs(Uc)  # first touch
    t0 = time.perf_counter()
    for s in range(steps):
        op.lserk_step(Uc, K, s * dt, dt)
    wall = time.perf_counter() - t0
    del op
    return dict(value=m.nDOF * NSTAGES * steps / wall, unit="DOF-updates/s", cores=threads, kind="port",
                sample=f"oracle FP64, {steps} LSERK step(s) (5 stages each) of the headline workload, TGV N={m.N} "
                       f"{m.nElems} elements ({m.nDOF} DOF), {wall:.2f} s wall on {threads} threads",
                cpu=cpu_model(), nproc=os.cpu_count(), affinity_cpus=len(os.sched_getaffinity(0)))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--ne", type=int, default=48, help="elements per direction of the headline workload")
    ap.add_argument("--compare-steps", type=int, default=1000,
                    help="steps of the FP64-vs-mixed comparison on the headline mesh (0: skip)")
    ap.add_argument("--adjoint-ne", type=int, default=32, help="mesh of the adjoint measurement (0: skip)")
    ap.add_argument("--adjoint-steps", type=int, default=400)
    ap.add_argument("--adjoint-ckpt", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=1, help="LSERK steps of the oracle's timed sample (0: skip)")
    a = ap.parse_args(argv)
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (one rank's element share per GPU; "
              "run with --gpus 1)", file=sys.stderr)
        return 2
    N = 7
    m, U = tgv_case(a.ne, N)
    nDOF = m.nDOF
    print(f"# headline: TGV N={N} {a.ne}^3 ({nDOF} DOF) FP64, {a.steps} steps", file=sys.stderr, flush=True)
    r = run_gpu(m, U, "fp64", a.steps, a.warmup)
    value = nDOF * NSTAGES * a.steps / r["wall"]
    ms_step = 1e3 * r["wall"] / a.steps
    lift_b, rhs_b = alg_bytes(N, 8)
    avg_s = {k: r["kernels"][k][0] / r["kernels"][k][1] / 1e3 for k in ("lift", "rhs")}
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
        "kernels_ms_avg": {k: 1e3 * v for k, v in avg_s.items()},
        "roofline": roofline("k_rhs<7,double,double,split,LF>", rhs_b, nDOF, avg_s["rhs"],
                             load_traffic("k_rhs_fp64", a.ne, N)),
        "roofline_k_lift": roofline("k_lift<7,double,double>", lift_b, nDOF, avg_s["lift"],
                                    load_traffic("k_lift_fp64", a.ne, N)),
        "stage_roofline": {"alg_bytes_per_dof_stage": lift_b + rhs_b,
                           "achieved_GBs": (lift_b + rhs_b) * nDOF / (avg_s["lift"] + avg_s["rhs"]) / 1e9},
        "device_memory_GB": r["mem"] / 1e9,
    }
    dt = r["dt"]
    del r
    # the box's achievable streaming rate: STREAM-like kernels over arrays of the headline state's size (2.26 GB)
    rates = stream_rates(5 * nDOF * 8)
    cgbs = max(rates.values())
    out["hbm_stream"] = {"GBs": cgbs, "by_kernel_GBs": rates, "array_bytes": 5 * nDOF * 8,
                         "what": "best of copy, scale, add, triad (torch kernels, float64; bytes read + written over "
                                 "the HIP-event time of 10 launches)"}
    for key in ("roofline", "roofline_k_lift"):
        out[key]["frac_of_stream"] = out[key]["achieved"] / cgbs
    if a.compare_steps:
        print(f"# compare: TGV N={N} {a.ne}^3 FP64 vs mixed, {a.compare_steps} steps", file=sys.stderr, flush=True)
        r64 = run_gpu(m, U, "fp64", a.compare_steps, 1, dt=dt)
        r32 = run_gpu(m, U, "mixed", a.compare_steps, 1, dt=dt)
        v64 = nDOF * NSTAGES * a.compare_steps / r64["wall"]
        v32 = nDOF * NSTAGES * a.compare_steps / r32["wall"]
        Ua, Ub = r64["U"], r32["U"]
        l2 = [float(np.linalg.norm((Ub[v] - Ua[v]).ravel()) / np.linalg.norm(Ua[v].ravel())) for v in range(5)]
        l2_all = float(np.linalg.norm((Ub - Ua).ravel()) / np.linalg.norm(Ua.ravel()))
        del Ua, Ub
        mlift_b, mrhs_b = alg_bytes(N, 4, 4)
        mavg = {k: r32["kernels"][k][0] / r32["kernels"][k][1] / 1e3 for k in ("lift", "rhs")}
        out["compare"] = {"workload": f"TGV N={N} {a.ne}^3 ({nDOF} DOF), {a.compare_steps}+1 steps from t=0 with the "
                                      f"headline dt",
                          "fp64_dof_updates_per_s": v64, "mixed_dof_updates_per_s": v32,
                          "mixed_over_fp64": v32 / v64,
                          "mixed_rel_l2_per_var_after_steps": l2, "mixed_rel_l2_all_vars": l2_all,
                          "mixed_alg_bytes_per_dof_stage": mlift_b + mrhs_b,
                          "fp64_kernels_ms_avg": {k: r64["kernels"][k][0] / r64["kernels"][k][1]
                                                  for k in ("lift", "rhs")},
                          "mixed_kernels_ms_avg": {k: 1e3 * v for k, v in mavg.items()},
                          "roofline_mixed": roofline("k_rhs<7,float,float,split,LF>", mrhs_b, nDOF, mavg["rhs"],
                                                     load_traffic("k_rhs_mixed", a.ne, N)),
                          "roofline_mixed_k_lift": roofline("k_lift<7,float,float>", mlift_b, nDOF, mavg["lift"],
                                                            load_traffic("k_lift_mixed", a.ne, N))}
        for key in ("roofline_mixed", "roofline_mixed_k_lift"):
            out["compare"][key]["frac_of_stream"] = out["compare"][key]["achieved"] / cgbs
        del r64, r32
    if a.cpu_steps:
        print(f"# cpu baseline: oracle, {a.cpu_steps} step(s) of the headline workload", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(m, U, a.cpu_steps)
    del m, U
    if a.adjoint_ne:
        print(f"# adjoint: curved TGV N={N} {a.adjoint_ne}^3, {a.adjoint_steps} steps, checkpoints every "
              f"{a.adjoint_ckpt}", file=sys.stderr, flush=True)
        out["adjoint"] = run_adjoint(a.adjoint_ne, a.adjoint_steps, a.adjoint_ckpt, (1e-1, 1e-2, 1e-3))
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
