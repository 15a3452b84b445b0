"""GPU tests of the discrete adjoint (split PI + LF, BR1, FP64; also along a mixed-precision forward) through the C
ABI.

Bar (stated per test): the vector-Jacobian product agrees with central finite differences of the FP64 oracle's
operator to 1e-6 relative (best over a step-size sweep), with the reference's matrix-free Jacobian-vector
product formula (src/implicit/implicit.f90:150-170,480-528) to 1e-6, and annihilates the discrete
conservation weights (exact identity, to rounding); the checkpointed LSERK adjoint agrees with central finite
differences of the time integration to 1e-6 and is bitwise independent of the checkpoint spacing.
"""
import numpy as np
import pytest

import oracle as O
from galaexi_amd import mesh as M
from galaexi_amd import testcases as T

pytestmark = pytest.mark.gpu

GAS = dict(kappa=1.4, R=1.0, Pr=0.72, mu0=2e-2)


def say(*a):
    print(*a, flush=True)


def dgmod():
    from galaexi_amd import dg
    return dg


def noisy_case(nel, N, seed=0, curved=True, rot=True, noise=1e-2):
    """Smooth periodic state plus seeded nodal noise, so that the two traces of every face differ (the LF
    dissipation and its max() switches are exercised)."""
    lo, hi = (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0)
    kw = {}
    if curved:
        kw.update(deform_amp=0.05, deform_phase=(0.13, 0.37, 0.71))
    if rot:
        kw.update(rotations="random", seed=seed + 11)
    m = M.box_mesh(nel, N, lo=lo, hi=hi, **kw)
    U = T.smooth_random(m.Elem_xGP, lo, hi, seed=seed)
    rng = np.random.default_rng(seed + 100)
    U = U * (1.0 + noise * rng.standard_normal(U.shape))
    return m, np.asfortranarray(U)


def direction(U, rng):
    """Random direction scaled per variable like the state."""
    s = np.abs(U).reshape(5, -1).mean(axis=1).reshape((5,) + (1,) * (U.ndim - 1))
    return np.asfortranarray(0.1 * s * rng.standard_normal(U.shape))


def gpu_op(m, **kw):
    g = dict(GAS)
    g.update(kw)
    op = dgmod().DGOperator(m.N, form="split", riemann="lf", precision="fp64", **g)
    op.set_mesh(m)
    return op


def mass_weights(m):
    w = O.dg_basis(m.N)["wGP"]
    W = w[:, None, None, None] * w[None, :, None, None] * w[None, None, :, None]
    return W / m.sJ  # (n,n,n,nElems): quadrature weight times Jacobian


def kinetic_energy(m, U):
    W = mass_weights(m)
    return float(np.sum(W * 0.5 * (U[1] ** 2 + U[2] ** 2 + U[3] ** 2) / U[0]))


def kinetic_energy_grad(m, U):
    W = mass_weights(m)
    g = np.zeros_like(U, order="F")
    q2 = (U[1] ** 2 + U[2] ** 2 + U[3] ** 2) / U[0] ** 2
    g[0] = -0.5 * W * q2
    for d in (1, 2, 3):
        g[d] = W * U[d] / U[0]
    return g


def best_fd_error(f, U, v, lam_dot_J_v_exact, eps_list):
    errs = []
    for eps in eps_list:
        fd = (f(U + eps * v) - f(U - eps * v)) / (2 * eps)
        errs.append(abs(fd - lam_dot_J_v_exact) / max(abs(lam_dot_J_v_exact), 1e-300))
    return min(errs), errs


def test_vjp_annihilates_conservation_weights():
    """Periodic mesh: sum_i w_i J_i Ut_i = 0 for every U, so (dR/dU)^T applied to the conservation weights is
    zero for each conservative variable (to rounding), while a random cotangent of the same size is not."""
    say("case: vjp conservation weights, (3,2,2) N=4 curved rotated")
    m, U = noisy_case((3, 2, 2), 4, seed=1)
    op = gpu_op(m)
    op.set_state(U)
    W = mass_weights(m)
    rng = np.random.default_rng(5)
    for v in range(5):
        lam = np.zeros_like(U, order="F")
        lam[v] = W
        Ub = op.rhs_vjp(lam)
        lam_r = np.zeros_like(U, order="F")
        lam_r[v] = W * rng.choice([-1.0, 1.0], size=W.shape)
        scale = np.abs(op.rhs_vjp(lam_r)).max()
        assert np.abs(Ub).max() <= 1e-12 * scale, (v, np.abs(Ub).max(), scale)
    op.close()


@pytest.mark.parametrize("N,nel,mu0", [(3, (3, 2, 2), 2e-2), (4, (2, 2, 3), 0.0), (5, (2, 2, 2), 5e-3),
                                       (7, (2, 2, 2), 2e-2)])
def test_vjp_vs_central_fd_of_oracle(N, nel, mu0):
    """lam . (R(U+eps v) - R(U-eps v))/(2 eps) of the FP64 oracle against Ubar . v: best over eps <= 1e-6."""
    say(f"case: vjp vs oracle FD, {nel} N={N} mu0={mu0}")
    m, U = noisy_case(nel, N, seed=N)
    op = gpu_op(m, mu0=mu0)
    op.set_state(U)
    ref = O.Operator(m, split=True, riemann=O.RIEMANN_LF, kappa=1.4, R=1.0, Pr=0.72, mu0=mu0)
    rng = np.random.default_rng(7 + N)
    lam = np.asfortranarray(rng.standard_normal(U.shape))
    Ub = op.rhs_vjp(lam)
    for it in range(3):
        v = direction(U, rng)
        exact = float(np.sum(Ub * v))
        best, errs = best_fd_error(lambda X: float(np.sum(lam * ref.rhs(np.asfortranarray(X)))), U, v, exact,
                                   [1e-3, 1e-4, 1e-5, 1e-6])
        say(f"  dir {it}: Ubar.v={exact:.6e} fd rel errs {['%.1e' % e for e in errs]}")
        assert best <= 1e-6
    op.close()


@pytest.mark.parametrize("fd_order", [1, 2])
def test_vjp_vs_reference_matrix_vector_fd(fd_order):
    """The reference's matrix-free Jacobian-vector product (MatrixVector, implicit.f90:480-528, Eps_Method 2:
    EpsFD = rEps0/|v| * sqrt(1+|X|), rEps0 = sqrt(eps) for FD_Order 1 and (eps/2)^(1/3) for FD_Order 2,
    implicit.f90:160-170) applied to the oracle, contracted with lam, against Ubar . v: within 1e-6."""
    say(f"case: vjp vs MatrixVector FD order {fd_order}")
    m, U = noisy_case((3, 2, 2), 3, seed=21)
    op = gpu_op(m)
    op.set_state(U)
    ref = O.Operator(m, split=True, riemann=O.RIEMANN_LF, kappa=1.4, R=1.0, Pr=0.72, mu0=GAS["mu0"])
    rng = np.random.default_rng(3)
    lam = np.asfortranarray(rng.standard_normal(U.shape))
    Ub = op.rhs_vjp(lam)
    eps = np.finfo(np.float64).eps
    r_eps0 = np.sqrt(eps) if fd_order == 1 else (0.5 * eps) ** (1.0 / 3.0)
    norm_x = np.sqrt(np.sum(U * U))
    R0 = ref.rhs(U)
    for it in range(3):
        v = direction(U, rng)
        eps_fd = r_eps0 / np.sqrt(np.sum(v * v)) * np.sqrt(1.0 + norm_x)
        Rp = ref.rhs(np.asfortranarray(U + eps_fd * v))
        if fd_order == 1:
            Jv = (Rp - R0) / eps_fd
        else:
            Rm = ref.rhs(np.asfortranarray(U - eps_fd * v))
            Jv = (Rp - Rm) / (2.0 * eps_fd)
        a, b = float(np.sum(lam * Jv)), float(np.sum(Ub * v))
        say(f"  dir {it}: lam.Jv={a:.8e} Ubar.v={b:.8e} rel {abs(a - b) / abs(b):.2e}")
        assert abs(a - b) <= 1e-6 * abs(b)
    op.close()


def test_vjp_linear_and_stateless():
    """Repeated and combined calls: VJP(a l1 + b l2) = a VJP(l1) + b VJP(l2) to rounding, and a call does not
    depend on earlier ones (no accumulation left in the device buffers)."""
    say("case: vjp linearity (2,3,2) N=5")
    m, U = noisy_case((2, 3, 2), 5, seed=4)
    op = gpu_op(m)
    op.set_state(U)
    rng = np.random.default_rng(11)
    l1 = np.asfortranarray(rng.standard_normal(U.shape))
    l2 = np.asfortranarray(rng.standard_normal(U.shape))
    g1 = op.rhs_vjp(l1)
    g2 = op.rhs_vjp(l2)
    g12 = op.rhs_vjp(np.asfortranarray(2.0 * l1 - 3.0 * l2))
    assert np.array_equal(op.rhs_vjp(l1), g1)
    err = np.abs(g12 - (2.0 * g1 - 3.0 * g2)).max()
    assert err <= 1e-12 * np.abs(g12).max()
    op.close()


C4_GAS = dict(kappa=1.4, R=1.0, Pr=0.72, mu0=1e-3)


def c4_case(kind):
    """16^3 elements, N=5, curved periodic box.  'smooth': [-1,1]^3 with the 0.05 sine mapping and the seeded
    smooth state (traces continuous across faces); 'tgv': [0,2pi]^3 with the same mapping scaled to the box,
    Taylor-Green state with 1e-2 nodal velocity noise (velocity jumps, density and pressure continuous: the LF
    sound-speed switch sits on its tie at every face node)."""
    if kind == "smooth":
        lo, hi = (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0)
        m = M.box_mesh((16, 16, 16), 5, lo=lo, hi=hi, deform_amp=0.05, deform_phase=(0.31, 0.17, 0.83))
        return m, T.smooth_random(m.Elem_xGP, lo, hi, seed=44), C4_GAS
    lo, hi = (0.0, 0.0, 0.0), (2 * np.pi,) * 3
    m = M.box_mesh((16, 16, 16), 5, lo=lo, hi=hi, deform_amp=0.05, deform_phase=(0.1, 0.2, 0.3))
    return m, T.taylor_green(m.Elem_xGP, Ma=0.1, noise=1e-2, seed=16), dict(kappa=1.4, R=71.42857, Pr=0.72,
                                                                              mu0=1.0 / 1600.0)


@pytest.mark.parametrize("kind", ["smooth", "tgv"])
def test_vjp_16cubed_n5_curved_ten_directions(kind):
    """16^3 elements, N=5, curved mesh, i.i.d. N(0,1) cotangent: ten seeded unit directions against central
    differences of the GPU forward operator (itself within 1e-12 of the oracle), eps in 1e-4..1e-8: the best
    relative error <= 1e-6 for each direction."""
    say(f"case: vjp 16^3 N=5 curved {kind}, 10 directions")
    m, U, gas = c4_case(kind)
    op = dgmod().DGOperator(5, form="split", riemann="lf", precision="fp64", **gas)
    op.set_mesh(m)
    op.set_state(U)
    rng = np.random.default_rng(16)
    lam = np.asfortranarray(rng.standard_normal(U.shape))
    Ub = op.rhs_vjp(lam)
    fwd = dgmod().DGOperator(5, form="split", riemann="lf", precision="fp64", **gas)
    fwd.set_mesh(m)

    def f(X):
        fwd.set_state(np.asfortranarray(X))
        return float(np.sum(lam * fwd.DGTimeDerivative_weakForm(0.0)))

    for it in range(10):
        v = np.asfortranarray(rng.standard_normal(U.shape))
        v /= np.linalg.norm(v.ravel())
        exact = float(np.sum(Ub * v))
        best, errs = best_fd_error(f, U, v, exact, [1e-4, 1e-5, 1e-6, 1e-7, 1e-8])
        say(f"  dir {it}: Ubar.v={exact:.6e} rel errs {['%.1e' % e for e in errs]}")
        assert best <= 1e-6
    op.close()
    fwd.close()


def lserk_case():
    m, U = noisy_case((3, 3, 2), 3, seed=8, noise=5e-3)
    return m, U


def test_lserk_adjoint_vs_fd():
    """J = kinetic energy after 12 LSERK steps (checkpoints every 5: segments 5,5,2); dJ/dU0 . v against
    central differences of the GPU time integration: best over eps <= 1e-6 for three directions."""
    say("case: lserk adjoint vs FD, (3,3,2) N=3, 12 steps, ckpt 5")
    m, U0 = lserk_case()
    op = gpu_op(m)
    op.set_state(U0)
    dt, err = op.CalcTimeStep()
    assert err == 0
    nSteps = 12
    op.forward_ckpt(0.0, dt, nSteps, 5)
    UN = op.get_state()
    g0 = op.lserk_adjoint(kinetic_energy_grad(m, UN))
    assert np.array_equal(op.get_state(), UN)  # context left at the final state
    fwd = gpu_op(m)

    def J(X):
        fwd.set_state(np.asfortranarray(X))
        fwd.TimeStepByLSERKW2(0.0, dt, nSteps)
        return kinetic_energy(m, fwd.get_state())

    rng = np.random.default_rng(12)
    for it in range(3):
        v = direction(U0, rng)
        exact = float(np.sum(g0 * v))
        best, errs = best_fd_error(J, U0, v, exact, [1e-3, 1e-4, 1e-5, 1e-6])
        say(f"  dir {it}: dJ.v={exact:.6e} rel errs {['%.1e' % e for e in errs]}")
        assert best <= 1e-6
    op.close()
    fwd.close()


def test_lserk_adjoint_checkpoint_spacing_invariant():
    """The reverse sweep recomputes segments from checkpoints: the result is bitwise the same for every
    checkpoint spacing, and a second sweep over the same trajectory repeats the first."""
    say("case: lserk adjoint ckpt spacing 1/4/12")
    m, U0 = lserk_case()
    op = gpu_op(m)
    op.set_state(U0)
    dt, _ = op.CalcTimeStep()
    rng = np.random.default_rng(2)
    Ubar_final = np.asfortranarray(rng.standard_normal(U0.shape))
    res = []
    for every in (1, 4, 12):
        op.set_state(U0)
        op.forward_ckpt(0.0, dt, 12, every)
        res.append(op.lserk_adjoint(Ubar_final))
    res.append(op.lserk_adjoint(Ubar_final))
    for r in res[1:]:
        assert np.array_equal(r, res[0])
    op.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_lserk_adjoint_step_storage_equals_stage_storage(precision):
    """GALX_ADJ_STEPS keeps a segment's step inputs and recomputes each step's stages (and, in both forms, the
    FP64 trace records) before its VJPs: the same VJPs on the same stage inputs, so the gradient is bitwise the
    GALX_ADJ_STAGES one, for segments of several steps, of one step and a short last segment, with less device
    memory; the context ends at the final state, and a mode change discards the recording."""
    from galaexi_amd._lib import GalxError

    say(f"case: adjoint step storage vs stage storage, {precision}, (3,3,2) N=3, 12 steps, ckpt 5/1/12")
    m, U0 = lserk_case()
    op = dgmod().DGOperator(3, form="split", riemann="lf", precision=precision, **GAS)
    op.set_mesh(m)
    op.set_state(U0)
    dt, _ = op.CalcTimeStep()
    rng = np.random.default_rng(3)
    Ubar_final = np.asfortranarray(rng.standard_normal(U0.shape))
    for every in (5, 1, 12):
        res, mem = {}, {}
        for mode in ("stages", "steps"):
            op.set_adjoint_storage(mode)
            op.set_state(U0)
            op.forward_ckpt(0.0, dt, 12, every)
            UN = op.get_state()
            mem[mode] = op.memory_bytes()
            res[mode] = op.lserk_adjoint(Ubar_final)
            assert np.array_equal(op.get_state(), UN)
        say(f"  ckpt {every}: device memory stages {mem['stages'] / 1e6:.1f} MB, steps {mem['steps'] / 1e6:.1f} MB")
        assert np.array_equal(res["steps"], res["stages"])
        if every > 1:
            assert mem["steps"] < mem["stages"]
    op.set_adjoint_storage("stages")
    with pytest.raises(GalxError, match="trajectory"):
        op.lserk_adjoint(Ubar_final)
    with pytest.raises(ValueError):
        op.set_adjoint_storage("none")
    op.close()


def test_autograd_wrappers():
    """torch.autograd wrappers: gradcheck (fast mode: random projections u^T J v against finite differences) of
    the RHS on a small mesh, and the LSERK wrapper's gradient equal to the direct ABI result."""
    torch = pytest.importorskip("torch")
    from galaexi_amd import adjoint as A

    say("case: autograd wrappers (2,2,2) N=2")
    m, U = noisy_case((2, 2, 2), 2, seed=30)
    op = gpu_op(m)
    n = m.N + 1
    Ut_layout = np.ascontiguousarray(np.transpose(U, (4, 3, 2, 1, 0)))  # (e,k,j,i,v) == U(5,i,j,k,e) in F order
    X = torch.tensor(Ut_layout, dtype=torch.float64, device="cuda", requires_grad=True)
    assert X.shape == (m.nElems, n, n, n, 5)
    assert torch.autograd.gradcheck(lambda x: A.rhs(x, op), (X,), eps=1e-6, atol=1e-6, rtol=1e-5, fast_mode=True)
    # RHS forward equals the ABI
    Ut = A.rhs(X, op).detach().cpu().numpy()
    op.set_state(U)
    assert np.array_equal(np.transpose(Ut, (4, 3, 2, 1, 0)), op.DGTimeDerivative_weakForm())
    # LSERK wrapper
    op.set_state(U)
    dt, _ = op.CalcTimeStep()
    Y = A.lserk(X, op, dt, 6, 4)
    w = torch.tensor(np.ascontiguousarray(np.transpose(mass_weights(m), (3, 2, 1, 0))), device="cuda")
    J = (w * 0.5 * (Y[..., 1] ** 2 + Y[..., 2] ** 2 + Y[..., 3] ** 2) / Y[..., 0]).sum()
    J.backward()
    op2 = gpu_op(m)
    op2.set_state(U)
    op2.forward_ckpt(0.0, dt, 6, 2)
    g = op2.lserk_adjoint(kinetic_energy_grad(m, op2.get_state()))
    gt = np.transpose(X.grad.cpu().numpy(), (4, 3, 2, 1, 0))
    assert np.abs(gt - g).max() <= 1e-13 * np.abs(g).max()
    op.close()
    op2.close()


def test_adjoint_api_errors():
    dg = dgmod()
    from galaexi_amd._lib import GalxError

    m, U = noisy_case((2, 2, 2), 3, seed=0, curved=False, rot=False)
    for kw in (dict(form="weak", riemann="lf", precision="fp64"), dict(form="split", riemann="roe", precision="fp64"),
               dict(form="split", riemann="roeentropyfix", precision="mixed")):
        op = dg.DGOperator(3, **kw, **GAS)
        op.set_mesh(m)
        op.set_state(U)
        with pytest.raises(GalxError, match="adjoint"):
            op.rhs_vjp(U)
        op.close()
    op = gpu_op(m)
    op.set_state(U)
    with pytest.raises(GalxError, match="trajectory"):
        op.lserk_adjoint(U)
    with pytest.raises(GalxError, match="bad step"):
        op.forward_ckpt(0.0, 1e-3, 0, 1)
    op.close()


def test_mixed_context_vjp_equals_fp64():
    """A mixed-precision context runs the same FP64 VJP kernels on FP64 trace records of its state: bitwise the
    FP64 context's VJP at the same U."""
    say("case: vjp mixed context vs fp64 context, (3,2,2) N=3 curved rotated")
    m, U = noisy_case((3, 2, 2), 3, seed=21)
    lam = np.asfortranarray(np.random.default_rng(4).standard_normal(U.shape))
    res = []
    for prec in ("fp64", "mixed"):
        op = dgmod().DGOperator(3, form="split", riemann="lf", precision=prec, **GAS)
        op.set_mesh(m)
        op.set_state(U)
        res.append(op.rhs_vjp(lam))
        op.close()
    assert np.array_equal(res[0], res[1])


def test_mixed_forward_lserk_adjoint():
    """Checkpointed adjoint along a mixed-precision forward (FP32 fluxes, FP64 adjoint accumulation): bitwise
    independent of the checkpoint spacing, and within 1e-5 relative (L2) of the FP64 gradient, the size of the
    mixed trajectory's own deviation."""
    say("case: lserk adjoint along the mixed forward, (3,3,2) N=3, 12 steps")
    m, U0 = lserk_case()
    ref = gpu_op(m)
    ref.set_state(U0)
    dt, _ = ref.CalcTimeStep()
    ref.forward_ckpt(0.0, dt, 12, 5)
    g64 = ref.lserk_adjoint(kinetic_energy_grad(m, ref.get_state()))
    ref.close()
    op = dgmod().DGOperator(3, form="split", riemann="lf", precision="mixed", **GAS)
    op.set_mesh(m)
    res = []
    for every in (1, 5, 12):
        op.set_state(U0)
        op.forward_ckpt(0.0, dt, 12, every)
        UN = op.get_state()
        res.append(op.lserk_adjoint(kinetic_energy_grad(m, UN)))
        assert np.array_equal(op.get_state(), UN)
    for r in res[1:]:
        assert np.array_equal(r, res[0])
    rel = np.linalg.norm((res[0] - g64).ravel()) / np.linalg.norm(g64.ravel())
    say(f"  mixed-forward gradient vs FP64: rel L2 {rel:.2e}")
    assert 0 < rel <= 1e-5
    op.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_step_vjp(precision):
    """galx_step_vjp: bitwise the adjoint of a recorded one-step trajectory, the context kept at its state, and
    (FP64) against central FD of one LSERK step: J(U) = w . step(U), best over eps within 1e-6."""
    say(f"case: step vjp {precision}, (3,3,2) N=3")
    m, U0 = lserk_case()
    rng = np.random.default_rng(17)
    w = np.asfortranarray(rng.standard_normal(U0.shape))
    op = dgmod().DGOperator(3, form="split", riemann="lf", precision=precision, **GAS)
    op.set_mesh(m)
    op.set_state(U0)
    dt, _ = op.CalcTimeStep()
    g = op.step_vjp(0.0, dt, w)
    assert np.array_equal(op.get_state(), U0)
    ref = dgmod().DGOperator(3, form="split", riemann="lf", precision=precision, **GAS)
    ref.set_mesh(m)
    ref.set_state(U0)
    ref.forward_ckpt(0.0, dt, 1, 1)
    assert np.array_equal(g, ref.lserk_adjoint(w))
    # the context steps on from the kept state exactly like a fresh one
    op.TimeStepByLSERKW2(0.0, dt, 1)
    ref.set_state(U0)
    ref.TimeStepByLSERKW2(0.0, dt, 1)
    assert np.array_equal(op.get_state(), ref.get_state())
    if precision == "fp64":
        def J(X):
            ref.set_state(np.asfortranarray(X))
            ref.TimeStepByLSERKW2(0.0, dt, 1)
            return float(np.sum(w * ref.get_state()))

        v = direction(U0, rng)
        exact = float(np.sum(g * v))
        best, errs = best_fd_error(J, U0, v, exact, [1e-3, 1e-4, 1e-5, 1e-6])
        say(f"  dJ.v={exact:.6e} rel errs {['%.1e' % e for e in errs]}")
        assert best <= 1e-6
    op.close()
    ref.close()


def test_adjoint_discarded_by_state_changes():
    """A recorded trajectory belongs to the state it ends at: set_state, a step, a stage or set_source afterwards
    discard it, and the sweep then reports GALX_E_STATE instead of sweeping a stale trajectory."""
    from galaexi_amd._lib import GalxError

    say("case: stale trajectory refused, (2,2,2) N=3")
    m, U = noisy_case((2, 2, 2), 3, seed=3)
    op = gpu_op(m)
    op.set_state(U)
    dt, _ = op.CalcTimeStep()
    for change in (lambda: op.set_state(U), lambda: op.TimeStepByLSERKW2(0.0, dt, 1),
                   lambda: op.lserk_stage(0.0, dt, 1), lambda: op.set_source(0)):
        op.set_state(U)
        op.forward_ckpt(0.0, dt, 3, 2)
        g = op.lserk_adjoint(U)  # the recording is usable, and stays so after a sweep
        assert np.array_equal(op.lserk_adjoint(U), g)
        change()
        with pytest.raises(GalxError, match="trajectory"):
            op.lserk_adjoint(U)
    op.close()


# C5 workload, reduced: curved Taylor-Green vortex (Re 1600, Ma 0.1) on [0,2pi]^3 with the 0.05 sine mapping, N=7,
# 4^3 elements, 30 LSERK steps, checkpoints every 5; J = int 1/2 rho |u|^2 dV at the end.
C5_GAS = dict(kappa=1.4, R=71.42857, Pr=0.72, mu0=1.0 / 1600.0, CFLScale=0.5, DFLScale=0.5)


def c5_case(ne=4):
    lo, hi = (0.0, 0.0, 0.0), (2 * np.pi,) * 3
    m = M.box_mesh((ne, ne, ne), 7, lo=lo, hi=hi, deform_amp=0.05, deform_phase=(0.31, 0.17, 0.83))
    return m, T.taylor_green(m.Elem_xGP, Ma=0.1, noise=1e-3, seed=1600)


def c5_op(m, precision):
    op = dgmod().DGOperator(7, form="split", riemann="lf", precision=precision, **C5_GAS)
    op.set_mesh(m)
    return op


def well_conditioned_direction(g, seed, scale=2.0):
    """scale * (g/|g| + 0.3 r/|r|), r seeded N(0,1): dJ.v is about scale*|g|, a sizeable part of J, so the central
    difference is not lost to cancellation in J (a random direction alone gives dJ.v ~ 1e-5 J)."""
    rng = np.random.default_rng(seed)
    r = rng.standard_normal(g.shape)
    return np.asfortranarray(scale * (g / np.linalg.norm(g.ravel()) + 0.3 * r / np.linalg.norm(r.ravel())))


def consecutive_pass(errs, bar):
    return any(errs[i] <= bar and errs[i + 1] <= bar for i in range(len(errs) - 1))


def test_c5_reduced_n7_adjoint_vs_fd():
    """The N=7 adjoint kernels on the C5 workload at 4^3: the FP64 checkpointed adjoint against central
    differences of the FP64 time integration along a well-conditioned direction (dJ.v >= 1e-2 J), relative error
    <= 1e-6 at two consecutive step sizes; then the C5 setting itself, the mixed-precision forward with the FP64
    adjoint accumulated along it: bitwise independent of the checkpoint spacing, and within 1e-5 of the FP64
    gradient and of the FP64 finite difference (the mixed trajectory's own deviation, which bounds how far the
    two gradients can agree)."""
    say("case: C5 reduced, curved TGV 4^3 N=7, 30 steps, ckpt 5, fp64 and mixed forward")
    m, U0 = c5_case()
    W = mass_weights(m)
    steps, every = 30, 5
    op = c5_op(m, "fp64")
    op.set_state(U0)
    dt, err = op.CalcTimeStep()
    assert err == 0
    op.forward_ckpt(0.0, dt, steps, every)
    UN = op.get_state()
    J0 = kinetic_energy(m, UN)
    g = op.lserk_adjoint(kinetic_energy_grad(m, UN))
    v = well_conditioned_direction(g, 5)
    exact = float(np.sum(g * v))
    say(f"  J={J0:.10e} dJ.v={exact:.6e} (dJ.v/J = {exact / J0:.2e})")
    assert abs(exact) >= 1e-2 * J0

    def J(X):
        op.set_state(np.asfortranarray(X))
        op.TimeStepByLSERKW2(0.0, dt, steps)
        return kinetic_energy(m, op.get_state())

    eps_list = [1e-1, 3e-2, 1e-2, 3e-3, 1e-3, 3e-4]
    fd = {}
    for eps in eps_list:
        fd[eps] = (J(U0 + eps * v) - J(U0 - eps * v)) / (2 * eps)
    errs = [abs(fd[e] - exact) / abs(exact) for e in eps_list]
    say(f"  fp64 adjoint vs central FD, eps {eps_list}: rel errs {['%.1e' % e for e in errs]}")
    assert consecutive_pass(errs, 1e-6), errs
    fd_ref = fd[eps_list[int(np.argmin(errs))]]
    op.close()
    # the C5 setting: mixed forward, FP64 adjoint
    mx = c5_op(m, "mixed")
    res = []
    for ck in (every, 1, steps):
        mx.set_state(U0)
        mx.forward_ckpt(0.0, dt, steps, ck)
        UNm = mx.get_state()
        res.append(mx.lserk_adjoint(kinetic_energy_grad(m, UNm)))
    for r in res[1:]:
        assert np.array_equal(r, res[0])
    gm = res[0]
    dev_state = np.linalg.norm((UNm - UN).ravel()) / np.linalg.norm(UN.ravel())
    rel_g = np.linalg.norm((gm - g).ravel()) / np.linalg.norm(g.ravel())
    rel_dir = abs(float(np.sum(gm * v)) - fd_ref) / abs(fd_ref)
    say(f"  mixed forward: final state vs fp64 {dev_state:.1e}; gradient vs fp64 adjoint {rel_g:.1e}; "
        f"dJ.v vs fp64 FD {rel_dir:.1e}")
    assert rel_g <= 1e-5 and rel_dir <= 1e-5
    mx.close()
