"""Device implicit stepping (BDF1, BDF2, Crank-Nicolson with Newton on the
exact Jacobian and a device BiCGSTAB) against the reference's implicit
golds and, for BDF2, against the discrete residual evaluated by the oracle.
"""
import os

import numpy as np
import pytest

import cases
import pressiodemoapps_amd as pda

pytestmark = pytest.mark.gpu
S = pda.InviscidFluxReconstruction
TRAJ = os.path.join(os.path.dirname(__file__), "golden", "ref", "traj")


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    if not _t.cuda.is_available():
        pytest.skip("no HIP device")
    return _t


@pytest.mark.parametrize("scheme,st,name", [(S.FirstOrder, 3, "firstorder"), (S.Weno3, 5, "weno3")])
def test_sod1d_bdf1_matches_reference_gold(scheme, st, name, torch):
    # tests_cpp/eigen_1d_euler_sod_implicit: 100 cells, BDF1, dt 1e-3 x 100,
    # pressio Newton tolerance 1e-6 (correction rule), rtol 1e-9 / atol 1e-11
    print(f"[gpu-implicit] sod1d bdf1 {name}", flush=True)
    mesh = pda.create_structured_mesh((100,), (-0.5, 0.5), st)
    prob = pda.create_problem(mesh, pda.Euler1d.Sod, scheme)
    ip = pda.ImplicitStepper(prob, pda.ImplicitScheme.BDF1)
    ip.set_options(pda.StopRule.Correction, 1e-6)
    U = prob.initialCondition()
    iters = []
    for k in range(100):
        iters.append(ip.step(U, k * 1e-3, 1e-3)["newton_iters"])
    gold = np.load(os.path.join(TRAJ, f"eigen_1d_euler_sod_implicit__{name}__gold.npy"))
    Uh = U.cpu().numpy()
    print(f"[gpu-implicit] sod1d {name}: newton iters/step {np.mean(iters):.2f}, "
          f"max |U-gold| {np.abs(Uh - gold).max():.2e}", flush=True)
    assert np.allclose(Uh, gold, rtol=1e-9, atol=1e-11)


def test_riemann2d_crank_nicolson_matches_reference_gold(torch):
    # tests_cpp/eigen_2d_euler_riemann_implicit/firstorder: 20x20, IC2, CN
    # dt 0.02 x 30, Newton tolerance 1e-5
    print("[gpu-implicit] riemann2d cn", flush=True)
    mesh = pda.create_structured_mesh((20, 20), (0.0, 1.0, 0.0, 1.0), 3)
    prob = pda.create_problem(mesh, pda.Euler2d.Riemann, S.FirstOrder, 2)
    ip = pda.ImplicitStepper(prob, pda.ImplicitScheme.CrankNicolson)
    ip.set_options(pda.StopRule.Correction, 1e-5)
    U = prob.initialCondition()
    for k in range(30):
        ip.step(U, k * 0.02, 0.02)
    Uh = U.cpu().numpy().reshape(-1, 4)
    rho = Uh[:, 0]
    p = 0.4 * (Uh[:, 3] - 0.5 * (Uh[:, 1] ** 2 + Uh[:, 2] ** 2) / rho)
    base = "eigen_2d_euler_riemann_implicit__firstorder__"
    assert np.allclose(rho, np.load(os.path.join(TRAJ, base + "rho_gold.npy")), rtol=1e-10, atol=1e-12)
    assert np.allclose(p, np.load(os.path.join(TRAJ, base + "p_gold.npy")), rtol=1e-10, atol=1e-12)


def test_burgers_bdf2_residual_rule(oracle, torch):
    """BDF2 on periodic Burgers (WENO5, exact WENO5 Jacobian), Newton to
    ||R|| <= 1e-10; the returned states satisfy the BDF2 residual evaluated
    with the oracle's velocity (parity of BDF2 itself is unpinned: no
    reference test uses it)."""
    print("[gpu-implicit] burgers bdf2", flush=True)
    case = [c for c in cases.CASES if c[0] == "burgers-periodic-Weno5"][0]
    mesh, prob = cases.make(case)
    ms = oracle.from_mesh(mesh)
    pb = cases.oracle_problem(case)
    ip = pda.ImplicitStepper(prob, pda.ImplicitScheme.BDF2)
    ip.set_options(pda.StopRule.Residual, 1e-10)
    dt = 5e-3
    U = prob.initialCondition()
    Uprev = None
    for k in range(6):
        y_n = U.clone()
        st = ip.step(U, k * dt, dt, prev=Uprev)
        assert st["converged"] and st["residual_norm"] <= 1e-10
        y = U.cpu().numpy()
        f = oracle.velocity_and_jacobian(ms, pb, y)[0]
        if Uprev is None:
            R = y - y_n.cpu().numpy() - dt * f
        else:
            R = y - (4. / 3.) * y_n.cpu().numpy() + (1. / 3.) * Uprev.cpu().numpy() - (2. / 3.) * dt * f
        print(f"[gpu-implicit] bdf2 step {k}: newton {st['newton_iters']} linear {st['linear_iters']} "
              f"||R|| dev {st['residual_norm']:.2e} oracle {np.linalg.norm(R):.2e}", flush=True)
        assert np.linalg.norm(R) <= 2e-10
        Uprev = y_n


@pytest.mark.parametrize("scheme,st,name", [(S.FirstOrder, 3, "firstorder"), (S.Weno3, 5, "weno3"),
                                            (S.Weno5, 7, "weno5")])
def test_adr_crank_nicolson_matches_reference_gold(scheme, st, name, torch):
    # tests_cpp/eigen_2d_advdiffreac_probA_implicit: 17x17 on [0,1]^2, zero
    # initial state, CN dt 0.05 x 40, Newton tolerance 1e-5
    print(f"[gpu-implicit] adr cn {name}", flush=True)
    mesh = pda.create_structured_mesh((17, 17), (0.0, 1.0, 0.0, 1.0), st)
    prob = pda.create_problem(mesh, pda.AdvectionDiffusionReaction2d.ProblemA, scheme)
    ip = pda.ImplicitStepper(prob, pda.ImplicitScheme.CrankNicolson)
    ip.set_options(pda.StopRule.Correction, 1e-5)
    U = prob.initialCondition()
    for k in range(40):
        ip.step(U, k * 0.05, 0.05)
    gold = np.load(os.path.join(TRAJ, f"eigen_2d_advdiffreac_probA_implicit__{name}__gold.npy"))
    Uh = U.cpu().numpy()
    print(f"[gpu-implicit] adr {name}: max |U-gold| {np.abs(Uh - gold).max():.2e}", flush=True)
    assert np.allclose(Uh, gold, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("scheme,st,name", [(S.FirstOrder, 3, "firstorder"), (S.Weno3, 5, "weno3")])
def test_burgers_outflow_crank_nicolson_matches_reference_gold(scheme, st, name, torch):
    # tests_cpp/eigen_2d_burgers_outflow_implicit: 20x20 on [-1,1]^2, default
    # pulse and diffusion, CN dt 0.01 x 200, Newton tolerance 1e-5
    print(f"[gpu-implicit] burgers outflow cn {name}", flush=True)
    mesh = pda.create_structured_mesh((20, 20), (-1.0, 1.0, -1.0, 1.0), st)
    prob = pda.create_problem(mesh, pda.AdvectionDiffusion2d.BurgersOutflow, scheme, 1)
    ip = pda.ImplicitStepper(prob, pda.ImplicitScheme.CrankNicolson)
    ip.set_options(pda.StopRule.Correction, 1e-5)
    U = prob.initialCondition()
    for k in range(200):
        ip.step(U, k * 0.01, 0.01)
    gold = np.load(os.path.join(TRAJ, f"eigen_2d_burgers_outflow_implicit__{name}__gold.npy"))
    Uh = U.cpu().numpy()
    print(f"[gpu-implicit] burgers outflow {name}: max |U-gold| {np.abs(Uh - gold).max():.2e}", flush=True)
    assert np.allclose(Uh, gold, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("stop", ["residual", "correction"])
def test_nonfinite_state_leaves_yn(stop, torch):
    """A NaN in y_n makes the residual non-finite: the step fails with
    PDA_E_STATE and the state still holds y_n (pressio's
    NonlinearSolveFailure, after which its stepper restores y_n,
    ode_implicit_stepper_standard.hpp:213-225)."""
    print(f"[gpu-implicit] non-finite state, {stop} rule", flush=True)
    mesh = pda.create_structured_mesh((32, 32), (-1.0, 1.0, -1.0, 1.0), 7, True)
    prob = pda.create_problem(mesh, pda.AdvectionDiffusion2d.BurgersPeriodic,
                              pda.InviscidFluxReconstruction.Weno5, 1, {"diffusion": 0.005})
    ip = pda.ImplicitStepper(prob, pda.ImplicitScheme.BDF1)
    rule = pda.StopRule.Residual if stop == "residual" else pda.StopRule.Correction
    ip.set_options(rule, 1e-10, 30, 1e-12, 0)
    U = prob.initialCondition()
    U[5] = float("nan")
    U0 = U.clone()
    with pytest.raises(pda.PdaError, match="not finite"):
        ip.step(U, 0.0, 5e-3)
    torch.cuda.synchronize()
    assert torch.equal(torch.isnan(U), torch.isnan(U0))
    assert torch.equal(torch.nan_to_num(U), torch.nan_to_num(U0))
