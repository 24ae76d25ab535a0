"""Explicit time integration on the device vs the CPU oracle.

Sod shock tube (1000 cells on [-0.5, 0.5], WENO5, SSPRK3, dt = 1e-4 to
t = 0.2, i.e. 2000 steps): the device velocity is checked against the oracle
on the device's own state at every step (1e-12), the graph-replayed stepper
against the eagerly launched one at every step (bit-exact), and the final
state against the oracle's own SSPRK3 run with pressio's stage formulas
(tests_cpp/pressio/include/pressio/ode/impl/ode_explicit_stepper_without_mass_matrix.hpp:230-281).
"""
import numpy as np
import pytest

import cases
import pressiodemoapps_amd as pda

pytestmark = pytest.mark.gpu
S = pda.InviscidFluxReconstruction


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    if not _t.cuda.is_available():
        pytest.skip("no HIP device")
    return _t


def ssprk3_cpu(f, y, dt, n):
    for _ in range(n):
        a = y + dt * f(y)
        a = 0.25 * a + 0.75 * y + (0.25 * dt) * f(a)
        y = (1. / 3.) * y + (2. / 3.) * a + (2. / 3. * dt) * f(a)
    return y


def rk4_cpu(f, y, dt, n):
    for _ in range(n):
        k1 = f(y)
        k2 = f(y + (dt / 2) * k1)
        k3 = f(y + (dt / 2) * k2)
        k4 = f(y + dt * k3)
        y = y + (dt / 6) * k1 + (dt / 3) * k2 + (dt / 3) * k3 + (dt / 6) * k4
    return y


def test_sod1d_weno5_ssprk3_2000_steps(oracle, torch):
    print("[gpu-step] sod1d weno5 ssprk3 1000 cells 2000 steps", flush=True)
    mesh = pda.create_structured_mesh((1000,), (-0.5, 0.5), 7)
    prob = pda.create_problem(mesh, pda.Euler1d.Sod, S.Weno5)
    ms = oracle.from_mesh(mesh)
    import pyoracle
    pb = pyoracle.Oracle.problem(1, int(pda.Euler1d.Sod), int(S.Weno5), [1.4])
    f = lambda U: oracle.velocity(ms, pb, U)  # noqa: E731
    U0 = prob.initialCondition()
    Ue, Ug = U0.clone(), U0.clone()
    eager = pda.ExplicitStepper(prob, pda.StepScheme.SSPRK3)
    graph = pda.ExplicitStepper(prob, pda.StepScheme.SSPRK3)
    V = prob.createRightHandSide()
    dt, nsteps = 1e-4, 2000
    worst_v = 0.0
    for k in range(nsteps):
        prob.rightHandSide(Ue, k * dt, V)
        torch.cuda.synchronize()
        Uh = Ue.cpu().numpy()
        ev = cases.rel_err(V.cpu().numpy(), f(Uh))
        worst_v = max(worst_v, ev)
        assert ev <= 1e-12, (k, ev)
        eager.advance(Ue, dt, 1, k * dt, use_graph=False)
        graph.advance(Ug, dt, 1, k * dt, use_graph=True)
        torch.cuda.synchronize()
        assert torch.equal(Ue, Ug), k
    Ucpu = ssprk3_cpu(f, U0.cpu().numpy(), dt, nsteps)
    err = cases.rel_err(Ug.cpu().numpy(), Ucpu)
    print(f"[gpu-step] sod1d: worst per-step V {worst_v:.2e}, final state vs oracle run {err:.2e}",
          flush=True)
    assert err <= 1e-10


@pytest.mark.parametrize("case", [c for c in cases.CASES if c[0] in (
    "euler2d-periodic-Weno5", "swe-slipwall-Weno3", "burgers-periodic-Weno5",
    "euler2d-riemann-FirstOrder", "euler3d-periodic-Weno3")], ids=lambda c: c[0])
def test_rk4_graph_matches_eager_and_oracle(case, oracle, torch):
    print(f"[gpu-step] rk4 {case[0]}", flush=True)
    mesh, prob = cases.make(case)
    dt, nsteps = 1e-3, 60
    U0 = torch.from_numpy(cases.perturbed_state(prob, 2)).cuda()
    Ue, Ug = U0.clone(), U0.clone()
    pda.ExplicitStepper(prob, pda.StepScheme.RK4).advance(Ue, dt, nsteps, use_graph=False)
    # graph replayed in chunks with eager work in between
    st = pda.ExplicitStepper(prob, pda.StepScheme.RK4)
    V = prob.createRightHandSide()
    for chunk in (1, 19, 40):
        st.advance(Ug, dt, chunk, use_graph=True)
        prob.rightHandSide(Ug, 0.0, V)
        torch.cuda.synchronize()
    assert torch.equal(Ue, Ug)
    ms = oracle.from_mesh(mesh)
    pb = cases.oracle_problem(case)
    Ucpu = rk4_cpu(lambda U: oracle.velocity(ms, pb, U), U0.cpu().numpy(), dt, nsteps)
    err = cases.rel_err(Ug.cpu().numpy(), Ucpu)
    print(f"[gpu-step] {case[0]}: final state vs oracle run {err:.2e}", flush=True)
    assert err <= 1e-11


def test_rk2_and_observer(oracle, torch):
    mesh, prob = cases.make([c for c in cases.CASES if c[0] == "euler1d-periodic-Weno3"][0])
    U = prob.initialCondition()
    seen = []
    pda.advanceRK2(prob, U, 1e-3, 5, observer=lambda k, s, v: seen.append(k))
    assert seen == [0, 1, 2, 3, 4]
    ms = oracle.from_mesh(mesh)
    pb = cases.oracle_problem([c for c in cases.CASES if c[0] == "euler1d-periodic-Weno3"][0])
    f = lambda y: oracle.velocity(ms, pb, y)  # noqa: E731
    y = prob.initialCondition().cpu().numpy()
    for _ in range(5):
        k1 = 1e-3 * f(y)
        k2 = 1e-3 * f(y + k1)
        y = y + k2 * 0.5 + k1 * 0.5
    torch.cuda.synchronize()
    assert cases.rel_err(U.cpu().numpy(), y) <= 1e-12


@pytest.mark.parametrize("scheme,st,name", [(S.FirstOrder, 3, "firstorder"), (S.Weno3, 5, "weno3"),
                                            (S.Weno5, 7, "weno5")])
def test_adr_rk4_matches_reference_gold(scheme, st, name, torch):
    # tests_cpp/eigen_2d_advdiffreac_probA_explicit: 18x18 on [0,1]^2, zero
    # initial state, default parameters, RK4 dt 0.01 x 500 (graph-replayed)
    import os
    print(f"[gpu-step] adr rk4 {name}", flush=True)
    mesh = pda.create_structured_mesh((18, 18), (0.0, 1.0, 0.0, 1.0), st)
    prob = pda.create_problem(mesh, pda.AdvectionDiffusionReaction2d.ProblemA, scheme)
    U = prob.initialCondition()
    pda.ExplicitStepper(prob, pda.StepScheme.RK4).advance(U, 0.01, 500)  # pressio's RK4, as the C++ test
    gold = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref", "traj",
                                f"eigen_2d_advdiffreac_probA_explicit__{name}__gold.npy"))
    assert np.allclose(U.cpu().numpy(), gold, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("scheme", ["ForwardEuler", "SSPRK3", "RK4", "RK2", "ModuleRK4", "ModuleSSP3"])
def test_stepper_matches_stage_formulas(scheme, torch):
    """The device stepper on a problem evaluated by the generic row kernel
    (1D here, where SSPRK3 and RK4 stages are formed in the velocity
    launches) equals the stage formulas evaluated through rightHandSide and
    separate vector updates (pressio's forward Euler, SSPRK3 and RK4
    stages, ode_explicit_stepper_without_mass_matrix.hpp:168-352; RK2, RK4
    and SSP3 as pressiodemoapps/__init__.py:90-185, whose algebra the
    device reproduces bit for bit), 40 steps of Sod on 1000 cells."""
    print(f"[gpu-step] stepper vs stage formulas, {scheme}", flush=True)
    mesh = pda.create_structured_mesh((1000,), (-0.5, 0.5), 7)
    prob = pda.create_problem(mesh, pda.Euler1d.Sod, S.Weno5)
    dt, nsteps = 1e-4, 40
    U = prob.initialCondition()
    y = U.clone()
    Ug = U.clone()
    pda.ExplicitStepper(prob, getattr(pda.StepScheme, scheme)).advance(U, dt, nsteps, use_graph=False)
    pda.ExplicitStepper(prob, getattr(pda.StepScheme, scheme)).advance(Ug, dt, nsteps, use_graph=True)

    def f(x):
        v = prob.createRightHandSide()
        prob.rightHandSide(x, 0.0, v)
        return v

    for _ in range(nsteps):
        if scheme == "ForwardEuler":
            y = y + dt * f(y)
        elif scheme == "SSPRK3":
            a = y + dt * f(y)
            a = 0.25 * a + 0.75 * y + (0.25 * dt) * f(a)
            y = (1.0 / 3.0) * y + (2.0 / 3.0) * a + ((2.0 / 3.0) * dt) * f(a)
        elif scheme == "RK4":
            k1 = f(y)
            k2 = f(y + (dt / 2) * k1)
            k3 = f(y + (dt / 2) * k2)
            k4 = f(y + dt * k3)
            y = y + (dt / 6) * k1 + (dt / 3) * k2 + (dt / 3) * k3 + (dt / 6) * k4
        elif scheme == "RK2":
            # pressiodemoapps/__init__.py:90-114, one rounding per operation
            k1 = dt * f(y)
            k2 = dt * f(y + k1)
            y = y + k2 * 0.5 + k1 * 0.5
        elif scheme == "ModuleRK4":
            # pressiodemoapps/__init__.py:117-152
            k1 = dt * f(y)
            k2 = dt * f(y + 0.5 * k1)
            k3 = dt * f(y + 0.5 * k2)
            k4 = dt * f(y + k3)
            y = y + (k1 + 2. * k2 + 2. * k3 + k4) * (1. / 6.)
        else:
            # pressiodemoapps/__init__.py:155-185
            y1 = y + dt * f(y)
            y2 = (3. / 4.) * y + (1. / 4.) * y1 + ((1. / 4.) * dt) * f(y1)
            y = (1. / 3.) * (y + 2. * y2 + (2. * dt) * f(y2))
    torch.cuda.synchronize()
    assert torch.equal(U, Ug)  # graph replay == eager
    err = cases.rel_err(U.cpu().numpy(), y.cpu().numpy())
    print(f"[gpu-step] {scheme}: vs stage formulas {err:.2e}", flush=True)
    if scheme in ("RK2", "ModuleRK4", "ModuleSSP3"):
        # the module's algebra, rounded as torch (and numpy) round it: bit for bit
        assert torch.equal(U, y), err
    else:
        assert err <= 1e-13


ONE_D = [
    ("sod-Weno5", pda.Euler1d.Sod, S.Weno5, (200,), (-0.5, 0.5), 7, (), 0, None),
    ("lax-Weno3", pda.Euler1d.Lax, S.Weno3, (150,), (-5.0, 5.0), 5, (), 0, None),
    ("shuosher-FirstOrder", pda.Euler1d.ShuOsher, S.FirstOrder, (130,), (-5.0, 5.0), 3, (), 0, None),
    ("periodic-Weno5", pda.Euler1d.PeriodicSmooth, S.Weno5, (40,), (-1.0, 1.0), 7, True, 0, None),
    ("periodic-Weno3-7", pda.Euler1d.PeriodicSmooth, S.Weno3, (7,), (-1.0, 1.0), 5, True, 0, None),
]


@pytest.mark.parametrize("scheme", ["SSPRK3", "RK4"])
@pytest.mark.parametrize("case", ONE_D, ids=[c[0] for c in ONE_D])
def test_1d_steps_per_launch(case, scheme, oracle, torch):
    """1D SSPRK3 / RK4 advance several steps per launch: 20 steps in one
    call (8 + 8 + 4 per launch), in 20 calls of one step and in calls of 3
    and 17 give the same states bit for bit, and match the oracle's own
    integration with pressio's stage formulas."""
    print(f"[gpu-step] 1d {case[0]} {scheme}", flush=True)
    mesh, prob = cases.make(case)
    sch = getattr(pda.StepScheme, scheme)
    dt, nsteps = 2e-4, 20
    U0 = torch.from_numpy(cases.perturbed_state(prob, 5)).cuda()
    Ua, Ub, Uc = U0.clone(), U0.clone(), U0.clone()
    pda.ExplicitStepper(prob, sch).advance(Ua, dt, nsteps)
    sb = pda.ExplicitStepper(prob, sch)
    for _ in range(nsteps):
        sb.advance(Ub, dt, 1)
    sc = pda.ExplicitStepper(prob, sch)
    sc.advance(Uc, dt, 3)
    sc.advance(Uc, dt, 17)
    torch.cuda.synchronize()
    assert torch.equal(Ua, Ub) and torch.equal(Ua, Uc)
    ms = oracle.from_mesh(mesh)
    pb = cases.oracle_problem(case)
    f = lambda U: oracle.velocity(ms, pb, U)  # noqa: E731
    ref = (ssprk3_cpu if scheme == "SSPRK3" else rk4_cpu)(f, U0.cpu().numpy(), dt, nsteps)
    err = cases.rel_err(Ua.cpu().numpy(), ref)
    print(f"[gpu-step] 1d {case[0]} {scheme}: vs oracle run {err:.2e}", flush=True)
    assert err <= 1e-11
