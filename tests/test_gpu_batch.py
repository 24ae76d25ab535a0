"""Batched evaluation and stepping of independent shallow-water instances
(different grids, gravity, Coriolis and pulse amplitude) in one launch,
against the CPU oracle per instance and against single-instance stepping."""
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


def _instances(scheme, sizes, seed):
    rng = np.random.default_rng(seed)
    out = []
    for n in sizes:
        params = {"gravity": float(rng.uniform(7.5, 12.5)), "coriolis": float(rng.uniform(-4, -2)),
                  "pulseMagnitude": float(rng.uniform(0.05, 0.2))}
        mesh = pda.create_structured_mesh((n, n), (-5, 5, -5, 5), 7)
        prob = pda.create_problem(mesh, pda.Swe2d.SlipWall, scheme, 1, params)
        out.append((mesh, prob, params))
    return out


@pytest.mark.parametrize("scheme", [S.FirstOrder, S.Weno3, S.Weno5], ids=lambda s: s.name)
def test_batched_velocity_matches_oracle_per_instance(scheme, oracle, torch):
    print(f"[gpu-batch] velocity {scheme.name}", flush=True)
    inst = _instances(scheme, (16, 24, 20, 33, 64), 1)
    batch = pda.Batch([p for _, p, _ in inst])
    U0 = batch.initialCondition()
    rng = np.random.default_rng(2)
    U = U0 * torch.from_numpy(1 + 0.01 * rng.uniform(-1, 1, batch.total)).cuda()
    V = batch.createRightHandSide()
    batch.rightHandSide(U, 0.0, V)
    torch.cuda.synchronize()
    import pyoracle
    for i, (mesh, prob, params) in enumerate(inst):
        Ui = batch.instance(U, i).cpu().numpy()
        pb = pyoracle.Oracle.problem(4, int(pda.Swe2d.SlipWall), int(scheme),
                                     [params["gravity"], params["coriolis"]])
        Vo = oracle.velocity(oracle.from_mesh(mesh), pb, Ui)
        e = cases.rel_err(batch.instance(V, i).cpu().numpy(), Vo)
        assert e <= 1e-12, (i, e)
    # the autotuned strip height gives the same values
    rps, us = batch.autotune(U)
    assert rps in (1, 2, 3, 4, 6, 8, 12, 16, 24, 32) and us > 0
    V2 = batch.createRightHandSide()
    batch.rightHandSide(U, 0.0, V2)
    torch.cuda.synchronize()
    assert torch.equal(V, V2)


def test_batched_rk4_matches_single_instances_and_oracle(oracle, torch):
    print("[gpu-batch] rk4", flush=True)
    inst = _instances(S.Weno5, (24, 32, 40), 3)
    batch = pda.Batch([p for _, p, _ in inst])
    U = batch.initialCondition()
    Ue = U.clone()
    dt, nsteps = 0.01, 25
    batch.advance(U, pda.StepScheme.RK4, dt, nsteps, steps_per_graph=10)
    batch.advance(Ue, pda.StepScheme.RK4, dt, nsteps, steps_per_graph=0)
    torch.cuda.synchronize()
    assert torch.equal(U, Ue)  # graph replay == eager
    import pyoracle
    for i, (mesh, prob, params) in enumerate(inst):
        Ui = prob.initialCondition()
        pda.ExplicitStepper(prob, pda.StepScheme.RK4).advance(Ui, dt, nsteps)
        torch.cuda.synchronize()
        assert cases.rel_err(batch.instance(U, i).cpu().numpy(), Ui.cpu().numpy()) <= 1e-13
    # one instance against the oracle's own RK4 run
    mesh, prob, params = inst[0]
    pb = pyoracle.Oracle.problem(4, int(pda.Swe2d.SlipWall), int(S.Weno5),
                                 [params["gravity"], params["coriolis"]])
    ms = oracle.from_mesh(mesh)
    f = lambda y: oracle.velocity(ms, pb, y)  # noqa: E731
    y = prob.initialCondition().cpu().numpy()
    for _ in range(nsteps):
        k1 = f(y)
        k2 = f(y + (dt / 2) * k1)
        k3 = f(y + (dt / 2) * k2)
        k4 = f(y + dt * k3)
        y = y + (dt / 6) * k1 + (dt / 3) * k2 + (dt / 3) * k3 + (dt / 6) * k4
    assert cases.rel_err(batch.instance(U, 0).cpu().numpy(), y) <= 1e-11


@pytest.mark.parametrize("family", ["burgers", "riemann"])
def test_batched_velocity_other_families(family, oracle, torch):
    """Burgers (periodic) and 2D Euler Riemann (mirror ghosts) instances of
    different sizes and parameters in one batched launch, each against the
    oracle."""
    print(f"[gpu-batch] {family}", flush=True)
    rng = np.random.default_rng(4)
    inst = []
    for n in (18, 40, 27, 64):
        if family == "burgers":
            params = {"diffusion": float(rng.uniform(1e-3, 1e-2)), "pulseMagnitude": float(rng.uniform(0.5, 1.0))}
            mesh = pda.create_structured_mesh((n, n), (-1, 1, -1, 1), 7, True)
            prob = pda.create_problem(mesh, pda.AdvectionDiffusion2d.BurgersPeriodic, S.Weno5, 1, params)
            orc = ([params["diffusion"]], 5, int(pda.AdvectionDiffusion2d.BurgersPeriodic))
        else:
            params = {"gamma": float(rng.uniform(1.3, 1.5))}
            mesh = pda.create_structured_mesh((n, n + 3), (0, 1, 0, 1), 7)
            prob = pda.create_problem(mesh, pda.Euler2d.Riemann, S.Weno5, 2, params)
            orc = ([params["gamma"]], 2, int(pda.Euler2d.Riemann))
        inst.append((mesh, prob, orc))
    batch = pda.Batch([p for _, p, _ in inst])
    U = batch.initialCondition()
    U = U * torch.from_numpy(1 + 0.01 * rng.uniform(-1, 1, batch.total)).cuda()
    V = batch.createRightHandSide()
    batch.rightHandSide(U, 0.0, V)
    torch.cuda.synchronize()
    import pyoracle
    for i, (mesh, prob, (pp, fam, enum)) in enumerate(inst):
        pb = pyoracle.Oracle.problem(fam, enum, int(S.Weno5), pp)
        Vo = oracle.velocity(oracle.from_mesh(mesh), pb, batch.instance(U, i).cpu().numpy())
        e = cases.rel_err(batch.instance(V, i).cpu().numpy(), Vo)
        assert e <= 1e-12, (i, e)


def test_batch_set_strip_height(torch):
    """pda_batch_set_velocity_strip_height: heights 1, 5 and 64 give the same
    batched velocity bit for bit; 0 and 65 are rejected."""
    print("[gpu-batch] set strip height", flush=True)
    inst = _instances(pda.InviscidFluxReconstruction.Weno5, (16, 24, 33, 64), 3)
    batch = pda.Batch([p for _, p, _ in inst])
    rng = np.random.default_rng(4)
    U = batch.initialCondition() * torch.from_numpy(1 + 0.01 * rng.uniform(-1, 1, batch.total)).cuda()
    V0 = batch.createRightHandSide()
    batch.rightHandSide(U, 0.0, V0)
    for h in (1, 5, 64):
        batch.set_velocity_strip_height(h)
        V = batch.createRightHandSide()
        batch.rightHandSide(U, 0.0, V)
        torch.cuda.synchronize()
        assert torch.equal(V, V0), h
    for bad in (0, 65):
        with pytest.raises(pda.PdaError):
            batch.set_velocity_strip_height(bad)


def test_batch_at_config_scale(oracle, torch):
    """24 slip-wall instances drawn by the batched workload's own generator
    (grids from {64, 96, 128, 192, 256}^2, gravity, Coriolis and pulse
    amplitude from the workload's ranges), autotuned strips: each instance's
    batched velocity against the oracle (1e-12) and against the same
    problem's own velocity() (bit-exact), then 5 batched RK4 steps (one
    graph) against single-instance stepping, bit-exact."""
    from pressiodemoapps_amd import batched
    print("[gpu-batch] config scale, 24 instances", flush=True)
    batch, meshes, sizes = batched.build(n_inst=24, seed=12)
    assert len(set(sizes)) == 5
    rng = np.random.default_rng(13)
    U = batch.initialCondition() * torch.from_numpy(1 + 0.01 * rng.uniform(-1, 1, batch.total)).cuda()
    rps, us = batch.autotune(U)
    print(f"[gpu-batch] config scale: rows per strip {rps}, {us:.1f} us", flush=True)
    V = batch.createRightHandSide()
    batch.rightHandSide(U, 0.0, V)
    torch.cuda.synchronize()
    import pyoracle
    for i, n in enumerate(sizes):
        prob = batch._probs[i]
        Ui = batch.instance(U, i).contiguous()
        Vs = prob.createRightHandSide()
        prob.rightHandSide(Ui, 0.0, Vs)
        torch.cuda.synchronize()
        assert torch.equal(batch.instance(V, i), Vs), i
        pb = pyoracle.Oracle.problem(4, int(pda.Swe2d.SlipWall), int(S.Weno5), [prob.gravity(), prob.coriolis()])
        Vo = oracle.velocity(oracle.from_mesh(meshes[n]), pb, Ui.cpu().numpy())
        e = cases.rel_err(Vs.cpu().numpy(), Vo)
        assert e <= 1e-12, (i, n, e)
    Ub = U.clone()
    batch.advance(Ub, pda.StepScheme.RK4, 0.01, 5, steps_per_graph=5)
    torch.cuda.synchronize()
    for i in range(len(sizes)):
        Ui = batch.instance(U, i).clone()
        pda.ExplicitStepper(batch._probs[i], pda.StepScheme.RK4).advance(Ui, 0.01, 5)
        torch.cuda.synchronize()
        assert torch.equal(batch.instance(Ub, i), Ui), i
