"""Velocity strip-height tuning (pda_problem_autotune): the kept height is
one of the candidates, velocity() after tuning is bit-identical to before
(every face has one evaluation site whatever the height) and still matches
the oracle within 1e-12; a problem whose velocity does not use the
structured kernel (a sample mesh) reports height 0."""
import numpy as np
import pytest

import cases
import pressiodemoapps_amd as pda

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    if not _t.cuda.is_available():
        pytest.skip("no HIP device")
    return _t


@pytest.mark.parametrize("name", ["euler2d-riemann-big-Weno5", "burgers-big-Weno5"])
def test_autotune_keeps_values(name, oracle, torch):
    case = [c for c in cases.BIG_CASES if c[0] == name][0]
    print(f"[gpu-autotune] {name}", flush=True)
    mesh, prob = cases.make(case)
    U = cases.perturbed_state(prob, 3)
    Ud = torch.from_numpy(U).cuda()
    V0 = prob.createRightHandSide()
    prob.rightHandSide(Ud, 0.0, V0)
    rps, us = prob.autotune(Ud)
    V1 = prob.createRightHandSide()
    prob.rightHandSide(Ud, 0.0, V1)
    torch.cuda.synchronize()
    print(f"[gpu-autotune] {name}: rows per strip {rps}, {us:.1f} us", flush=True)
    assert rps in (1, 2, 3, 4, 5, 6, 8, 12, 16, 24, 32) and us > 0
    assert torch.equal(V0, V1)
    Vo = oracle.velocity(oracle.from_mesh(mesh), cases.oracle_problem(case), U)
    assert cases.rel_err(V1.cpu().numpy(), Vo) <= 1e-12


def test_autotune_without_structured_kernel(torch):
    full = pda.create_structured_mesh((40, 30), (0.0, 1.0, 0.0, 1.0), 7)
    gids = np.arange(0, full.sampleMeshSize(), 7)
    sm = pda.create_sample_mesh(full, gids)
    prob = pda.create_problem(sm, pda.Euler2d.Riemann, pda.InviscidFluxReconstruction.Weno5, 2)
    U = prob.initialCondition()
    rps, us = prob.autotune(U)
    torch.cuda.synchronize()
    assert rps == 0 and us > 0
    # a fixed height is accepted and ignored: the velocity does not change
    V0 = prob.createRightHandSide()
    prob.rightHandSide(U, 0.0, V0)
    prob.set_velocity_strip_height(5)
    V1 = prob.createRightHandSide()
    prob.rightHandSide(U, 0.0, V1)
    torch.cuda.synchronize()
    assert torch.equal(V0, V1)


def test_set_strip_height(torch):
    """pda_problem_set_velocity_strip_height: heights 1, 7, 64 and 0 (the
    grid-size rule) give bit-identical velocities; heights outside 0..64 are
    rejected with an error and leave the problem usable."""
    case = [c for c in cases.BIG_CASES if c[0] == "euler2d-riemann-big-Weno5"][0]
    print("[gpu-autotune] set strip height", flush=True)
    mesh, prob = cases.make(case)
    Ud = torch.from_numpy(cases.perturbed_state(prob, 5)).cuda()
    V0 = prob.createRightHandSide()
    prob.rightHandSide(Ud, 0.0, V0)
    for h in (1, 7, 64, 0):
        prob.set_velocity_strip_height(h)
        V = prob.createRightHandSide()
        prob.rightHandSide(Ud, 0.0, V)
        torch.cuda.synchronize()
        assert torch.equal(V, V0), h
    for bad in (-1, 65):
        with pytest.raises(pda.PdaError):
            prob.set_velocity_strip_height(bad)
    V = prob.createRightHandSide()
    prob.rightHandSide(Ud, 0.0, V)
    torch.cuda.synchronize()
    assert torch.equal(V, V0)
