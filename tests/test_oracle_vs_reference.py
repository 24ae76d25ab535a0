"""The CPU oracle against the reference itself (oracle/_ref, compiled from the
reference headers), and the product's host-side set-up (pattern, initial
condition, parameters) against the same.  Skipped where oracle/_ref was not
built (it needs the reference sources at build time)."""
import ctypes
import os

import numpy as np
import pytest

import cases
import pyoracle
import pressiodemoapps_amd as pda

S = pda.InviscidFluxReconstruction
pytestmark = pytest.mark.skipif(not pyoracle.ref_available(), reason="oracle/_ref not built")

FAM_PARAMS = {
    4: {"gravity": 9.8, "coriolis": -3.0},
}


def _ref_for(case, tmp_path, mesh=None):
    _, enum, scheme, n, bounds, st, per, ic, params = case
    mesh = mesh or cases.make(case)[0]
    d = str(tmp_path / "mesh")
    mesh.write(d)
    fam = pda.FAMILY[type(enum)]
    return pyoracle.RefProblem(d, fam, int(enum), int(scheme), ic, params)


# the reference builds no 3D mesh with a 7-point stencil (impl/mesh_ccu.hpp:367-369),
# so 3D WENO5 has no reference run to compare with
REF_CASES = [c for c in cases.CASES if not (c[0].startswith("euler3d") and c[0].endswith("Weno5"))]


@pytest.mark.parametrize("case", REF_CASES, ids=[c[0] for c in REF_CASES])
def test_oracle_bit_exact_with_reference(case, oracle, tmp_path):
    mesh, prob = cases.make(case)
    ref = _ref_for(case, tmp_path, mesh)
    ms, pb = oracle.from_mesh(mesh), cases.oracle_problem(case)
    for seed in (0, 1):
        U = cases.perturbed_state(prob, seed)
        assert np.array_equal(oracle.velocity(ms, pb, U), ref.velocity(U))
        Vo, rpo, cio, Jo = oracle.velocity_and_jacobian(ms, pb, U)
        Vr, rpr, cir, Jr = ref.velocity_and_jacobian(U)
        assert np.array_equal(rpo, rpr) and np.array_equal(cio, cir)
        assert np.array_equal(Vo, Vr)
        assert np.array_equal(Jo, Jr)


@pytest.mark.parametrize("case", REF_CASES, ids=[c[0] for c in REF_CASES])
def test_product_setup_matches_reference(case, tmp_path):
    """Jacobian sparsity (bit-exact), sizes and initial condition."""
    mesh, prob = cases.make(case)
    ref = _ref_for(case, tmp_path, mesh)
    assert prob.totalDofSampleMesh() == ref.rows
    assert prob.totalDofStencilMesh() == ref.cols
    assert prob.nonZeros() == ref.nnz
    rp, ci = prob.jacobian_pattern_host()
    U = cases.perturbed_state(prob, 0)
    _, rpr, cir, _ = ref.velocity_and_jacobian(U)
    assert np.array_equal(rp, rpr) and np.array_equal(ci, cir)
    U0 = np.zeros(prob.totalDofStencilMesh())
    assert pda.lib().pda_initial_condition_host(prob.handle, U0.ctypes.data_as(ctypes.c_void_p)) == 0
    Ur = ref.initial_condition()
    assert np.abs(U0 - Ur).max() <= 1e-14 * max(1.0, np.abs(Ur).max())


SAMPLE_CASES = [c for c in REF_CASES if c[0] in (
    "euler2d-riemann-FirstOrder", "euler2d-riemann-Weno5", "swe-slipwall-Weno3",
    "burgers-periodic-Weno5", "euler1d-sod-Weno5", "euler3d-sedov-Weno3", "adr-probA-FirstOrder",
    "adr-custom-Weno5", "burgers-outflow-Weno5", "burgers-outflow-FirstOrder")]


@pytest.mark.parametrize("case", SAMPLE_CASES, ids=lambda c: c[0])
def test_sample_mesh_oracle_matches_reference(case, oracle, tmp_path):
    """Sample/stencil mesh: rows are a seeded random subset, columns the
    stencil-mesh cells (meshing_scripts/create_sample_mesh.py)."""
    full, fprob = cases.make(case)
    rng = np.random.default_rng(1)
    nsamp = max(4, full.sampleMeshSize() // 5)
    gids = np.sort(rng.choice(full.sampleMeshSize(), size=nsamp, replace=False))
    sm = pda.create_sample_mesh(full, gids)
    ref = _ref_for(case, tmp_path, sm)
    _, st = sm.sample_gids()
    Ufull = cases.perturbed_state(fprob, 4)
    m = fprob.numDofPerCell()
    U = Ufull.reshape(-1, m)[st].ravel()
    ms = oracle.from_mesh(sm)
    pb = cases.oracle_problem(case)
    assert np.array_equal(oracle.velocity(ms, pb, U), ref.velocity(U))
    Vo, rpo, cio, Jo = oracle.velocity_and_jacobian(ms, pb, U)
    Vr, rpr, cir, Jr = ref.velocity_and_jacobian(U)
    assert np.array_equal(rpo, rpr) and np.array_equal(cio, cir)
    assert np.array_equal(Vo, Vr) and np.array_equal(Jo, Jr)
    # and the sample-mesh residual equals the full-mesh residual on the sample rows
    Vfull = oracle.velocity(oracle.from_mesh(full), pb, Ufull)
    assert np.array_equal(Vo[::1].reshape(-1, m), Vfull.reshape(-1, m)[gids]) or \
        np.abs(Vo.reshape(-1, m) - Vfull.reshape(-1, m)[gids]).max() <= 1e-14 * np.abs(Vfull).max()


def test_custom_parameter_creators_match_reference(tmp_path):
    """create_slip_wall_swe_2d_problem / create_adv_diff_reac_2d_problem_A /
    create_burgers_2d_problem give the reference's set-up."""
    mesh = pda.create_structured_mesh((20, 22), (-5, 5, -5, 5), 7)
    d = str(tmp_path / "m1")
    mesh.write(d)
    p = pda.create_slip_wall_swe_2d_problem(mesh, S.Weno3, 9.0, -2.5, 0.3)
    ref = pyoracle.RefProblem(d, 4, 0, int(S.Weno3), 1, {"gravity": 9.0, "coriolis": -2.5, "pulseMagnitude": 0.3})
    U0 = np.zeros(p.totalDofStencilMesh())
    assert pda.lib().pda_initial_condition_host(p.handle, U0.ctypes.data_as(ctypes.c_void_p)) == 0
    assert np.abs(U0 - ref.initial_condition()).max() <= 1e-14
    assert p.nonZeros() == ref.nnz
    q = pda.create_adv_diff_reac_2d_problem_A(mesh, S.Weno5, 0.3, -0.2, 0.01, 0.5)
    refq = pyoracle.RefProblem(d, 6, 0, int(S.Weno5), 0,
                               {"ux": 0.3, "uy": -0.2, "diffusion": 0.01, "sigma": 0.5})
    assert q.nonZeros() == refq.nnz and q.totalDofSampleMesh() == refq.rows
    mb = pda.create_structured_mesh((16, 18), (-1, 1, -1, 1), 7, True)
    db = str(tmp_path / "m2")
    mb.write(db)
    params = {"diffusion": 2e-3, "pulseMagnitude": 0.7, "pulseX": 0.1}
    b = pda.create_burgers_2d_problem(mb, pda.AdvectionDiffusion2d.BurgersPeriodic, S.FirstOrder,
                                      pda.ViscousFluxReconstruction.FirstOrder, params)
    refb = pyoracle.RefProblem(db, 5, 0, int(S.FirstOrder), 1, params)
    Ub = np.zeros(b.totalDofStencilMesh())
    assert pda.lib().pda_initial_condition_host(b.handle, Ub.ctypes.data_as(ctypes.c_void_p)) == 0
    assert np.abs(Ub - refb.initial_condition()).max() <= 1e-14
