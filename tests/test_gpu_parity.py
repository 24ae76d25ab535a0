"""HIP engine vs the CPU oracle on the same seeded inputs, through the C ABI.

Bar: Jacobian sparsity bit-exact; velocity and Jacobian values within 1e-12
of the oracle, norm-wise (max |diff| / max |oracle|).
"""
import ctypes
import sys

import numpy as np
import pytest

import cases
import pressiodemoapps_amd as pda

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    if not _t.cuda.is_available():
        pytest.skip("no HIP device")
    return _t


def _run_gpu(prob, U, torch):
    Ud = torch.from_numpy(U).cuda()
    V = prob.createRightHandSide()
    prob.rightHandSide(Ud, 0.0, V)
    J = prob.createJacobian()
    V2 = prob.createRightHandSide()
    prob.rightHandSideAndJacobian(Ud, 0.0, V2, J)
    Jv = prob.createJacobian()
    prob.jacobian(Ud, 0.0, Jv)
    torch.cuda.synchronize()
    return (V.cpu().numpy(), V2.cpu().numpy(), J.values.cpu().numpy(), Jv.values.cpu().numpy(),
            J.rowptr.cpu().numpy(), J.colidx.cpu().numpy())


@pytest.mark.parametrize("case", cases.CASES + cases.BIG_CASES, ids=cases.IDS + cases.BIG_IDS)
def test_velocity_and_jacobian_match_oracle(case, oracle, torch):
    print(f"[gpu-parity] {case[0]}", flush=True)
    mesh, prob = cases.make(case)
    ms = oracle.from_mesh(mesh)
    pb = cases.oracle_problem(case)
    for seed in (0, 1):
        U = cases.perturbed_state(prob, seed)
        Vg, V2g, Jg, Jvg, rp, ci = _run_gpu(prob, U, torch)
        Vo = oracle.velocity(ms, pb, U)
        V2o, rpo, cio, Jo = oracle.velocity_and_jacobian(ms, pb, U)
        assert np.array_equal(rp, rpo), "row pointers differ"
        assert np.array_equal(ci, cio), "column indices differ"
        print(f"[gpu-parity] {case[0]} seed {seed}: V {cases.rel_err(Vg, Vo):.2e} "
              f"V(J path) {cases.rel_err(V2g, V2o):.2e} J {cases.rel_err(Jg, Jo):.2e}", flush=True)
        assert cases.rel_err(Vg, Vo) <= TOL, cases.rel_err(Vg, Vo)
        assert cases.rel_err(V2g, V2o) <= TOL, cases.rel_err(V2g, V2o)
        assert cases.rel_err(Jg, Jo) <= TOL, cases.rel_err(Jg, Jo)
        # jacobian() without V gives the same values up to rounding: the two
        # calls run separately compiled kernels, which may contract
        # multiply-adds differently
        assert cases.rel_err(Jvg, Jg) <= 1e-14, cases.rel_err(Jvg, Jg)


# grid widths around the 2D velocity march's segment width (64 - 2W owned
# cells per wavefront: 58, 60, 62 for W = 3, 2, 1), the narrowest grids a
# full-grid box allows (2W + 1), mirror ghosts with sign flips (slip walls),
# zero ghosts (Burgers outflow's minus sides) and periodic wraps
S = pda.InviscidFluxReconstruction
SEGMENT_EDGE_CASES = (
    [(f"riemann-{nx}x9-Weno5", pda.Euler2d.Riemann, S.Weno5, (nx, 9), (0, 1, 0, 1), 7, (), 2, None)
     for nx in (7, 57, 58, 59, 116, 117)]
    + [(f"periodic-{nx}x8-Weno5", pda.Euler2d.PeriodicSmooth, S.Weno5, (nx, 8), (-1, 1, -1, 1), 7, True, 0, None)
       for nx in (7, 58, 59)]
    + [(f"swe-{nx}x{ny}-{s.name}", pda.Swe2d.SlipWall, s, (nx, ny), (-5, 5, -5, 5), 7, (), 1,
        {"gravity": 9.8, "coriolis": -3.0})
       for s, nx, ny in ((S.Weno5, 59, 9), (S.Weno3, 60, 7), (S.Weno3, 61, 7), (S.FirstOrder, 62, 5),
                         (S.FirstOrder, 63, 5))]
    + [("burgers-outflow-59x9-Weno5", pda.AdvectionDiffusion2d.BurgersOutflow, S.Weno5, (59, 9), (-1, 1, -1, 1), 7,
        (), 1, {"diffusion": 5e-3, "pulseX": 0.6, "pulseY": 0.5}),
       ("burgers-periodic-61x7-Weno3", pda.AdvectionDiffusion2d.BurgersPeriodic, S.Weno3, (61, 7), (-1, 1, -1, 1), 5,
        True, 1, {"diffusion": 5e-3})])


@pytest.mark.parametrize("case", SEGMENT_EDGE_CASES, ids=[c[0] for c in SEGMENT_EDGE_CASES])
def test_velocity_segment_edges(case, oracle, torch):
    """velocity() at the widths where the march's last segment is full,
    one cell short or one cell over, at strip heights 1, 4 and 64 (equal
    bit for bit) against the oracle."""
    print(f"[gpu-parity] velocity {case[0]}", flush=True)
    mesh, prob = cases.make(case)
    ms = oracle.from_mesh(mesh)
    pb = cases.oracle_problem(case)
    U = cases.perturbed_state(prob, 3)
    Ud = torch.from_numpy(U).cuda()
    out = []
    for h in (0, 1, 4, 64):
        prob.set_velocity_strip_height(h)
        V = prob.createRightHandSide()
        prob.rightHandSide(Ud, 0.0, V)
        torch.cuda.synchronize()
        out.append(V.cpu().numpy())
    prob.set_velocity_strip_height(0)
    Vo = oracle.velocity(ms, pb, U)
    e = cases.rel_err(out[0], Vo)
    print(f"[gpu-parity] velocity {case[0]}: {e:.2e}", flush=True)
    assert e <= TOL, e
    for V in out[1:]:
        assert np.array_equal(V, out[0])


@pytest.mark.parametrize("case", [c for c in cases.CASES if "riemann-Weno5" in c[0]
                                  or "burgers-periodic-Weno5" in c[0]
                                  or "euler1d-sod-Weno3" in c[0]],
                         ids=lambda c: c[0])
def test_apply_jacobian_matches_csr_product(case, oracle, torch):
    mesh, prob = cases.make(case)
    U = cases.perturbed_state(prob, 3)
    Ud = torch.from_numpy(U).cuda()
    ms = oracle.from_mesh(mesh)
    _, rp, ci, Jo = oracle.velocity_and_jacobian(ms, cases.oracle_problem(case), U)
    import scipy.sparse as sp
    Jsp = sp.csr_matrix((Jo, ci, rp), shape=(prob.totalDofSampleMesh(), prob.totalDofStencilMesh()))
    rng = np.random.default_rng(5)
    B = rng.standard_normal((prob.totalDofStencilMesh(), 3))
    ref = Jsp @ B
    for Bt in (torch.from_numpy(B).cuda(), torch.from_numpy(np.asfortranarray(B)).cuda().t().contiguous().t()):
        AJ = prob.createApplyJacobianResult(Bt)
        prob.applyJacobian(Ud, Bt, 0.0, AJ)
        torch.cuda.synchronize()
        assert cases.rel_err(AJ.cpu().numpy(), ref) <= TOL
    b = torch.from_numpy(B[:, 0].copy()).cuda()
    AJ = prob.createApplyJacobianResult(b)
    prob.applyJacobian(Ud, b, 0.0, AJ)
    torch.cuda.synchronize()
    assert cases.rel_err(AJ.cpu().numpy(), ref[:, 0]) <= TOL


SAMPLE_CASES = [c for c in cases.CASES if c[0].endswith(("FirstOrder", "Weno3", "Weno5"))]


@pytest.mark.parametrize("case", SAMPLE_CASES, ids=[c[0] for c in SAMPLE_CASES])
def test_sample_mesh_matches_oracle(case, oracle, torch):
    """Sample/stencil mesh (meshing_scripts/create_sample_mesh.py): a seeded
    random 20 % of the cells as rows, their stencil cells as columns."""
    print(f"[gpu-parity-sample] {case[0]}", flush=True)
    full, _ = cases.make(case)
    rng = np.random.default_rng(11)
    nsamp = max(4, full.sampleMeshSize() // 5)
    gids = np.sort(rng.choice(full.sampleMeshSize(), size=nsamp, replace=False))
    sm = pda.create_sample_mesh(full, gids)
    _, enum, scheme, n, bounds, st, per, ic, params = case
    prob = pda.create_problem(sm, enum, scheme, ic, params)
    U = cases.perturbed_state(prob, 5)
    Vg, V2g, Jg, Jvg, rp, ci = _run_gpu(prob, U, torch)
    ms = oracle.from_mesh(sm)
    pb = cases.oracle_problem(case)
    Vo = oracle.velocity(ms, pb, U)
    V2o, rpo, cio, Jo = oracle.velocity_and_jacobian(ms, pb, U)
    assert np.array_equal(rp, rpo) and np.array_equal(ci, cio)
    print(f"[gpu-parity-sample] {case[0]}: rows {nsamp} V {cases.rel_err(Vg, Vo):.2e} "
          f"V(J path) {cases.rel_err(V2g, V2o):.2e} J {cases.rel_err(Jg, Jo):.2e}", flush=True)
    assert cases.rel_err(Vg, Vo) <= TOL
    assert cases.rel_err(V2g, V2o) <= TOL
    assert cases.rel_err(Jg, Jo) <= TOL
    assert cases.rel_err(Jvg, Jo) <= TOL
    assert cases.rel_err(Jvg, Jg) <= 1e-14


@pytest.mark.parametrize("dim", [2, 3])
def test_sample_mesh_contiguous_runs(dim, oracle, torch):
    """A sample mesh of whole runs of cells along x: the staged first-order
    Jacobian sees 16 consecutive inner rows in one wavefront and copies them
    as one range; the values still match the oracle."""
    print(f"[gpu-parity-sample] contiguous runs, {dim}D first order", flush=True)
    if dim == 2:
        full = pda.create_structured_mesh((60, 20), (0.0, 1.0, 0.0, 1.0), 7)
        nx, ny = 60, 20
        gids = np.array([j * nx + i for j in range(5, 15) for i in range(3, 57)])
        enum, ic, params, fam = pda.Euler2d.Riemann, 2, None, 2
    else:
        full = pda.create_structured_mesh((40, 8, 7), (0.0, 1.2, 0.0, 1.2, 0.0, 1.2), 7)
        nx, ny = 40, 8
        gids = np.array([(k * ny + j) * nx + i for k in range(2, 5) for j in range(2, 6) for i in range(2, 38)])
        enum, ic, params, fam = pda.Euler3d.SedovSymmetry, 0, None, 3
    sm = pda.create_sample_mesh(full, gids)
    prob = pda.create_problem(sm, enum, pda.InviscidFluxReconstruction.FirstOrder, ic, params)
    U = cases.perturbed_state(prob, 6)
    Vg, V2g, Jg, Jvg, rp, ci = _run_gpu(prob, U, torch)
    import pyoracle
    ms = oracle.from_mesh(sm)
    pb = pyoracle.Oracle.problem(fam, int(enum), int(pda.InviscidFluxReconstruction.FirstOrder), [1.4])
    V2o, rpo, cio, Jo = oracle.velocity_and_jacobian(ms, pb, U)
    assert np.array_equal(rp, rpo) and np.array_equal(ci, cio)
    print(f"[gpu-parity-sample] contiguous {dim}D: J {cases.rel_err(Jvg, Jo):.2e}", flush=True)
    assert cases.rel_err(Jg, Jo) <= TOL
    assert cases.rel_err(Jvg, Jo) <= TOL


def test_sample_mesh_5pct_sedov3d(oracle, torch):
    """3D Sedov with symmetry BCs on a 5 % seeded random sample mesh of a
    24^3 grid, first-order and WENO3."""
    for scheme in (pda.InviscidFluxReconstruction.FirstOrder, pda.InviscidFluxReconstruction.Weno3):
        full = pda.create_structured_mesh((24, 24, 24), (0, 1.2, 0, 1.2, 0, 1.2), 5)
        rng = np.random.default_rng(7)
        gids = np.sort(rng.choice(full.sampleMeshSize(), size=full.sampleMeshSize() // 20,
                                  replace=False))
        sm = pda.create_sample_mesh(full, gids)
        prob = pda.create_problem(sm, pda.Euler3d.SedovSymmetry, scheme, 0, None)
        U = cases.perturbed_state(prob, 9)
        Vg, V2g, Jg, Jvg, rp, ci = _run_gpu(prob, U, torch)
        import pyoracle
        pb = pyoracle.Oracle.problem(3, int(pda.Euler3d.SedovSymmetry), int(scheme), [1.4])
        ms = oracle.from_mesh(sm)
        Vo = oracle.velocity(ms, pb, U)
        V2o, rpo, cio, Jo = oracle.velocity_and_jacobian(ms, pb, U)
        assert np.array_equal(rp, rpo) and np.array_equal(ci, cio)
        print(f"[gpu-parity-sample] sedov3d 5% {scheme.name}: V {cases.rel_err(Vg, Vo):.2e} "
              f"J {cases.rel_err(Jg, Jo):.2e}", flush=True)
        assert cases.rel_err(Vg, Vo) <= TOL and cases.rel_err(V2g, V2o) <= TOL
        assert cases.rel_err(Jg, Jo) <= TOL
        assert cases.rel_err(Jvg, Jo) <= TOL  # jacobian() alone (staged first-order kernel)


def test_rhs_and_jacobian_aliases(torch):
    """The adapter's other names (adapter_cpp.hpp:116, :210-221): createRhs,
    rhsAndJacobian with and without a Jacobian."""
    case = [c for c in cases.BIG_CASES if c[0] == "euler2d-riemann-big-Weno5"][0]
    _, prob = cases.make(case)
    U = torch.from_numpy(cases.perturbed_state(prob, 4)).cuda()
    V1, V2, V3 = prob.createRhs(), prob.createRightHandSide(), prob.createRhs()
    J1, J2 = prob.createJacobian(), prob.createJacobian()
    prob.rhsAndJacobian(U, 0.0, V1)
    prob.rightHandSide(U, 0.0, V2)
    prob.rhsAndJacobian(U, 0.0, V3, J1)
    Vj = prob.createRhs()
    prob.rightHandSideAndJacobian(U, 0.0, Vj, J2)
    torch.cuda.synchronize()
    assert torch.equal(V1, V2)
    assert torch.equal(V3, Vj) and torch.equal(J1.values, J2.values)


def _random_cases():
    """Seeded random 2D shapes (7..133 per side) over the box kernels'
    families, schemes and boundary kinds: partial march segments, partial
    WENO tiles and strips at arbitrary positions."""
    rng = np.random.default_rng(2024)
    kinds = [
        ("riemann", pda.Euler2d.Riemann, (0, 1, 0, 1), (), 2, None),
        ("euler-periodic", pda.Euler2d.PeriodicSmooth, (-1, 1, -1, 1), True, 0, None),
        ("swe", pda.Swe2d.SlipWall, (-5, 5, -5, 5), (), 1, {"gravity": 9.8, "coriolis": -3.0}),
        ("burgers-periodic", pda.AdvectionDiffusion2d.BurgersPeriodic, (-1, 1, -1, 1), True, 1,
         {"diffusion": 5e-3}),
        ("burgers-outflow", pda.AdvectionDiffusion2d.BurgersOutflow, (-1, 1, -1, 1), (), 1,
         {"diffusion": 5e-3, "pulseX": 0.6, "pulseY": 0.5}),
    ]
    out = []
    for kname, enum, bounds, per, ic, params in kinds:
        for s in (S.FirstOrder, S.Weno3, S.Weno5):
            nx, ny = (int(v) for v in rng.integers(7, 134, size=2))
            st = {S.FirstOrder: 3, S.Weno3: 5, S.Weno5: 7}[s]
            out.append((f"rand-{kname}-{s.name}-{nx}x{ny}", enum, s, (nx, ny), bounds, st, per, ic, params))
    return out


RANDOM_CASES = _random_cases()


@pytest.mark.parametrize("case", RANDOM_CASES, ids=[c[0] for c in RANDOM_CASES])
def test_random_shapes_match_oracle(case, oracle, torch):
    print(f"[gpu-parity] {case[0]}", flush=True)
    mesh, prob = cases.make(case)
    ms = oracle.from_mesh(mesh)
    pb = cases.oracle_problem(case)
    U = cases.perturbed_state(prob, 11)
    Vg, V2g, Jg, Jvg, rp, ci = _run_gpu(prob, U, torch)
    Vo = oracle.velocity(ms, pb, U)
    V2o, rpo, cio, Jo = oracle.velocity_and_jacobian(ms, pb, U)
    assert np.array_equal(rp, rpo) and np.array_equal(ci, cio)
    e = (cases.rel_err(Vg, Vo), cases.rel_err(V2g, V2o), cases.rel_err(Jg, Jo), cases.rel_err(Jvg, Jo))
    print(f"[gpu-parity] {case[0]}: V {e[0]:.2e} V(J path) {e[1]:.2e} J {e[2]:.2e} J alone {e[3]:.2e}", flush=True)
    assert max(e) <= TOL, e
