"""The kernel shapes the measured workloads run, against the oracle.

The marching kernels process a strip of grid rows per workgroup; the strip
height follows the grid size (or a tuner), so small test grids and the
1024 x 1024 workload take different code paths (the first-order 2D Euler
Jacobian carries each row's y-face sides to the next row and prefetches the
next row's states only when a strip has two or more rows).  Here:
  - the first-order 2D Euler Jacobian at strip heights 1, 2, 3, 8 and the
    grid-size rule gives bit-identical values, within 1e-12 of the oracle;
  - the 3D velocity march at heights 1, 4, 12 and the rule likewise;
  - the first-order 3D Euler Jacobian at strip heights 1, 2, 3, 8 and the
    rule on three shapes, and at 256^3 (the 3D timing's shape) on ~3,000
    sampled rows against the oracle on a sample mesh of those rows;
  - the 1024 x 1024 Riemann workload itself (WENO5 velocity at the heights
    the step loop picks from, first-order Jacobian at the grid-size rule of
    2 rows): pattern bit-exact, values within 1e-12;
  - velocity and WENO5 / first-order Jacobians on states evolved into shocks
    by the device stepper, within 1e-12.
"""
import numpy as np
import pytest

import cases
import pressiodemoapps_amd as pda

pytestmark = pytest.mark.gpu
S = pda.InviscidFluxReconstruction
TOL = 1e-12


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    if not _t.cuda.is_available():
        pytest.skip("no HIP device")
    return _t


def _big(name):
    return [c for c in cases.BIG_CASES if c[0] == name][0]


def _jac(prob, Ud, torch):
    J = prob.createJacobian()
    prob.jacobian(Ud, 0.0, J)
    torch.cuda.synchronize()
    return J


def test_jacobian_strip_heights(oracle, torch):
    # 150 x 37: box rows 1..35 (35 rows: strips of 8 end with a 3-row strip),
    # five 31-cell segments per row, the last one partial
    case = _big("euler2d-riemann-big-FirstOrder")
    mesh, prob = cases.make(case)
    ms, pb = oracle.from_mesh(mesh), cases.oracle_problem(case)
    for seed in (0, 1):
        U = cases.perturbed_state(prob, seed)
        Ud = torch.from_numpy(U).cuda()
        _, rpo, cio, Jo = oracle.velocity_and_jacobian(ms, pb, U)
        ref = None
        for h in (1, 2, 3, 8, 0):
            print(f"[gpu-strips] jacobian 150x37 seed {seed} height {h}", flush=True)
            prob.set_jacobian_strip_height(h)
            J = _jac(prob, Ud, torch)
            if ref is None:
                ref = J.values.clone()
                assert np.array_equal(J.rowptr.cpu().numpy(), rpo)
                assert np.array_equal(J.colidx.cpu().numpy(), cio)
            else:
                assert torch.equal(J.values, ref), h
        err = cases.rel_err(ref.cpu().numpy(), Jo)
        print(f"[gpu-strips] jacobian 150x37 seed {seed}: {err:.2e}", flush=True)
        assert err <= TOL, err
    for bad in (-1, 65):
        with pytest.raises(pda.PdaError):
            prob.set_jacobian_strip_height(bad)


@pytest.mark.parametrize("shape", [(40, 6, 7), (37, 5, 21), (20, 3, 3)])
def test_jacobian_3d_strip_heights(shape, oracle, torch):
    """First-order Sedov 3D: the box's x lines in 21-cell segments (38 cells:
    21 + 17; 35: 21 + 14; 18: one partial segment), marched in z over strips of 1,
    2, 3, 8 planes and the grid-size rule (19 box planes: strips of 8 end
    with a 3-plane strip); pattern bit-exact, values bit-identical across
    heights and within 1e-12 of the oracle."""
    case = ("euler3d-sedov-strips-FirstOrder", pda.Euler3d.SedovSymmetry, S.FirstOrder, shape,
            (0, 1.2, 0, 1.2, 0, 1.2), 3, (), 0, None)
    mesh, prob = cases.make(case)
    ms, pb = oracle.from_mesh(mesh), cases.oracle_problem(case)
    U = cases.perturbed_state(prob, 5)
    Ud = torch.from_numpy(U).cuda()
    _, rpo, cio, Jo = oracle.velocity_and_jacobian(ms, pb, U)
    ref = None
    for h in (1, 2, 3, 8, 0):
        print(f"[gpu-strips] 3D jacobian {shape} height {h}", flush=True)
        prob.set_jacobian_strip_height(h)
        J = _jac(prob, Ud, torch)
        if ref is None:
            ref = J.values.clone()
            assert np.array_equal(J.rowptr.cpu().numpy(), rpo)
            assert np.array_equal(J.colidx.cpu().numpy(), cio)
        else:
            assert torch.equal(J.values, ref), h
    err = cases.rel_err(ref.cpu().numpy(), Jo)
    print(f"[gpu-strips] 3D jacobian {shape}: {err:.2e}", flush=True)
    assert err <= TOL, err


def test_sedov3d_256_jacobian_sampled_rows(oracle, torch):
    """The 256^3 first-order Sedov Jacobian the 3D timing measures: box x
    lines in 21-cell segments (the last of 2 cells) marched over 6-plane
    strips, boundary rows staged with their ghosts.  ~3,000 seeded rows, plus
    the last segment's cells, cells beside segment, strip and plane edges,
    and boundary cells, are checked against the oracle on a sample mesh of
    those rows: same columns (mapped back to full-mesh ids), values within
    1e-12.  The same rows of the WENO5 velocity (the 3D march at 12-row
    strips and at the rule) likewise; the rows include the blast corner, so
    that their largest value, which rel_err divides by, is the field's."""
    import pyoracle
    n = 256
    full = pda.create_structured_mesh((n, n, n), (0.0, 1.2, 0.0, 1.2, 0.0, 1.2), 7)
    ic = {"sedovRadiusCells": 2.0, "sedovAmbientPressure": 1e-5, "sedovEnergy": 0.851072}
    prob = pda.create_problem(full, pda.Euler3d.SedovSymmetry, S.FirstOrder, 0, ic)
    U = cases.perturbed_state(prob, 11)
    Ud = torch.from_numpy(U).cuda()
    print("[gpu-strips] 256^3 first-order jacobian", flush=True)
    J = _jac(prob, Ud, torch)
    rng = np.random.default_rng(11)
    pick = set(rng.choice(n ** 3, 3000, replace=False).tolist())
    for i, j, k in [(253, 1, 1), (254, 1, 1), (254, 254, 254), (253, 100, 6), (253, 100, 7), (1, 1, 1),
                    (0, 0, 0), (255, 255, 255), (0, 128, 128), (128, 0, 128), (128, 128, 0), (255, 17, 200),
                    (21, 5, 6), (20, 5, 6), (22, 5, 6), (42, 200, 12), (43, 200, 13), (231, 7, 253)]:
        pick.add((k * n + j) * n + i)
    # the blast corner, where |V| and the Jacobian entries peak, so that the
    # sampled rows' scale is the field's (rel_err divides by the largest
    # oracle value; rows of cancelling fluxes alone would understate it)
    for k in range(4):
        for j in range(4):
            for i in range(4):
                pick.add((k * n + j) * n + i)
    gids = np.array(sorted(pick), dtype=np.int32)
    g = torch.from_numpy(gids.astype(np.int64)).cuda()
    starts, ends = J.rowptr[5 * g], J.rowptr[5 * g + 5]
    lens = ends - starts
    offs = torch.cumsum(lens, 0) - lens
    idx = torch.repeat_interleave(starts - offs, lens) + torch.arange(int(lens.sum()), device="cuda")
    vals = J.values[idx].cpu().numpy()
    cols = J.colidx[idx].cpu().numpy().astype(np.int64)
    del J, idx
    torch.cuda.empty_cache()
    sm = pda.create_sample_mesh(full, gids)
    _, stencil = sm.sample_gids()
    Us = U.reshape(-1, 5)[stencil].ravel()
    pb = pyoracle.Oracle.problem(3, int(pda.Euler3d.SedovSymmetry), int(S.FirstOrder), [1.4])
    _, rpo, cio, Jo = oracle.velocity_and_jacobian(oracle.from_mesh(sm), pb, Us)
    assert np.array_equal(stencil[cio // 5].astype(np.int64) * 5 + cio % 5, cols), "columns differ"
    err = cases.rel_err(vals, Jo)
    print(f"[gpu-strips] 256^3 jacobian, {len(gids)} rows: {err:.2e}", flush=True)
    assert err <= TOL, err
    # the WENO5 velocity of the same state at the march's timed strip height
    # (12 rows, the autotuner's pick at 256^3) and the rule, on the same rows
    # (the sample mesh of the 7-point full mesh holds the WENO5 stencils)
    pw = pda.create_problem(full, pda.Euler3d.SedovSymmetry, S.Weno5, 0, ic)
    pbw = pyoracle.Oracle.problem(3, int(pda.Euler3d.SedovSymmetry), int(S.Weno5), [1.4])
    Vo = oracle.velocity(oracle.from_mesh(sm), pbw, Us)
    for h in (12, 0):
        pw.set_velocity_strip_height(h)
        V = pw.createRightHandSide()
        pw.rightHandSide(Ud, 0.0, V)
        torch.cuda.synchronize()
        Vg = V.view(-1, 5)[g].cpu().numpy().ravel()
        assert np.abs(Vg).max() == float(V.abs().max()), "the sampled rows miss the field's peak"
        err = cases.rel_err(Vg, Vo)
        print(f"[gpu-strips] 256^3 WENO5 velocity, height {h}, {len(gids)} rows: {err:.2e}", flush=True)
        assert err <= TOL, err


def test_velocity_march_3d_strip_heights(oracle, torch):
    # Sedov 3D, WENO5: 40 grid rows per plane, so 12-row strips end with a
    # 4-row strip; the default rule gives 2 rows here and 12 at 256^3
    case = ("euler3d-sedov-tall-Weno5", pda.Euler3d.SedovSymmetry, S.Weno5, (24, 40, 9),
            (0, 1.2, 0, 1.2, 0, 1.2), 7, (), 0, None)
    mesh, prob = cases.make(case)
    ms, pb = oracle.from_mesh(mesh), cases.oracle_problem(case)
    U = cases.perturbed_state(prob, 4)
    Ud = torch.from_numpy(U).cuda()
    Vo = oracle.velocity(ms, pb, U)
    ref = None
    for h in (1, 4, 12, 0):
        print(f"[gpu-strips] 3D velocity height {h}", flush=True)
        prob.set_velocity_strip_height(h)
        V = prob.createRightHandSide()
        prob.rightHandSide(Ud, 0.0, V)
        torch.cuda.synchronize()
        if ref is None:
            ref = V.clone()
        else:
            assert torch.equal(V, ref), h
    err = cases.rel_err(ref.cpu().numpy(), Vo)
    print(f"[gpu-strips] 3D velocity: {err:.2e}", flush=True)
    assert err <= TOL, err


def test_riemann_1024_workload(oracle, torch):
    """The measured 1024^2 workload: WENO5 velocity (problem A) at the strip
    heights the step loop chooses from, first-order jacobian() (problem B)
    at its grid-size rule (2 rows per strip) and at 3 rows."""
    n = 1024
    mesh = pda.create_structured_mesh((n, n), (0.0, 1.0, 0.0, 1.0), 7)
    pa = pda.create_problem(mesh, pda.Euler2d.Riemann, S.Weno5, 2)
    pb = pda.create_problem(mesh, pda.Euler2d.Riemann, S.FirstOrder, 2)
    U = cases.perturbed_state(pb, 7)
    Ud = torch.from_numpy(U).cuda()
    ms = oracle.from_mesh(mesh)
    print("[gpu-strips] 1024^2 velocity", flush=True)
    Vo = oracle.velocity(ms, oracle.problem(2, 4, 2, [1.4]), U)
    ref = None
    for h in (2, 3, 4, 6, 0):
        pa.set_velocity_strip_height(h)
        V = pa.createRightHandSide()
        pa.rightHandSide(Ud, 0.0, V)
        torch.cuda.synchronize()
        if ref is None:
            ref = V.clone()
        else:
            assert torch.equal(V, ref), h
    err = cases.rel_err(ref.cpu().numpy(), Vo)
    print(f"[gpu-strips] 1024^2 velocity: {err:.2e}", flush=True)
    assert err <= TOL, err
    del Vo, ref
    print("[gpu-strips] 1024^2 first-order jacobian", flush=True)
    _, rpo, cio, Jo = oracle.velocity_and_jacobian(ms, oracle.problem(2, 4, 0, [1.4]), U)
    J = _jac(pb, Ud, torch)
    assert np.array_equal(J.rowptr.cpu().numpy(), rpo), "row pointers differ"
    assert np.array_equal(J.colidx.cpu().numpy(), cio), "column indices differ"
    del rpo, cio
    err = cases.rel_err(J.values.cpu().numpy(), Jo)
    print(f"[gpu-strips] 1024^2 jacobian (rule): {err:.2e}", flush=True)
    assert err <= TOL, err
    pb.set_jacobian_strip_height(3)
    J3 = _jac(pb, Ud, torch)
    assert torch.equal(J3.values, J.values)


@pytest.mark.parametrize("scheme", [S.Weno5, S.FirstOrder], ids=lambda s: s.name)
def test_jacobian_on_evolved_shock_state(scheme, oracle, torch):
    """Riemann IC2 on 64 x 64 advanced by the device SSPRK3 stepper to
    t = 0.3 (shocks and contacts inside the grid), then velocity, velocity
    from the Jacobian path and Jacobian values against the oracle on that
    state."""
    mesh = pda.create_structured_mesh((64, 64), (0.0, 1.0, 0.0, 1.0), 7)
    prob = pda.create_problem(mesh, pda.Euler2d.Riemann, scheme, 2)
    U = prob.initialCondition()
    pda.advanceSSP3(prob, U, 2e-3, 150)
    torch.cuda.synchronize()
    Uh = U.cpu().numpy()
    assert np.isfinite(Uh).all()
    rho = Uh[0::4]
    assert rho.max() / rho.min() > 5.0  # strong gradients are present
    V = prob.createRightHandSide()
    prob.rightHandSide(U, 0.0, V)
    V2 = prob.createRightHandSide()
    J = prob.createJacobian()
    prob.rightHandSideAndJacobian(U, 0.0, V2, J)
    torch.cuda.synchronize()
    ms = oracle.from_mesh(mesh)
    pb = oracle.problem(2, 4, int(scheme), [1.4])
    Vo = oracle.velocity(ms, pb, Uh)
    V2o, rpo, cio, Jo = oracle.velocity_and_jacobian(ms, pb, Uh)
    e = (cases.rel_err(V.cpu().numpy(), Vo), cases.rel_err(V2.cpu().numpy(), V2o),
         cases.rel_err(J.values.cpu().numpy(), Jo))
    print(f"[gpu-strips] evolved Riemann {scheme.name}: V {e[0]:.2e} V(J) {e[1]:.2e} J {e[2]:.2e}", flush=True)
    assert np.array_equal(J.colidx.cpu().numpy(), cio)
    assert max(e) <= TOL, e
