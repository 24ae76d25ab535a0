"""The reference's own Python tests (tests_py) replayed through this
package: each loads the reference's mesh directory with
load_cellcentered_uniform_mesh, creates the problem with the same call,
integrates with advanceRK4 / advanceSSP3 (the module's integrators, on the
device) for the same dt and step count, and compares with the reference's
gold file at the reference test's own rtol / atol.  Meshes and golds are
data copied by tests/golden/collect_reference_fixtures.py (meshes
gzip-compressed, golds as .npy)."""
import gzip
import os
import shutil

import numpy as np
import pytest

import pressiodemoapps_amd as pda

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "ref", "tests_py")
S = pda.InviscidFluxReconstruction


def mesh_dir(case, tmp_path):
    """Unpack the case's mesh files into a directory, as the reference keeps them."""
    d = tmp_path / case
    d.mkdir()
    for f in ("info.dat", "connectivity.dat", "coordinates.dat"):
        with gzip.open(os.path.join(FIX, case, f + ".gz"), "rb") as fi, open(d / f, "wb") as fo:
            shutil.copyfileobj(fi, fo)
    return str(d)


def gold(case, name):
    return np.load(os.path.join(FIX, case, name))


def pressure(rho, u, v, E):
    # computePressure of tests_py/2d_sedov_full/weno5/test_2d_sedov_weno5.py:12-15
    gamma_ = (5. + 2.) / 5.
    vel = u ** 2 + v ** 2
    return (gamma_ - 1.) * (E - rho * vel * 0.5)


# (case, problem factory, integrator, dt, steps, quantity compared, gold file, rtol, atol, source)
def _default(enum, scheme):
    return lambda mesh: pda.create_problem(mesh, enum, scheme)


def _h(y):
    return y.reshape(-1, 3)[:, 0]


def _p(y):
    s = y.reshape(-1, 4)
    rho = s[:, 0]
    return pressure(rho, s[:, 1] / rho, s[:, 2] / rho, s[:, 3])


CASES = [
    ("1d_sod__weno3", _default(pda.Euler1d.Sod, S.Weno3), pda.advanceRK4, 0.001, 100, None, "gold.npy",
     1e-9, 1e-11, "tests_py/1d_sod/weno3/test_1d_sod_weno3.py:20-32"),
    ("1d_sod__weno5", _default(pda.Euler1d.Sod, S.Weno5), pda.advanceRK4, 0.001, 100, None, "gold.npy",
     1e-9, 1e-11, "tests_py/1d_sod/weno5/test_1d_sod_weno5.py:9-24"),
    ("1d_shu_osher__weno5", _default(pda.Euler1d.ShuOsher, S.Weno5), pda.advanceSSP3, 0.001, int(1.8 / 0.001),
     None, "gold.npy", 1e-9, 1e-11, "tests_py/1d_shu_osher/weno5/test_1d_shu_osher_weno5.py"),
    ("2d_sedov_full__weno3", _default(pda.Euler2d.SedovFull, S.Weno3), pda.advanceSSP3, 0.0001,
     int(0.05 / 0.0001), _p, "p_gold.npy", 1e-10, 1e-12, "tests_py/2d_sedov_full/weno3/test_2d_sedov_weno3.py"),
    ("2d_sedov_full__weno5", _default(pda.Euler2d.SedovFull, S.Weno5), pda.advanceSSP3, 0.0001,
     int(0.05 / 0.0001), _p, "p_gold.npy", 1e-10, 1e-12, "tests_py/2d_sedov_full/weno5/test_2d_sedov_weno5.py"),
    ("2d_swe_slip_wall__weno3", _default(pda.Swe2d.SlipWall, S.Weno3), pda.advanceRK4, 0.005, int(7.5 / 0.005),
     _h, "h_gold.npy", 1e-10, 1e-12, "tests_py/2d_swe_slip_wall/weno3/test_2d_swe_slip_wall_weno3.py"),
    ("2d_swe_slip_wall__weno5", _default(pda.Swe2d.SlipWall, S.Weno5), pda.advanceRK4, 0.005, int(7.5 / 0.005),
     _h, "h_gold.npy", 1e-10, 1e-12,
     "tests_py/2d_swe_slip_wall/weno5/test_2d_swe_gaussian_pulse_weno5.py"),
    ("2d_burgers_periodic",
     lambda mesh: pda.create_problem(mesh, pda.AdvectionDiffusion2d.BurgersPeriodic, S.Weno5,
                                     pda.ViscousFluxReconstruction.FirstOrder),
     pda.advanceRK4, 0.005, 2000, None, "gold.npy", 1e-9, 1e-11,
     "tests_py/2d_burgers_periodic/test_2d_burgers_periodic.py"),
    ("2d_advection_diffusion_reaction", _default(pda.AdvectionDiffusionReaction2d.ProblemA, S.Weno3),
     pda.advanceSSP3, 0.005, 1000, None, "gold1.npy", 1e-9, 1e-11,
     "tests_py/2d_advection_diffusion_reaction/test_2d_adv_diff_reac.py:11-26"),
    ("2d_advection_diffusion_reaction",
     lambda mesh: pda.create_adv_diff_reac_2d_problem_A(mesh, S.Weno3, 0.5 * np.cos(np.pi / 3),
                                                        0.5 * np.sin(np.pi / 3), 0.0005, 2.0),
     pda.advanceSSP3, 0.005, 1000, None, "gold2.npy", 1e-9, 1e-11,
     "tests_py/2d_advection_diffusion_reaction/test_2d_adv_diff_reac.py:28-45"),
    ("3d_sedov_symmetry__firstorder_nxnynz_equal", _default(pda.Euler3d.SedovSymmetry, S.FirstOrder),
     pda.advanceSSP3, 0.0001, int(0.025 / 0.0001), None, "gold_state.npy", 1e-8, 1e-10,
     "tests_py/3d_sedov_symmetry/firstorder_nxnynz_equal/test_3d_sedov_firstorder_nxnynz_equal.py"),
    ("3d_sedov_symmetry__weno3_nxnynz_equal", _default(pda.Euler3d.SedovSymmetry, S.Weno3),
     pda.advanceSSP3, 0.0001, int(0.025 / 0.0001), None, "gold_state.npy", 1e-8, 1e-10,
     "tests_py/3d_sedov_symmetry/weno3_nxnynz_equal/test_3d_sedov_weno3_nxnynz_equal.py"),
]
IDS = [f"{c[0]}-{c[6][:-4]}" for c in CASES]


@pytest.mark.gpu
@pytest.mark.parametrize("case,make,integrate,dt,nsteps,quantity,goldname,rtol,atol,source", CASES, ids=IDS)
def test_reference_python_gold(case, make, integrate, dt, nsteps, quantity, goldname, rtol, atol, source,
                               tmp_path):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    print(f"[tests_py] {case} {goldname}: {integrate.__name__} dt {dt} x {nsteps} ({source})", flush=True)
    mesh = pda.load_cellcentered_uniform_mesh(mesh_dir(case, tmp_path))
    app = make(mesh)
    yn = app.initialCondition()
    integrate(app, yn, dt, nsteps)
    torch.cuda.synchronize()
    y = yn.cpu().numpy()
    val = quantity(y) if quantity else y
    g = gold(case, goldname)
    assert val.shape == g.shape
    assert not np.isnan(val).any()
    err = np.abs(val - g).max() / np.abs(g).max()
    print(f"[tests_py] {case} {goldname}: max |diff| / max |gold| = {err:.2e}", flush=True)
    assert np.allclose(val, g, rtol=rtol, atol=atol)


@pytest.mark.gpu
def test_sod_weno3_velocity_size(tmp_path):
    # tests_py/1d_sod/weno3/test_1d_sod_weno3.py:12-18
    mesh = pda.load_cellcentered_uniform_mesh(mesh_dir("1d_sod__weno3", tmp_path))
    app = pda.create_problem(mesh, pda.Euler1d.Sod, S.Weno3)
    assert app.createRightHandSide().shape[0] == 300


def test_ccu_mesh_s3_5x5(tmp_path):
    """tests_py/ccu_mesh_s3_5x5/test_ccu_mesh_s3_5x5.py: sizes, spacing and
    cell centres of the reference's 5 x 5 periodic mesh, through both
    load_cellcentered_uniform_mesh and CellCenteredUniformMesh(path); the
    graph equals the reference's connectivity file (which the test's gold
    graph restates)."""
    d = mesh_dir("ccu_mesh_s3_5x5", tmp_path)
    for m in (pda.load_cellcentered_uniform_mesh(d), pda.CellCenteredUniformMesh(d)):
        assert m.dimensionality() == 2
        assert m.stencilMeshSize() == 25 and m.sampleMeshSize() == 25
        assert m.stencilSize() == 3
        assert m.graph().shape == (25, 5)
        assert m.dx() == 0.2 and m.dxInv() == 5.0
        assert m.dy() == 0.2 and m.dyInv() == 5.0
        centres = 0.1 + 0.2 * np.arange(5)
        assert np.allclose(m.viewX(), np.tile(centres, 5))
        assert np.allclose(m.viewY(), np.repeat(centres, 5))
        conn = np.loadtxt(os.path.join(d, "connectivity.dat"), dtype=np.int64)
        assert np.array_equal(m.graph(), conn)
        # periodic in both axes: left/front/right/back neighbours of cell 0
        assert list(m.graph()[0]) == [0, 4, 5, 1, 20]
