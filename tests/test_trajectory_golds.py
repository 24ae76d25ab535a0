"""Trajectory golds the reference's own tests commit (tests_cpp/*/gold*.txt),
reproduced with the CPU oracle's residual and the pressio integrator formulas
(tests_cpp/pressio/include/pressio/ode/impl/ode_explicit_stepper_without_mass_matrix.hpp:230-352).
Meshes come from the product generator (pinned by tests_mesh) and initial
conditions from the product set-up; both are what the reference test builds."""
import ctypes
import os

import numpy as np
import pytest

import pyoracle
import pressiodemoapps_amd as pda

TRAJ = os.path.join(os.path.dirname(__file__), "golden", "ref", "traj")
S = pda.InviscidFluxReconstruction


def rk4(f, y, dt, n):
    # ode_explicit_stepper_without_mass_matrix.hpp:284-352
    for _ in range(n):
        k1 = f(y)
        k2 = f(y + (dt / 2) * k1)
        k3 = f(y + (dt / 2) * k2)
        k4 = f(y + dt * k3)
        y = y + (dt / 6) * k1 + (dt / 3) * k2 + (dt / 3) * k3 + (dt / 6) * k4
    return y


def ssprk3(f, y, dt, n):
    # ode_explicit_stepper_without_mass_matrix.hpp:230-281
    for _ in range(n):
        a = y + dt * f(y)
        a = 0.25 * a + 0.75 * y + (0.25 * dt) * f(a)
        y = (1. / 3.) * y + (2. / 3.) * a + (2. / 3. * dt) * f(a)
    return y


def _setup(enum, scheme, n, bounds, st, per=(), ic=0, params=None, oparams=(1.4,)):
    mesh = pda.create_structured_mesh(n, bounds, st, per)
    prob = pda.create_problem(mesh, enum, scheme, ic, params)
    U0 = np.zeros(prob.totalDofStencilMesh())
    assert pda.lib().pda_initial_condition_host(prob.handle, U0.ctypes.data_as(ctypes.c_void_p)) == 0
    orc = pyoracle.Oracle()
    ms = orc.from_mesh(mesh)
    pb = orc.problem(pda.FAMILY[type(enum)], int(enum), int(scheme), list(oparams))
    return U0, (lambda U: orc.velocity(ms, pb, U))


@pytest.mark.parametrize("scheme,st,name", [(S.FirstOrder, 3, "firstorder"), (S.Weno3, 5, "weno3"),
                                            (S.Weno5, 7, "weno5")])
def test_sod1d_rk4(scheme, st, name):
    # tests_cpp/eigen_1d_euler_sod_explicit: 100 cells on [-0.5,0.5], RK4 dt 1e-3 x 100
    U0, f = _setup(pda.Euler1d.Sod, scheme, (100,), (-0.5, 0.5), st)
    U = rk4(f, U0, 1e-3, 100)
    gold = np.load(os.path.join(TRAJ, f"eigen_1d_euler_sod_explicit__{name}__gold.npy"))
    assert np.allclose(U, gold, rtol=1e-9, atol=1e-11)


@pytest.mark.parametrize("scheme,st,name", [(S.FirstOrder, 3, "firstorder"), (S.Weno3, 7, "weno3"),
                                            (S.Weno5, 7, "weno5")])
def test_riemann2d_ssprk3(scheme, st, name):
    # tests_cpp/eigen_2d_euler_riemann_explicit: 20x20 on [0,1]^2, IC2, SSPRK3 dt 0.01 x 60
    U0, f = _setup(pda.Euler2d.Riemann, scheme, (20, 20), (0.0, 1.0, 0.0, 1.0), st, ic=2)
    U = ssprk3(f, U0, 0.01, 60).reshape(-1, 4)
    rho = U[:, 0]
    p = 0.4 * (U[:, 3] - 0.5 * (U[:, 1] ** 2 + U[:, 2] ** 2) / rho)
    base = f"eigen_2d_euler_riemann_explicit__{name}__"
    assert np.allclose(rho, np.load(os.path.join(TRAJ, base + "rho_gold.npy")), rtol=1e-10, atol=1e-12)
    assert np.allclose(p, np.load(os.path.join(TRAJ, base + "p_gold.npy")), rtol=1e-9, atol=1e-11)


@pytest.mark.parametrize("scheme,st,name", [(S.FirstOrder, 3, "firstorder"), (S.Weno3, 5, "weno3"),
                                            (S.Weno5, 7, "weno5")])
def test_swe_slipwall_rk4(scheme, st, name):
    # tests_cpp/eigen_2d_swe_slip_wall_explicit: 25x25 on [-5,5]^2, RK4 dt 0.01 x 200
    U0, f = _setup(pda.Swe2d.SlipWall, scheme, (25, 25), (-5.0, 5.0, -5.0, 5.0), st,
                   oparams=(9.8, -3.0))
    U = rk4(f, U0, 0.01, 200).reshape(-1, 3)
    gold = np.load(os.path.join(TRAJ, f"eigen_2d_swe_slip_wall_explicit__{name}__h_gold.npy"))
    assert np.allclose(U[:, 0], gold, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("scheme,st,name", [(S.FirstOrder, 3, "firstorder"), (S.Weno3, 5, "weno3"),
                                            (S.Weno5, 7, "weno5")])
def test_burgers_periodic_rk4(scheme, st, name):
    # tests_cpp/eigen_2d_burgers_periodic_explicit: 20x20 on [-1,1]^2 periodic, RK4 dt 0.01 x 200
    U0, f = _setup(pda.AdvectionDiffusion2d.BurgersPeriodic, scheme, (20, 20), (-1.0, 1.0, -1.0, 1.0),
                   st, per=True, oparams=(1e-5,))
    U = rk4(f, U0, 0.01, 200)
    gold = np.load(os.path.join(TRAJ, f"eigen_2d_burgers_periodic_explicit__{name}__gold.npy"))
    assert np.allclose(U, gold, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("scheme,st,name", [(S.FirstOrder, 3, "firstorder"), (S.Weno3, 5, "weno3")])
def test_sedov3d_ssprk3(scheme, st, name):
    # tests_cpp/eigen_3d_euler_sedov_symmetry: 20^3 on [0,0.4]^3, SSPRK3 dt 1e-4 x 250
    U0, f = _setup(pda.Euler3d.SedovSymmetry, scheme, (20, 20, 20), (0.0, 0.4, 0.0, 0.4, 0.0, 0.4), st)
    U = ssprk3(f, U0, 1e-4, 250)
    gold = np.load(os.path.join(TRAJ, f"eigen_3d_euler_sedov_symmetry__{name}_nxnynz_equal__gold_state.npy"))
    assert np.allclose(U, gold, rtol=1e-8, atol=1e-10)


# ------------------------------------------------------------- implicit
def newton_implicit(scheme, y_n, y_nm1, dt, fj, tol, f_n=None):
    """One implicit step with pressio's discrete residual/Jacobian
    (ode_implicit_discrete_residual.hpp:112-240, ode_implicit_discrete_jacobian.hpp:60-121)
    and its Newton loop: solve, stop when ||dx|| < tol without applying dx
    (solvers_nonlinear/impl/nonlinear_least_squares.hpp:70-84); the linear
    solves are direct (pressio's BiCGSTAB runs to machine precision)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    cnp1, cn, cnm1, cf = 1.0, -1.0, 0.0, -dt
    if scheme == "bdf2":
        cn, cnm1, cf = -4.0 / 3.0, 1.0 / 3.0, -(2.0 / 3.0) * dt
    elif scheme == "cn":
        cf = -0.5 * dt
    y = y_n.copy()
    n = len(y)
    for _ in range(100):
        F, rp, ci, vals = fj(y)
        R = cnp1 * y + cn * y_n + cf * F
        if scheme == "bdf2":
            R += cnm1 * y_nm1
        if scheme == "cn":
            R += -0.5 * dt * f_n
        A = sp.csr_matrix((cf * vals, ci, rp), shape=(n, n)) + cnp1 * sp.identity(n, format="csr")
        dx = spla.spsolve(A.tocsc(), -R)
        if np.linalg.norm(dx) < tol:
            return y
        y = y + dx
    raise AssertionError("Newton did not converge")


@pytest.mark.parametrize("scheme,st,name", [(S.FirstOrder, 3, "firstorder"), (S.Weno3, 5, "weno3")])
def test_sod1d_bdf1_implicit(scheme, st, name):
    # tests_cpp/eigen_1d_euler_sod_implicit: 100 cells on [-0.5,0.5], BDF1,
    # dt 1e-3 x 100, Newton tolerance 1e-6, gold compared at rtol 1e-9, atol 1e-11
    mesh = pda.create_structured_mesh((100,), (-0.5, 0.5), st)
    prob = pda.create_problem(mesh, pda.Euler1d.Sod, scheme)
    U = np.zeros(prob.totalDofStencilMesh())
    assert pda.lib().pda_initial_condition_host(prob.handle, U.ctypes.data_as(ctypes.c_void_p)) == 0
    orc = pyoracle.Oracle()
    ms = orc.from_mesh(mesh)
    pb = orc.problem(1, int(pda.Euler1d.Sod), int(scheme), [1.4])
    fj = lambda y: orc.velocity_and_jacobian(ms, pb, y)  # noqa: E731
    for _ in range(100):
        U = newton_implicit("bdf1", U, None, 1e-3, fj, 1e-6)
    gold = np.load(os.path.join(TRAJ, f"eigen_1d_euler_sod_implicit__{name}__gold.npy"))
    assert np.allclose(U, gold, rtol=1e-9, atol=1e-11)


def test_riemann2d_cn_implicit():
    # tests_cpp/eigen_2d_euler_riemann_implicit/firstorder: 20x20, IC2,
    # Crank-Nicolson dt 0.02 x 30, Newton tolerance 1e-5
    mesh = pda.create_structured_mesh((20, 20), (0.0, 1.0, 0.0, 1.0), 3)
    prob = pda.create_problem(mesh, pda.Euler2d.Riemann, S.FirstOrder, 2)
    U = np.zeros(prob.totalDofStencilMesh())
    assert pda.lib().pda_initial_condition_host(prob.handle, U.ctypes.data_as(ctypes.c_void_p)) == 0
    orc = pyoracle.Oracle()
    ms = orc.from_mesh(mesh)
    pb = orc.problem(2, int(pda.Euler2d.Riemann), 0, [1.4])
    fj = lambda y: orc.velocity_and_jacobian(ms, pb, y)  # noqa: E731
    for _ in range(30):
        fn = orc.velocity_and_jacobian(ms, pb, U)[0]
        U = newton_implicit("cn", U, None, 0.02, fj, 1e-5, f_n=fn)
    U = U.reshape(-1, 4)
    rho = U[:, 0]
    p = 0.4 * (U[:, 3] - 0.5 * (U[:, 1] ** 2 + U[:, 2] ** 2) / rho)
    base = "eigen_2d_euler_riemann_implicit__firstorder__"
    assert np.allclose(rho, np.load(os.path.join(TRAJ, base + "rho_gold.npy")), rtol=1e-10, atol=1e-12)
    assert np.allclose(p, np.load(os.path.join(TRAJ, base + "p_gold.npy")), rtol=1e-10, atol=1e-12)


ADR_SCHEMES = [(S.FirstOrder, 3, "firstorder"), (S.Weno3, 5, "weno3"), (S.Weno5, 7, "weno5")]
ADR_DEFAULT = (0.5 * np.cos(np.pi / 3), 0.5 * np.sin(np.pi / 3), 0.001, 1.0)


@pytest.mark.parametrize("scheme,st,name", ADR_SCHEMES)
def test_adr_probA_rk4(scheme, st, name):
    # tests_cpp/eigen_2d_advdiffreac_probA_explicit: 18x18 on [0,1]^2, zero
    # initial state, default parameters, RK4 dt 0.01 x 500
    U0, f = _setup(pda.AdvectionDiffusionReaction2d.ProblemA, scheme, (18, 18), (0.0, 1.0, 0.0, 1.0), st,
                   oparams=ADR_DEFAULT)
    assert not U0.any()
    U = rk4(f, U0, 0.01, 500)
    gold = np.load(os.path.join(TRAJ, f"eigen_2d_advdiffreac_probA_explicit__{name}__gold.npy"))
    assert np.allclose(U, gold, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("scheme,st,name", ADR_SCHEMES)
def test_adr_probA_cn_implicit(scheme, st, name):
    # tests_cpp/eigen_2d_advdiffreac_probA_implicit: 17x17 on [0,1]^2,
    # Crank-Nicolson dt 0.05 x 40, Newton tolerance 1e-5
    mesh = pda.create_structured_mesh((17, 17), (0.0, 1.0, 0.0, 1.0), st)
    orc = pyoracle.Oracle()
    ms = orc.from_mesh(mesh)
    pb = orc.problem(6, 0, int(scheme), list(ADR_DEFAULT))
    fj = lambda y: orc.velocity_and_jacobian(ms, pb, y)  # noqa: E731
    U = np.zeros(mesh.stencilMeshSize())
    for _ in range(40):
        fn = orc.velocity(ms, pb, U)
        U = newton_implicit("cn", U, None, 0.05, fj, 1e-5, f_n=fn)
    gold = np.load(os.path.join(TRAJ, f"eigen_2d_advdiffreac_probA_implicit__{name}__gold.npy"))
    assert np.allclose(U, gold, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("scheme,st,name", [(S.FirstOrder, 3, "firstorder"), (S.Weno3, 5, "weno3")])
def test_burgers_outflow_cn_implicit(scheme, st, name):
    # tests_cpp/eigen_2d_burgers_outflow_implicit: 20x20 on [-1,1]^2, default
    # pulse and diffusion, Crank-Nicolson dt 0.01 x 200, Newton tolerance
    # 1e-5, final state compared at rtol 1e-10, atol 1e-12 (compare.py)
    mesh = pda.create_structured_mesh((20, 20), (-1.0, 1.0, -1.0, 1.0), st)
    prob = pda.create_problem(mesh, pda.AdvectionDiffusion2d.BurgersOutflow, scheme, 1)
    U = np.zeros(prob.totalDofStencilMesh())
    assert pda.lib().pda_initial_condition_host(prob.handle, U.ctypes.data_as(ctypes.c_void_p)) == 0
    orc = pyoracle.Oracle()
    ms = orc.from_mesh(mesh)
    pb = orc.problem(5, int(pda.AdvectionDiffusion2d.BurgersOutflow), int(scheme), [1e-5])
    fj = lambda y: orc.velocity_and_jacobian(ms, pb, y)  # noqa: E731
    for _ in range(200):
        fn = orc.velocity(ms, pb, U)
        U = newton_implicit("cn", U, None, 0.01, fj, 1e-5, f_n=fn)
    gold = np.load(os.path.join(TRAJ, f"eigen_2d_burgers_outflow_implicit__{name}__gold.npy"))
    assert np.allclose(U, gold, rtol=1e-10, atol=1e-12)
