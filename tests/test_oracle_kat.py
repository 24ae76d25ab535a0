"""Pins of the CPU oracle that do not need the reference build:
known-answer vectors from the reference's own tests, and properties the
mathematics fixes (consistency, exactness, derivative checks, conservation).
Chosen so a dropped term, a sign error or a transposed operand in the oracle
fails at least one of them."""
import numpy as np
import pytest

import cases
import pyoracle
import pressiodemoapps_amd as pda


# ---------------------------------------------------------------- WENO KATs
def test_weno5_known_answer(oracle):
    # tests_cpp/weno5/main.cc:12-33: q = {8,5,0,1,2,1,4}, tolerance 1e-13
    q = [8., 5., 0., 1., 2., 1., 4.]
    gold = [0.498169371091894, 0.498239224509689, 1.50249211669906, 1.53308488724067]
    lm, lp = oracle.weno5(q[0:6])
    rm, rp = oracle.weno5(q[1:7])
    assert np.allclose([lm, lp, rm, rp], gold, rtol=0, atol=1e-13)


def test_weno3_known_answer(oracle):
    # tests_cpp/weno3/main.cc:10-33: q = {5,3,0,1,2}
    q = [5., 3., 0., 1., 2.]
    gold = [1.85840705145274, -0.451807145304084, 0.487730039670285, 0.5]
    lm, lp = oracle.weno3(q[0:4])
    rm, rp = oracle.weno3(q[1:5])
    assert np.allclose([lm, lp, rm, rp], gold, rtol=0, atol=1e-13)


@pytest.mark.parametrize("w", [3, 5])
def test_weno_exact_on_linear_and_constant_data(oracle, w):
    f = oracle.weno5 if w == 5 else oracle.weno3
    n = 6 if w == 5 else 4
    c = n // 2 - 1  # face between points c and c+1
    for a, b in [(0.0, 2.5), (1.3, -0.7), (-4.0, 0.0)]:
        q = a + b * np.arange(n)
        un, up = f(q)
        exact = a + b * (c + 0.5)
        assert abs(un - exact) < 1e-12 and abs(up - exact) < 1e-12


@pytest.mark.parametrize("w", [3, 5])
def test_weno_gradient_matches_finite_differences_of_values(oracle, w):
    """tests_cpp/weno5/main_jacobians.cc checks the gradient by complex step;
    here: Richardson-extrapolated central differences of the oracle's own
    value function (no formula retyped)."""
    f = oracle.weno5 if w == 5 else oracle.weno3
    n = 6 if w == 5 else 4
    rng = np.random.default_rng(7)
    samples = [np.array([0.4, 8., 3., 0.5, 0.9, 128.3][:n])] + [rng.uniform(0.5, 2.0, n) for _ in range(20)]
    for q in samples:
        un, up, gn, gp = f(q, grad=True)
        v0 = f(q)
        assert un == v0[0] or abs(un - v0[0]) < 1e-14 * max(1, abs(un))
        for i in range(n):
            def d(h):
                qp, qm = q.copy(), q.copy()
                qp[i] += h
                qm[i] -= h
                a, b = np.array(f(qp)), np.array(f(qm))
                return (a - b) / (2 * h)
            h = 1e-4 * max(1.0, abs(q[i]))
            fd = (4 * d(h / 2) - d(h)) / 3
            scale = max(1.0, np.abs([gn[i], gp[i]]).max())
            assert abs(fd[0] - gn[i]) < 1e-6 * scale, (i, fd[0], gn[i])
            assert abs(fd[1] - gp[i]) < 1e-6 * scale, (i, fd[1], gp[i])


# -------------------------------------------------------------- flux checks
def _phys_flux(fam, q, n, p):
    """closed-form physical flux f(q).n of each family"""
    if fam in (1, 2, 3):
        g = p[0]
        dim = {1: 1, 2: 2, 3: 3}[fam]
        r = q[0]
        vel = q[1:1 + dim] / r
        E = q[-1]
        pr = (g - 1) * (E - 0.5 * r * vel @ vel)
        un = vel @ n[:dim]
        return np.concatenate([[r * un], r * un * vel + pr * n[:dim], [(E + pr) * un]])
    if fam == 4:
        g = p[0]
        h, u, v = q[0], q[1] / q[0], q[2] / q[0]
        un = u * n[0] + v * n[1]
        return np.array([h * un, h * un * u + 0.5 * g * h * h * n[0], h * un * v + 0.5 * g * h * h * n[1]])
    if fam == 5:
        u, v = q
        return np.array([0.5 * (u * u * n[0] + u * v * n[1]), 0.5 * (u * v * n[0] + v * v * n[1])])
    raise ValueError


def _states(fam, rng):
    if fam == 1:
        r, u, pr = rng.uniform(0.5, 2), rng.uniform(-1, 1), rng.uniform(0.5, 2)
        return np.array([r, r * u, pr / 0.4 + 0.5 * r * u * u])
    if fam == 2:
        r, u, v, pr = rng.uniform(0.5, 2), rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(0.5, 2)
        return np.array([r, r * u, r * v, pr / 0.4 + 0.5 * r * (u * u + v * v)])
    if fam == 3:
        r, pr = rng.uniform(0.5, 2), rng.uniform(0.5, 2)
        vel = rng.uniform(-1, 1, 3)
        return np.concatenate([[r], r * vel, [pr / 0.4 + 0.5 * r * vel @ vel]])
    if fam == 4:
        h = rng.uniform(0.8, 1.5)
        return np.array([h, h * rng.uniform(-1, 1), h * rng.uniform(-1, 1)])
    return rng.uniform(-1.0, 1.0, 2)


FAMS = [(1, 1, [1.4]), (2, 2, [1.4]), (3, 3, [1.4]), (4, 2, [9.8, -3.0]), (5, 2, [1e-3])]


@pytest.mark.parametrize("fam,dim,p", FAMS, ids=["euler1d", "euler2d", "euler3d", "swe", "burgers"])
def test_rusanov_flux_is_consistent(oracle, fam, dim, p):
    """F(q, q) = f(q).n: the dissipation term vanishes for equal states."""
    rng = np.random.default_rng(fam)
    pb = oracle.problem(fam, 0, 0, p)
    for _ in range(20):
        q = _states(fam, rng)
        for a in range(dim):
            n = np.zeros(3)
            n[a] = 1.0
            F = oracle.face_flux(pb, dim, q, q, n)
            assert np.allclose(F, _phys_flux(fam, q, n, p), rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("fam,dim,p", FAMS, ids=["euler1d", "euler2d", "euler3d", "swe", "burgers"])
def test_flux_jacobians_match_finite_differences(oracle, fam, dim, p):
    """dF/dqL and dF/dqR of the oracle's flux, including the wave-speed
    derivative (tests_cpp/eigen_rusanov_flux_jacobians_euler checks this by FD)."""
    rng = np.random.default_rng(10 + fam)
    pb = oracle.problem(fam, 0, 0, p)
    m = len(_states(fam, rng))
    for _ in range(10):
        qL, qR = _states(fam, rng), _states(fam, rng)
        if fam == 5:  # keep clear of the |qL|=|qR| kink of the Burgers wave speed
            qR = qR + np.sign(qR) * 0.05
        for a in range(dim):
            n = np.zeros(3)
            n[a] = 1.0
            JL, JR = oracle.face_flux_jac(pb, dim, qL, qR, n)
            for side, J in ((0, JL), (1, JR)):
                fd = np.zeros((m, m))
                for j in range(m):
                    h = 1e-6 * max(1.0, abs((qL, qR)[side][j]))
                    qp = [qL.copy(), qR.copy()]
                    qm = [qL.copy(), qR.copy()]
                    qp[side][j] += h
                    qm[side][j] -= h
                    fd[:, j] = (oracle.face_flux(pb, dim, qp[0], qp[1], n)
                                - oracle.face_flux(pb, dim, qm[0], qm[1], n)) / (2 * h)
                assert np.allclose(J, fd, rtol=1e-6, atol=1e-7), (side, a, J, fd)


@pytest.mark.parametrize("fam,dim,p", FAMS[:3], ids=["euler1d", "euler2d", "euler3d"])
def test_flux_jacobians_sum_to_physical_jacobian(oracle, fam, dim, p):
    """At qL = qR the dissipation parts cancel: JL + JR = A(q) = df/dq."""
    rng = np.random.default_rng(20 + fam)
    pb = oracle.problem(fam, 0, 0, p)
    for _ in range(10):
        q = _states(fam, rng)
        m = len(q)
        for a in range(dim):
            n = np.zeros(3)
            n[a] = 1.0
            JL, JR = oracle.face_flux_jac(pb, dim, q, q, n)
            A = np.zeros((m, m))
            for j in range(m):
                h = 1e-6
                qp, qm = q.copy(), q.copy()
                qp[j] += h
                qm[j] -= h
                A[:, j] = (_phys_flux(fam, qp, n, p) - _phys_flux(fam, qm, n, p)) / (2 * h)
            assert np.allclose(JL + JR, A, rtol=1e-6, atol=1e-7)


# ----------------------------------------------------------- conservation
@pytest.mark.parametrize("case", [c for c in cases.CASES if "periodic" in c[0]], ids=lambda c: c[0])
def test_residual_is_conservative_on_periodic_meshes(oracle, case):
    """On a fully periodic mesh every face flux enters two cells with opposite
    signs and the Laplacian telescopes: sum over cells of V = 0 per dof."""
    mesh, prob = cases.make(case)
    U = cases.perturbed_state(prob, 11)
    V = oracle.velocity(oracle.from_mesh(mesh), cases.oracle_problem(case), U)
    m = prob.numDofPerCell()
    S = V.reshape(-1, m).sum(axis=0)
    assert np.abs(S).max() <= 1e-10 * np.abs(V).max() * len(V)


@pytest.mark.parametrize("case", [c for c in cases.CASES if "Weno5" in c[0]], ids=lambda c: c[0])
def test_jacobian_matches_finite_difference_of_residual(oracle, case):
    """J a ~ (V(U + eps a) - V(U)) / eps on inner rows, as
    tests_cpp/eigen_1d_euler_jacobian_fd/main.cc:76-101 checks; near-boundary
    rows carry the reference's first-order Jacobian by design and are
    excluded."""
    import scipy.sparse as sp
    mesh, prob = cases.make(case)
    rng = np.random.default_rng(13)
    # a small random offset keeps momenta away from exactly zero, where the
    # |u| in the Rusanov wave speed has its kink
    m = prob.numDofPerCell()
    # momenta only for Euler; small for the Sedov background (energy ~1e-5)
    amp = 1e-3 if case[1] == pda.Euler3d.SedovSymmetry else 0.01
    off = amp * rng.uniform(-1, 1, prob.totalDofStencilMesh())
    if type(case[1]) in (pda.Euler1d, pda.Euler2d, pda.Euler3d):
        off.reshape(-1, m)[:, [0, m - 1]] = 0.0
    U = cases.perturbed_state(prob, 12) + off
    ms, pb = oracle.from_mesh(mesh), cases.oracle_problem(case)
    V, rp, ci, vals = oracle.velocity_and_jacobian(ms, pb, U)
    J = sp.csr_matrix((vals, ci, rp), shape=(len(V), len(U)))
    a = rng.uniform(-1, 1, len(U))
    eps = 1e-7
    # central difference of the velocity path of the Jacobian evaluation
    Vp = oracle.velocity_and_jacobian(ms, pb, U + eps * a)[0]
    Vm = oracle.velocity_and_jacobian(ms, pb, U - eps * a)[0]
    fd = (Vp - Vm) / (2 * eps)
    ja = J @ a
    inner = np.repeat(_inner_rows(mesh), m)
    # scale by |J||a| per row: J a itself cancels strongly on smooth data
    scale = (abs(J) @ np.abs(a))[inner].max()
    err = np.abs(ja[inner] - fd[inner]).max() / scale
    assert err < 1e-7, err


def _inner_rows(mesh):
    G = mesh.graph()
    return np.all(G >= 0, axis=1)


@pytest.mark.parametrize("scheme", [pda.InviscidFluxReconstruction.FirstOrder, pda.InviscidFluxReconstruction.Weno3,
                                    pda.InviscidFluxReconstruction.Weno5], ids=lambda s: s.name)
def test_sedov3d_axis_permutation_equivariance(oracle, scheme):
    """On a cube with the same boundary treatment on every axis, exchanging
    two axes of the state (cells and momentum components) exchanges them in
    the velocity: V(P U) = P V(U).  Pins the 3D axis handling, including the
    WENO5 ghost layers the reference does not define in 3D."""
    import pyoracle
    n = 9
    st = 7 if scheme == pda.InviscidFluxReconstruction.Weno5 else 5
    mesh = pda.create_structured_mesh((n, n, n), (0, 1.2, 0, 1.2, 0, 1.2), st)
    prob = pda.create_problem(mesh, pda.Euler3d.SedovSymmetry, scheme, 0, None)
    ms = oracle.from_mesh(mesh)
    pb = pyoracle.Oracle.problem(3, int(pda.Euler3d.SedovSymmetry), int(scheme), [1.4])
    U = cases.perturbed_state(prob, 13, amp=0.05).reshape(n, n, n, 5)  # [k][j][i][dof]
    V = oracle.velocity(ms, pb, U.ravel()).reshape(n, n, n, 5)
    for (a, b) in ((0, 1), (0, 2), (1, 2)):
        # axis a of the cell grid (x = last index) <-> axis b; momentum 1+a <-> 1+b
        ga, gb = 2 - a, 2 - b
        Up = np.swapaxes(U, ga, gb).copy()
        Up[..., [1 + a, 1 + b]] = Up[..., [1 + b, 1 + a]]
        Vp = oracle.velocity(ms, pb, Up.ravel()).reshape(n, n, n, 5)
        Ve = np.swapaxes(V, ga, gb).copy()
        Ve[..., [1 + a, 1 + b]] = Ve[..., [1 + b, 1 + a]]
        assert np.abs(Vp - Ve).max() <= 1e-12 * np.abs(V).max(), (a, b)
