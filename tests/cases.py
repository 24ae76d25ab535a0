"""Shared problem cases for the parity tests (family, problem enum, scheme,
mesh shape).  Small enough that the CPU oracle finishes in milliseconds."""
import numpy as np

import pressiodemoapps_amd as pda

S = pda.InviscidFluxReconstruction

# (id, problem enum, scheme, n, bounds, mesh stencil, periodic, ic flag, params)
CASES = []
for s in (S.FirstOrder, S.Weno3, S.Weno5):
    CASES += [
        (f"euler1d-sod-{s.name}", pda.Euler1d.Sod, s, (64,), (-0.5, 0.5), 7, (), 0, None),
        (f"euler1d-periodic-{s.name}", pda.Euler1d.PeriodicSmooth, s, (40,), (-1, 1), 7, True, 0, None),
        (f"euler2d-riemann-{s.name}", pda.Euler2d.Riemann, s, (24, 20), (0, 1, 0, 1), 7, (), 2, None),
        (f"euler2d-periodic-{s.name}", pda.Euler2d.PeriodicSmooth, s, (16, 18), (-1, 1, -1, 1), 7, True, 0, None),
        (f"swe-slipwall-{s.name}", pda.Swe2d.SlipWall, s, (20, 22), (-5, 5, -5, 5), 7, (), 1,
         {"gravity": 9.8, "coriolis": -3.0}),
        (f"burgers-periodic-{s.name}", pda.AdvectionDiffusion2d.BurgersPeriodic, s, (18, 16),
         (-1, 1, -1, 1), 7, True, 1, {"diffusion": 1e-4}),
        # homogeneous Dirichlet at the left and back sides, Neumann at the
        # right and front ones; a pulse near the right/front corner
        (f"burgers-outflow-{s.name}", pda.AdvectionDiffusion2d.BurgersOutflow, s, (18, 16),
         (-1, 1, -1, 1), 7, (), 1, {"diffusion": 5e-3, "pulseX": 0.6, "pulseY": 0.5}),
    ]
for s in (S.FirstOrder, S.Weno3):
    CASES += [
        (f"euler3d-sedov-{s.name}", pda.Euler3d.SedovSymmetry, s, (10, 9, 8), (0, 1.2, 0, 1.2, 0, 1.2), 5, (), 0, None),
        (f"euler3d-periodic-{s.name}", pda.Euler3d.PeriodicSmooth, s, (8, 8, 8), (-1, 1, -1, 1, -1, 1), 5, True, 0, None),
    ]

# advection-diffusion-reaction problem A (homogeneous Dirichlet walls, zero
# initial state), default parameters and a stronger-diffusion set
ADR_PARAMS = {"ux": 0.35, "uy": -0.6, "diffusion": 0.02, "sigma": 0.7}
for s in (S.FirstOrder, S.Weno3, S.Weno5):
    CASES += [
        (f"adr-probA-{s.name}", pda.AdvectionDiffusionReaction2d.ProblemA, s, (18, 16), (0, 1, 0, 1), 7, (), 0,
         None),
        (f"adr-custom-{s.name}", pda.AdvectionDiffusionReaction2d.ProblemA, s, (15, 19), (0, 1, 0, 1), 7, (), 0,
         dict(ADR_PARAMS)),
    ]

# 3D WENO5 (build-defined: the reference stops at WENO3 in 3D); pinned by
# conservation and axis-permutation equivariance in test_oracle_kat.py
CASES += [
    ("euler3d-sedov-Weno5", pda.Euler3d.SedovSymmetry, S.Weno5, (10, 9, 8), (0, 1.2, 0, 1.2, 0, 1.2), 7, (), 0, None),
    ("euler3d-periodic-Weno5", pda.Euler3d.PeriodicSmooth, S.Weno5, (8, 8, 8), (-1, 1, -1, 1, -1, 1), 7, True, 0, None),
]

IDS = [c[0] for c in CASES]

# larger meshes: several 64-wide tiles, partial segments, tall and wide boxes
BIG_CASES = [
    ("euler2d-riemann-big-FirstOrder", pda.Euler2d.Riemann, S.FirstOrder, (150, 37), (0, 1, 0, 1), 7, (), 2, None),
    ("euler2d-riemann-big-Weno5", pda.Euler2d.Riemann, S.Weno5, (150, 37), (0, 1, 0, 1), 7, (), 2, None),
    ("euler2d-periodic-big-Weno3", pda.Euler2d.PeriodicSmooth, S.Weno3, (70, 131), (-1, 1, -1, 1), 5, True, 0, None),
    ("swe-big-FirstOrder", pda.Swe2d.SlipWall, S.FirstOrder, (97, 70), (-5, 5, -5, 5), 7, (), 1,
     {"gravity": 9.8, "coriolis": -3.0}),
    # WENO Jacobian with each face once over 8 x 3 tiles (partial tiles at
    # the box's far edges; periodic tiles at the wrap take graph and slot
    # tables)
    ("swe-big-Weno5", pda.Swe2d.SlipWall, S.Weno5, (61, 44), (-5, 5, -5, 5), 7, (), 1,
     {"gravity": 9.8, "coriolis": -3.0}),
    ("burgers-big-Weno3", pda.AdvectionDiffusion2d.BurgersPeriodic, S.Weno3, (45, 38), (-1, 1, -1, 1), 5,
     True, 1, {"diffusion": 5e-3}),
    ("burgers-big-Weno5", pda.AdvectionDiffusion2d.BurgersPeriodic, S.Weno5, (130, 66), (-1, 1, -1, 1), 7,
     True, 1, {"diffusion": 5e-3}),
    ("burgers-big-FirstOrder", pda.AdvectionDiffusion2d.BurgersPeriodic, S.FirstOrder, (130, 66), (-1, 1, -1, 1), 3,
     True, 1, {"diffusion": 5e-3}),
    ("burgers-outflow-big-Weno5", pda.AdvectionDiffusion2d.BurgersOutflow, S.Weno5, (130, 66), (-1, 1, -1, 1), 7,
     (), 1, {"diffusion": 5e-3, "pulseX": 0.7, "pulseY": 0.6}),
    ("burgers-outflow-big-FirstOrder", pda.AdvectionDiffusion2d.BurgersOutflow, S.FirstOrder, (97, 70),
     (-1, 1, -1, 1), 3, (), 1, {"diffusion": 5e-3, "pulseX": 0.7, "pulseY": 0.6}),
    # box lines of 38 cells: the 3D first-order Jacobian copies runs of 16
    # consecutive rows as one range
    ("euler3d-sedov-big-FirstOrder", pda.Euler3d.SedovSymmetry, S.FirstOrder, (40, 7, 6), (0, 1.2, 0, 1.2, 0, 1.2),
     7, (), 0, None),
    ("euler3d-sedov-big-Weno3", pda.Euler3d.SedovSymmetry, S.Weno3, (40, 7, 6), (0, 1.2, 0, 1.2, 0, 1.2),
     7, (), 0, None),
]
BIG_IDS = [c[0] for c in BIG_CASES]

ORACLE_PARAMS = {
    1: lambda p: [1.4],
    2: lambda p: [(p or {}).get("gamma", 1.4)],
    3: lambda p: [1.4],
    4: lambda p: [(p or {}).get("gravity", 9.8), (p or {}).get("coriolis", -3.0)],
    5: lambda p: [(p or {}).get("diffusion", 1e-5)],
    6: lambda p: [(p or {}).get("ux", 0.5 * np.cos(np.pi / 3)), (p or {}).get("uy", 0.5 * np.sin(np.pi / 3)),
                  (p or {}).get("diffusion", 0.001), (p or {}).get("sigma", 1.0)],
}


def make(case):
    _, enum, scheme, n, bounds, st, per, ic, params = case
    mesh = pda.create_structured_mesh(n, bounds, st, per)
    prob = pda.create_problem(mesh, enum, scheme, ic, params)
    return mesh, prob


def oracle_problem(case):
    import pyoracle
    _, enum, scheme, n, bounds, st, per, ic, params = case
    fam = pda.FAMILY[type(enum)]
    return pyoracle.Oracle.problem(fam, int(enum), int(scheme), ORACLE_PARAMS[fam](params))


def perturbed_state(prob, seed, amp=0.01):
    """IC times (1 + U(-amp, amp)) per entry, computed on the host."""
    import ctypes
    U0 = np.zeros(prob.totalDofStencilMesh())
    rc = pda.lib().pda_initial_condition_host(prob.handle, U0.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0
    rng = np.random.default_rng(seed)
    if not U0.any():  # zero initial state (advection-diffusion-reaction): a smooth bump plus noise
        n = U0.size
        return np.sin(np.linspace(0.0, 3.0, n)) ** 2 + amp * rng.uniform(-1.0, 1.0, n)
    return U0 * (1.0 + amp * rng.uniform(-1.0, 1.0, U0.shape))


def rel_err(a, b):
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / (den if den > 0 else 1.0))
