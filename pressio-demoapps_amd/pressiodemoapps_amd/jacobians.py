"""Timing of jacobian(), rightHandSideAndJacobian() and rightHandSide() on
the structured workloads outside the headline: WENO Jacobians (the tiled
and staged kernels) and first-order ones, each on a seeded +-1 % perturbation
of the problem's initial condition, with the algorithmic bytes (8 nnz + 8 m
N: values written once, state read once; velocity 8 m (Nst + N)) per call.

    python -m pressiodemoapps_amd.jacobians [--cases euler2d-weno5,burgers-weno5] [--reps 5]
        [--velocity-strip-height H]   (rows per strip of the velocity march; 0: the grid rule)

Cases:
  euler2d-weno5   2D Euler Riemann IC2, 1024^2, WENO5 (13 4x4 blocks per row)
  euler2d-weno3   2D Euler Riemann IC2, 1024^2, WENO3
  burgers-weno5   2D viscous Burgers, periodic, 512^2, WENO5 (the implicit workload's Jacobian)
  burgers-weno3   the same with WENO3
  burgers-outflow-weno5  2D viscous Burgers, outflow boundaries, 512^2, WENO5
  swe-weno3       shallow water, slip walls, 512^2, WENO3
  euler3d-weno3   3D Euler Sedov, 128^3, WENO3
"""
import argparse
import json
import sys

from . import (AdvectionDiffusion2d, Euler2d, Euler3d, InviscidFluxReconstruction as S, Swe2d, create_problem,
               create_structured_mesh)

CASES = {
    "euler2d-weno5": lambda: (create_structured_mesh((1024, 1024), (0.0, 1.0, 0.0, 1.0), 7), Euler2d.Riemann,
                              S.Weno5, 2, None),
    "euler2d-weno3": lambda: (create_structured_mesh((1024, 1024), (0.0, 1.0, 0.0, 1.0), 5), Euler2d.Riemann,
                              S.Weno3, 2, None),
    "burgers-weno5": lambda: (create_structured_mesh((512, 512), (-1.0, 1.0, -1.0, 1.0), 7, True),
                              AdvectionDiffusion2d.BurgersPeriodic, S.Weno5, 1, {"diffusion": 5e-3}),
    "burgers-weno3": lambda: (create_structured_mesh((512, 512), (-1.0, 1.0, -1.0, 1.0), 5, True),
                              AdvectionDiffusion2d.BurgersPeriodic, S.Weno3, 1, {"diffusion": 5e-3}),
    "burgers-outflow-weno5": lambda: (create_structured_mesh((512, 512), (-1.0, 1.0, -1.0, 1.0), 7),
                                      AdvectionDiffusion2d.BurgersOutflow, S.Weno5, 1,
                                      {"diffusion": 5e-3, "pulseX": 0.3, "pulseY": 0.2}),
    "swe-weno3": lambda: (create_structured_mesh((512, 512), (-5.0, 5.0, -5.0, 5.0), 5), Swe2d.SlipWall, S.Weno3,
                          1, None),
    "euler3d-weno3": lambda: (create_structured_mesh((128, 128, 128), (0.0, 1.2, 0.0, 1.2, 0.0, 1.2), 5),
                              Euler3d.SedovSymmetry, S.Weno3, 0, None),
}


def measure(name, reps, log, strip_height=0):
    import torch
    from .sedov3d import _time
    mesh, enum, scheme, ic, params = CASES[name]()
    p = create_problem(mesh, enum, scheme, ic, params)
    p.set_velocity_strip_height(strip_height)
    U0 = p.initialCondition()
    g = torch.Generator(device="cuda").manual_seed(0)
    U = U0 * (1.0 + 0.01 * (torch.rand(U0.numel(), dtype=torch.float64, device="cuda", generator=g) * 2 - 1))
    J = p.createJacobian()
    V = p.createRightHandSide()
    tj = _time(lambda: p.jacobian(U, 0.0, J), reps)
    tvj = _time(lambda: p.rightHandSideAndJacobian(U, 0.0, V, J), reps)
    tv = _time(lambda: p.rightHandSide(U, 0.0, V), reps)
    m, n_rows, nst = p.numDofPerCell(), p.totalDofSampleMesh(), p.totalDofStencilMesh()
    bytes_j = 8 * J.nonZeros() + 8 * nst
    r = {"case": name, "rows": n_rows // m, "jacobian_nnz": J.nonZeros(), "jacobian_us": tj * 1e6,
         "jacobian_GBps": bytes_j / tj / 1e9, "rhs_and_jacobian_us": tvj * 1e6,
         "rhs_and_jacobian_GBps": (bytes_j + 8 * n_rows) / tvj / 1e9,
         "rhs_us": tv * 1e6, "rhs_GBps": 8 * (nst + n_rows) / tv / 1e9,
         "finite": bool(torch.isfinite(J.values).all().item() and torch.isfinite(V).all().item())}
    log(f"[jacobians] {json.dumps(r)}")
    return r


def main(argv=None):
    ap = argparse.ArgumentParser(prog="pressiodemoapps_amd.jacobians")
    ap.add_argument("--cases", default="euler2d-weno5,burgers-weno5,swe-weno3")
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--velocity-strip-height", type=int, default=0)
    a = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    print(json.dumps([measure(c, a.reps, log, a.velocity_strip_height) for c in a.cases.split(",")]))


if __name__ == "__main__":
    main()
