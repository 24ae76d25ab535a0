"""3D Euler Sedov blast with symmetry boundaries on [0, 1.2]^3: WENO5
velocity and first-order Jacobian (5 x 5 blocks, 7-point stencil) in FP64 on
the full mesh and on a sample mesh of a seeded random 5 % of the cells with
their WENO5 stencil cells.  The IC deposits energy 0.851072 in the cells
within 2 dx of the origin on a background of density 1 and pressure 1e-5.
WENO5 in 3D is build-defined (the reference stops at WENO3 in 3D).

    python -m pressiodemoapps_amd.sedov3d --n 256
"""
import argparse
import json
import sys
import time

import numpy as np

from . import Euler3d, InviscidFluxReconstruction as S, create_problem, create_sample_mesh, \
    create_structured_mesh

IC = {"sedovRadiusCells": 2.0, "sedovAmbientPressure": 1e-5, "sedovEnergy": 0.851072}


def _time(fn, reps):
    import torch
    fn()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    for _ in range(reps):
        fn()
    ev[1].record()
    torch.cuda.synchronize()
    return ev[0].elapsed_time(ev[1]) / reps * 1e-3


def measure(mesh, label, reps, log, jheights=()):
    import torch
    pa = create_problem(mesh, Euler3d.SedovSymmetry, S.Weno5, 0, dict(IC))
    pb = create_problem(mesh, Euler3d.SedovSymmetry, S.FirstOrder, 0, dict(IC))
    U0 = pa.initialCondition()
    g = torch.Generator(device="cuda").manual_seed(0)
    U = U0 * (1.0 + 0.01 * (torch.rand(U0.numel(), dtype=torch.float64, device="cuda", generator=g) * 2 - 1))
    V = pa.createRightHandSide()
    J = pb.createJacobian()
    rps, _ = pa.autotune(U)  # set-up: the velocity kernel's strip height
    rows = pa.totalDofSampleMesh() // 5
    nnz = J.nonZeros()
    tv = _time(lambda: pa.rightHandSide(U, 0.0, V), reps)
    tj = _time(lambda: pb.jacobian(U, 0.0, J), reps)
    bytes_v = 8 * 5 * (pa.totalDofStencilMesh() // 5 + rows)
    bytes_j = 8 * nnz + 8 * 5 * (pa.totalDofStencilMesh() // 5)
    r = {"rows": rows, "stencil_cells": pa.totalDofStencilMesh() // 5, "jacobian_nnz": nnz,
         "velocity_rows_per_strip": rps,
         "velocity_us": tv * 1e6, "velocity_rows_per_s": rows / tv, "velocity_GBps": bytes_v / tv / 1e9,
         "jacobian_us": tj * 1e6, "jacobian_rows_per_s": rows / tj, "jacobian_GBps": bytes_j / tj / 1e9,
         "finite": bool(torch.isfinite(V).all().item() and torch.isfinite(J.values).all().item())}
    if jheights:
        # the first-order Jacobian at fixed strip heights (lines per workgroup)
        by = {}
        for h in jheights:
            pb.set_jacobian_strip_height(h)
            by[h] = _time(lambda: pb.jacobian(U, 0.0, J), reps) * 1e6
        pb.set_jacobian_strip_height(0)
        r["jacobian_us_by_strip_height"] = by
    log(f"[sedov3d] {label}: {json.dumps(r)}")
    del J, V
    return r


def main(argv=None):
    ap = argparse.ArgumentParser(prog="pressiodemoapps_amd.sedov3d")
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--sample", type=float, default=0.05)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--jacobian-strip-heights", default="",
                    help="comma-separated strip heights at which the full mesh's Jacobian is also timed")
    a = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    t0 = time.perf_counter()
    full = create_structured_mesh((a.n, a.n, a.n), (0.0, 1.2, 0.0, 1.2, 0.0, 1.2), 7)
    log(f"[sedov3d] mesh {a.n}^3 in {time.perf_counter() - t0:.1f} s")
    jh = [int(h) for h in a.jacobian_strip_heights.split(",") if h]
    out = {"n": a.n, "full": measure(full, "full mesh", a.reps, log, jh)}
    rng = np.random.default_rng(5)
    N = full.sampleMeshSize()
    gids = np.sort(rng.choice(N, size=int(round(a.sample * N)), replace=False))
    t0 = time.perf_counter()
    sm = create_sample_mesh(full, gids)
    log(f"[sedov3d] sample mesh ({len(gids)} rows) in {time.perf_counter() - t0:.1f} s")
    out["sample"] = measure(sm, f"{a.sample:.0%} sample mesh", a.reps, log)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
