"""Batched shallow-water workload: many independent slip-wall instances on
[-5, 5]^2 with grids drawn from {64, 96, 128, 192, 256}^2, gravity
U(7.5, 12.5), Coriolis U(-4, -2) and a Gaussian height pulse of amplitude
U(0.05, 0.2) and unit width on unit depth (centre (1, 1)), WENO5 and
Rusanov fluxes in FP64, stepped together with RK4 (dt = 0.01) by one batched
velocity launch per stage, several steps per hipGraph replay."""
import time

import numpy as np

from . import Batch, InviscidFluxReconstruction, StepScheme, Swe2d, create_problem, create_structured_mesh

SIZES = (64, 96, 128, 192, 256)


def build(n_inst=384, seed=7, scheme=InviscidFluxReconstruction.Weno5):
    rng = np.random.default_rng(seed)
    meshes = {n: create_structured_mesh((n, n), (-5.0, 5.0, -5.0, 5.0), 7) for n in SIZES}
    probs, sizes = [], []
    for _ in range(n_inst):
        n = int(rng.choice(SIZES))
        params = {"gravity": float(rng.uniform(7.5, 12.5)), "coriolis": float(rng.uniform(-4.0, -2.0)),
                  "pulseMagnitude": float(rng.uniform(0.05, 0.2))}
        probs.append(create_problem(meshes[n], Swe2d.SlipWall, scheme, 1, params))
        sizes.append(n)
    return Batch(probs), meshes, sizes


def run(batch, sizes, steps=500, dt=0.01, steps_per_graph=10, log=print):
    import torch
    U0 = batch.initialCondition()
    U = U0.clone()
    rps, vel_us = batch.autotune(U)
    log(f"[batched] autotuned rows per strip {rps}: batched velocity {vel_us:.1f} us")
    # record and warm the graph on the tensor that is timed (the recorded
    # graph is tied to the state's address), then restore the initial state
    batch.advance(U, StepScheme.RK4, dt, steps_per_graph, steps_per_graph=steps_per_graph)
    U.copy_(U0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    batch.advance(U, StepScheme.RK4, dt, steps, steps_per_graph=steps_per_graph)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    cells = int(sum(n * n for n in sizes))
    return {"instances": len(sizes), "cells": cells, "steps": steps, "dt": dt, "wall_s": wall,
            "steps_per_s": steps / wall, "cell_steps_per_s": cells * steps / wall,
            "rhs_evals_per_s": 4 * steps / wall, "rows_per_strip": rps, "batched_velocity_us": vel_us,
            "finite": bool(torch.isfinite(U).all().item())}
