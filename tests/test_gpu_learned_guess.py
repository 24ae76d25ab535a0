"""The trained guess network (package data burgers_guessnet_512.npz) on two
held-out 512x512 periodic Burgers instances, BDF2 dt 5e-3 with Newton on the
exact WENO5 Jacobian to ||R|| <= 1e-10: every step converges with each
guess and the final states agree to the solver tolerance.

The network picks the extrapolation order of the FP64 predictor per step
from its estimates of what each order would cost.  Asserted: it needs fewer
Newton + Krylov iterations per step than the quadratic predictor (order 2)
and than the previous-step guess on both instances, and it picks a higher
order on the viscous instance (smooth fields, where high-order
extrapolation pays) than on the nearly inviscid one (steep fronts, where it
does not).  Its wall-clock rate is reported by learned.compare and
bench.py."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    if not _t.cuda.is_available():
        pytest.skip("no HIP device")
    return _t


def _instances(L):
    viscous = L.draw_params(np.random.default_rng(987654))  # not a training seed
    fronts = dict(L.draw_params(np.random.default_rng(424242)), diffusion=0.0025)
    return {"viscous": viscous, "fronts": fronts}


def test_learned_guess_gain(torch):
    from pressiodemoapps_amd import learned as L
    assert os.path.exists(L.DEFAULT_WEIGHTS)
    n, nsteps = 512, 150
    net, meta = L.load(L.DEFAULT_WEIGHTS, n)
    mesh = L.burgers_mesh(n)
    orders_mean = {}
    for inst, params in _instances(L).items():
        prob = L.make_problem(mesh, params)
        print(f"[gpu-learned] {inst}: {params}", flush=True)
        runs = {}
        for name, g in (("previous_step", None), ("quadratic", 2.0), ("learned", net)):
            orders = []
            U, stats, _ = L.run_trajectory(prob, nsteps, guess=g, nu=params["diffusion"], orders=orders)
            torch.cuda.synchronize()
            s = stats[5:]  # the learned guess starts at step 5 (five velocities)
            runs[name] = {"U": U.cpu().numpy(), "stats": stats,
                          "newton": float(np.mean([x["newton_iters"] for x in s])),
                          "linear": float(np.mean([x["linear_iters"] for x in s]))}
            r = runs[name]
            r["cost"] = r["newton"] + r["linear"]
            extra = ""
            if orders:
                orders_mean[inst] = float(np.mean(orders))
                extra = f", order mean {orders_mean[inst]:.2f} (min {min(orders):.2f}, max {max(orders):.2f})"
            print(f"[gpu-learned] {inst} {name}: Newton/step {r['newton']:.3f}, linear/step {r['linear']:.2f}"
                  f"{extra}", flush=True)
        for name, r in runs.items():
            assert all(x["converged"] and x["residual_norm"] <= L.TOL for x in r["stats"]), (inst, name)
        prev, quad, lrn = runs["previous_step"], runs["quadratic"], runs["learned"]
        assert lrn["cost"] < prev["cost"]
        assert lrn["cost"] <= quad["cost"], (inst, lrn["cost"], quad["cost"])
        for name in ("quadratic", "learned"):
            diff = np.abs(runs[name]["U"] - prev["U"]).max()
            print(f"[gpu-learned] {inst} {name}: final state max |diff| vs previous-step run {diff:.2e}", flush=True)
            assert diff <= 1e-9
    assert orders_mean["viscous"] > orders_mean["fronts"] + 0.5
