"""The trained guess network (package data burgers_guessnet_512.npz) on a
held-out 512x512 periodic Burgers instance, BDF2 dt 5e-3 with Newton on the
exact WENO5 Jacobian to ||R|| <= 1e-10: every step converges with either
guess, the network starts Newton much closer to the solution than the
previous-step guess and needs no more Newton iterations, and the final
states agree to the solver tolerance."""
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


def test_learned_guess_beats_previous_step_guess(torch):
    from pressiodemoapps_amd import learned as L
    assert os.path.exists(L.DEFAULT_WEIGHTS)
    n, nsteps = 512, 40
    net, meta = L.load(L.DEFAULT_WEIGHTS, n)
    params = L.draw_params(np.random.default_rng(987654))  # not a training seed
    prob = L.make_problem(L.burgers_mesh(n), params)
    print(f"[gpu-learned] instance {params}", flush=True)
    runs = {}
    for name, g in (("previous_step", None), ("learned", net)):
        U, stats, _ = L.run_trajectory(prob, nsteps, guess=g, nu=params["diffusion"])
        torch.cuda.synchronize()
        runs[name] = (U.cpu().numpy(), stats)
        r0 = np.median([s["initial_residual_norm"] for s in stats[2:]])
        nw = np.mean([s["newton_iters"] for s in stats[2:]])
        print(f"[gpu-learned] {name}: median initial ||R|| {r0:.3e}, Newton/step {nw:.2f}", flush=True)
    for name, (_, stats) in runs.items():
        assert all(s["converged"] and s["residual_norm"] <= L.TOL for s in stats), name
    prev, lrn = runs["previous_step"][1][2:], runs["learned"][1][2:]
    r_prev = np.median([s["initial_residual_norm"] for s in prev])
    r_lrn = np.median([s["initial_residual_norm"] for s in lrn])
    assert r_lrn <= 0.1 * r_prev
    assert np.mean([s["newton_iters"] for s in lrn]) <= np.mean([s["newton_iters"] for s in prev])
    diff = np.abs(runs["learned"][0] - runs["previous_step"][0]).max()
    print(f"[gpu-learned] final state max |diff| {diff:.2e}", flush=True)
    assert diff <= 1e-9
