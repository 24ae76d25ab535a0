"""Learned initial guesses for implicit Newton steps of periodic 2D viscous
Burgers (WENO5, BDF2), trained on trajectories this package's own solver
generates, with the device network of GuessNet.  The FP64 predictors of
y_{n+1} are the BDF2 relation with f(y_{n+1}) extrapolated in time with
order k from f_n, ..., f_{n-k} (linear in a fractional order t between
integer orders); the network picks the order t in [1, 4] per step, trained
on what the solver measured: for each training sample the Newton + Krylov
iterations c_j the step took from each of the orders 1, 1.5, ..., 4, which
the network learns to estimate (relative to their mean); it steps with the
order of the smallest estimate (a soft arg-min).  (Which
order pays changes with the flow: the highest on smooth fields, a low one
once steep fronts form, where higher-order extrapolation in time of a front
crossing a cell amplifies grid-scale error.)

Workload: 512 x 512 cells on [-1, 1]^2, diffusion nu ~ U(1e-3, 1e-2), Gaussian
pulse of amplitude U(0.5, 1.5) centred at U(-0.5, 0.5)^2 (spread 0.15), BDF2
with dt = 5e-3 (first step BDF1); Newton on the exact WENO5 Jacobian to
||R||_2 <= 1e-10 decides every step, whatever the guess.

    python -m pressiodemoapps_amd.learned train --out weights.npz   # generate data, train
    python -m pressiodemoapps_amd.learned compare --weights weights.npz

The weights file is a plain .npz (float32 parameters and float64 scales),
loaded without pickle.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import (AdvectionDiffusion2d, GuessNet, ImplicitScheme, ImplicitStepper,
               InviscidFluxReconstruction, StopRule, create_problem, create_structured_mesh)

DT = 5e-3
TOL = 1e-10
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                               "burgers_guessnet_512.npz")


def draw_params(rng):
    """One instance of the workload's parameter distribution."""
    return {"diffusion": float(rng.uniform(1e-3, 1e-2)),
            "pulseMagnitude": float(rng.uniform(0.5, 1.5)),
            "pulseX": float(rng.uniform(-0.5, 0.5)),
            "pulseY": float(rng.uniform(-0.5, 0.5)),
            "pulseSpread": 0.15}


def make_problem(mesh, params):
    return create_problem(mesh, AdvectionDiffusion2d.BurgersPeriodic, InviscidFluxReconstruction.Weno5,
                          1, dict(params))


def burgers_mesh(n):
    return create_structured_mesh((n, n), (-1.0, 1.0, -1.0, 1.0), 7, True)


ORDERS = GuessNet.ORDERS  # 1, 1.5, ..., 4: the orders a training sample is costed at
# binomial extrapolation weights of f_n, ..., f_{n-4} for integer orders 1..4
EXTRAP = {1: (2.0, -1.0, 0.0, 0.0, 0.0), 2: (3.0, -3.0, 1.0, 0.0, 0.0), 3: (4.0, -6.0, 4.0, -1.0, 0.0),
          4: (5.0, -10.0, 10.0, -5.0, 1.0)}
NAMED = {"bdf2": 1.0, "bdf2q": 2.0, "bdf2c": 3.0, "bdf2q4": 4.0}


def extrap_weights(t):
    """weights of f_n, ..., f_{n-4} in the order-t extrapolation, linear in
    t between floor(t) and floor(t) + 1 (t clamped to [1, 4])"""
    t = min(max(float(t), 1.0), 4.0)
    k = 3 if t >= 4.0 else int(t)
    w = t - k
    return tuple((1.0 - w) * a + w * b for a, b in zip(EXTRAP[k], EXTRAP[k + 1]))


def predictor(order, yn, ynm1, f, out):
    """FP64 guess: the BDF2 relation (4/3) y_n - (1/3) y_{n-1} + (2/3) dt
    f(y_{n+1}) with f(y_{n+1}) extrapolated in time with the given
    (possibly fractional) order from f = [f_n, f_{n-1}, ...]."""
    out.copy_(yn).mul_(4.0 / 3.0).add_(ynm1, alpha=-1.0 / 3.0)
    for c, fk in zip(extrap_weights(order), f):
        if c != 0.0:
            out.add_(fk, alpha=(2.0 / 3.0) * DT * c)
    return out


def bdf2_predictor(yn, ynm1, fn, fnm1, out):
    """the linear extrapolation, (4/3) y_n - (1/3) y_{n-1} + (2/3) dt (2 f_n - f_{n-1})"""
    return predictor(1, yn, ynm1, [fn, fnm1], out)


def _order_of(guess):
    if isinstance(guess, str):
        return NAMED[guess]
    return float(guess)


def run_trajectory(prob, nsteps, guess=None, nu=None, keep=None, tol=TOL, on_step=None, lin_rtol=1e-12,
                   orders=None):
    """BDF2 trajectory from the IC.  guess: None (previous step, y_n),
    "extrap" (2 y_n - y_{n-1}), an extrapolation order in [1, 4] (or one of
    the names "bdf2" / "bdf2q" / "bdf2c" / "bdf2q4" for 1 / 2 / 3 / 4; on
    the first steps the highest order the history allows) or a GuessNet (the
    cubic predictor until five velocities exist).  keep: set of step indices
    whose training samples are returned: (f = [f_k, ..., f_{k-4}], costs),
    costs[j] = Newton + Krylov iterations of the step solved from the
    predictor of order ORDERS[j] (the trajectory itself goes on from the
    cubic one).  orders: a list that receives the network's order per step.
    Returns (final state, per-step stats, kept samples)."""
    import torch
    ip = ImplicitStepper(prob, ImplicitScheme.BDF2)
    ip.set_options(StopRule.Residual, tol, 30, lin_rtol, 0)
    U = prob.initialCondition()
    prev = None
    stats, samples = [], []
    g = torch.empty_like(U)
    need_f = keep is not None or (guess is not None and not (isinstance(guess, str) and guess == "extrap"))
    hist = []  # f_k, f_{k-1}, ... (at most 5)
    for k in range(nsteps):
        yn = U.clone()
        if need_f:
            fn = hist.pop() if len(hist) == 5 else torch.empty_like(U)
            prob.rightHandSide(yn, k * DT, fn)
            hist.insert(0, fn)
        gs = None
        if prev is not None and guess is not None:
            if isinstance(guess, str) and guess == "extrap":
                torch.mul(yn, 2.0, out=g)
                g.sub_(prev)
            elif isinstance(guess, GuessNet):
                if len(hist) == 5:
                    guess.predict(yn, prev, hist, nu, out=g)
                    if orders is not None:
                        orders.append(guess.policy()[0])
                else:
                    predictor(min(3, len(hist) - 1), yn, prev, hist, g)
            else:
                predictor(min(_order_of(guess), len(hist) - 1), yn, prev, hist, g)
            gs = g
        if keep is not None and k in keep and prev is not None and len(hist) == 5:
            costs = []
            y = torch.empty_like(U)
            for t in ORDERS:
                y.copy_(yn)
                s = ip.step(y, k * DT, DT, prev=prev, guess=predictor(t, yn, prev, hist, torch.empty_like(U)))
                costs.append(s["newton_iters"] + s["linear_iters"])
            samples.append(([h.clone() for h in hist], np.asarray(costs, dtype=np.float64)))
        st = ip.step(U, k * DT, DT, prev=prev, guess=gs)
        stats.append(st)
        if on_step:
            on_step(k, st)
        prev = yn
    return U, stats, samples


def generate(n, n_traj, nsteps, keep_every, seed, log=print):
    """Training data: every keep_every-th step of n_traj trajectories, as
    (nu, f = [f_k, ..., f_{k-4}], costs at ORDERS)."""
    rng = np.random.default_rng(seed)
    mesh = burgers_mesh(n)
    data = []
    keep = set(range(keep_every, nsteps, keep_every))
    t0 = time.perf_counter()
    for i in range(n_traj):
        params = draw_params(rng)
        prob = make_problem(mesh, params)
        # the cubic predictor only speeds the Newton solves up: every step
        # converges to the same tolerance whatever the guess
        _, stats, tr = run_trajectory(prob, nsteps, guess="bdf2c", keep=keep)
        data += [(params["diffusion"], f, c) for f, c in tr]
        best = np.bincount([int(np.argmin(c)) for _, c in tr], minlength=len(ORDERS))
        log(f"[learned] trajectory {i + 1}/{n_traj} nu={params['diffusion']:.4f}: {len(tr)} samples, "
            f"cheapest order histogram {best.tolist()}, {time.perf_counter() - t0:.0f} s")
    return data


def init_params(net, seed):
    """He-style initial weights, small last layer."""
    from . import _guessnet_layers
    rng = np.random.default_rng(seed)
    w = np.zeros(net.nparams, dtype=np.float32)
    o = 0
    layers = _guessnet_layers()
    for li, (ci, co) in enumerate(layers):
        nW = 9 * ci * co
        std = np.sqrt(2.0 / (9 * ci)) * (0.1 if li == len(layers) - 1 else 1.0)
        w[o:o + nW] = rng.standard_normal(nW) * std
        o += nW + co
    net.set_params(w)


def fit(net, data, steps, lr, seed, log=print):
    """Adam on random samples, cosine learning-rate decay; the loss of a
    sample is the squared error of the network's cost estimates against the
    sample's measured costs (both relative to their mean over the orders)."""
    rng = np.random.default_rng(seed)
    losses = []
    for s in range(steps):
        nu, f, c = data[int(rng.integers(len(data)))]
        lr_s = lr * 0.5 * (1.0 + np.cos(np.pi * s / steps))
        want = (s % 100 == 0) or s == steps - 1
        loss = net.train_step(f, nu, c, max(lr_s, 1e-6), want_loss=want)
        if want:
            losses.append((s, loss))
            log(f"[learned] step {s}: loss {loss:.4e}")
    return losses


def evaluate(net, data):
    """Mean over samples of the excess iterations (over the sample's
    cheapest order) at the network's order, read off the sample's costs
    (linear between the costed orders), and the same for each fixed order."""
    exc, fixed = [], np.zeros(len(ORDERS))
    scratch = None
    for nu, f, c in data:
        # the policy depends on f and nu only; the guess formed from
        # (f_n, f_n) as stand-ins for (y_n, y_{n-1}) is discarded
        scratch = net.predict(f[0], f[0], f, nu, out=scratch)
        t = net.policy()[0]
        exc.append(float(np.interp(t, ORDERS, c) - c.min()))
        fixed += c - c.min()
    return float(np.mean(exc)), (fixed / len(data)).tolist()


def scales_from(data):
    """Spreads of dt D^k f for k = 1..4 over the data (the network's input
    scales)."""
    d = [[] for _ in range(4)]
    coef = ((1, -1), (1, -2, 1), (1, -3, 3, -1), (1, -4, 6, -4, 1))
    for _, f, _ in data:
        for k in range(4):
            dk = sum(c * fj for c, fj in zip(coef[k], f))
            d[k].append((DT * dk).flatten()[::97].cpu().numpy())
    return tuple(float(np.std(np.concatenate(x))) for x in d)


def save(path, net, meta):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    dt, spreads, nu_mean, nu_scale = net.scales
    np.savez(path, params=net.get_params(),
             scales=np.asarray([dt, *spreads, nu_mean, nu_scale], dtype=np.float64),
             meta=np.asarray(json.dumps(meta)))


def load(path, n):
    z = np.load(path, allow_pickle=False)
    net = GuessNet(n, n)
    sc = [float(v) for v in z["scales"]]
    net.set_scales(sc[0], sc[1:5], sc[5], sc[6])
    net.set_params(z["params"])
    return net, json.loads(str(z["meta"]))


def _fixed_name(t):
    return f"predictor_order_{t:g}"


def compare(net, n, nsteps, seed, log=print, lin_rtol=1e-12):
    """Same held-out instance stepped with the previous-step guess, the
    linear extrapolation 2 y_n - y_{n-1}, the FP64 predictor at each fixed
    order 1, 1.5, ..., 4 and the learned guess (the network's order per
    step); every run converges each step to ||R|| <= TOL with the same
    Newton and linear-solver settings.  Returns a dict of steps/s, Newton
    and linear iterations, and the final-state differences to the
    previous-step (cold start) run."""
    import torch
    rng = np.random.default_rng(seed)
    params = draw_params(rng)
    mesh = burgers_mesh(n)
    prob = make_problem(mesh, params)
    out = {"params": params, "steps": nsteps, "lin_rtol": lin_rtol}
    finals = {}
    methods = [("previous_step", None), ("extrapolation", "extrap")]
    methods += [(_fixed_name(t), t) for t in ORDERS]
    methods += [("learned", net)]
    for name, g in methods:
        # warm-up (library set-up, first-touch allocations)
        run_trajectory(prob, 6, guess=g, nu=params["diffusion"], lin_rtol=lin_rtol)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        U, stats, _ = run_trajectory(prob, nsteps, guess=g, nu=params["diffusion"], lin_rtol=lin_rtol)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        orders = None
        if name == "learned":
            # the network's orders, from a second (untimed) run: reading
            # them back each step would add a host synchronization
            orders = []
            run_trajectory(prob, nsteps, guess=g, nu=params["diffusion"], lin_rtol=lin_rtol, orders=orders)
        finals[name] = U.cpu().numpy()
        out[name] = {"steps_per_s": nsteps / el,
                     "newton_per_step": float(np.mean([s["newton_iters"] for s in stats])),
                     "linear_per_step": float(np.mean([s["linear_iters"] for s in stats])),
                     "median_initial_residual": float(np.median([s["initial_residual_norm"] for s in stats[2:]])),
                     "max_final_residual": float(max(s["residual_norm"] for s in stats))}
        if orders:
            o = np.asarray(orders)
            out[name]["order_mean"] = float(o.mean())
            out[name]["order_min"] = float(o.min())
            out[name]["order_max"] = float(o.max())
        log(f"[learned] {name}: {out[name]}")
    ref = finals["previous_step"]
    for name, _ in methods[1:]:
        out[name]["final_state_max_abs_diff_vs_cold_start"] = float(np.abs(finals[name] - ref).max())
    sp = lambda a, b: out[a]["steps_per_s"] / out[b]["steps_per_s"]  # noqa: E731
    best = max((_fixed_name(t) for t in ORDERS), key=lambda k: out[k]["steps_per_s"])
    out["best_fixed_order"] = best
    out["speedup_learned_vs_previous_step"] = sp("learned", "previous_step")
    out["speedup_learned_vs_quadratic_predictor"] = sp("learned", _fixed_name(2.0))
    out["speedup_learned_vs_best_fixed_order"] = sp("learned", best)
    for t in ORDERS:
        out[f"speedup_{_fixed_name(t)}_vs_previous_step"] = sp(_fixed_name(t), "previous_step")
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(prog="pressiodemoapps_amd.learned")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--n", type=int, default=512)
    t.add_argument("--trajectories", type=int, default=48)
    t.add_argument("--steps", type=int, default=400)
    t.add_argument("--keep-every", type=int, default=10)
    t.add_argument("--train-steps", type=int, default=3000)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--seed", type=int, default=2024)
    t.add_argument("--out", default=DEFAULT_WEIGHTS)
    c = sub.add_parser("compare")
    c.add_argument("--n", type=int, default=512)
    c.add_argument("--steps", type=int, default=400)
    c.add_argument("--seed", type=int, default=77)
    c.add_argument("--weights", default=DEFAULT_WEIGHTS)
    a = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    if a.cmd == "train":
        data = generate(a.n, a.trajectories, a.steps, a.keep_every, a.seed, log)
        spreads = scales_from(data)
        net = GuessNet(a.n, a.n)
        net.set_scales(DT, spreads, 5.5e-3, 4.5e-3)
        init_params(net, a.seed + 1)
        losses = fit(net, data, a.train_steps, a.lr, a.seed + 2, log)
        meta = {"n": a.n, "trajectories": a.trajectories, "steps": a.steps, "keep_every": a.keep_every,
                "samples": len(data), "train_steps": a.train_steps, "lr": a.lr, "seed": a.seed,
                "spreads": list(spreads), "final_loss": losses[-1][1],
                "mean_loss_last_1000": float(np.mean([l for st, l in losses if st >= a.train_steps - 1000]))}
        exc, fixed = evaluate(net, data)
        meta["training_excess_iterations_learned"] = exc
        meta["training_excess_iterations_fixed_orders"] = fixed
        save(a.out, net, meta)
        print(json.dumps(meta))
    else:
        net, meta = load(a.weights, a.n)
        print(json.dumps(compare(net, a.n, a.steps, a.seed, log)))


if __name__ == "__main__":
    main()
