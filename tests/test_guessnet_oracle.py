"""CPU pinning of the learned-guess network's oracle: its FP64 gradient
against central finite differences of its FP64 loss, and the BF16 rounding
helper against exact cases."""
import numpy as np

import guessnet_oracle as G


def test_bf16_rounding():
    # 1 + 2^-8 is exactly between 1 and 1 + 2^-7: ties to even -> 1
    assert G.bf16(1.0 + 2.0 ** -8) == 1.0
    assert G.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6
    assert G.bf16(-2.5) == -2.5
    x = np.random.default_rng(0).standard_normal(1000)
    r = G.bf16(x)
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8 + 1e-300)


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    H, W = 6, 5
    w = rng.standard_normal(G.num_params()) * 0.2
    un = rng.standard_normal(H * W * 2)
    unm1 = un + 0.01 * rng.standard_normal(H * W * 2)
    unp1 = un + 0.02 * rng.standard_normal(H * W * 2)
    fn = rng.standard_normal(H * W * 2)
    fnm1 = fn + 0.1 * rng.standard_normal(H * W * 2)
    args = (un, unm1, fn, fnm1, unp1, 0.004, H, W, 5e-3, 1e-2, 1e-3, 1e-2, 5.5e-3, 4.5e-3)
    kw = {}
    _, _, g = G.loss_and_grad(w, *args, bf=False, **kw)
    idx = rng.choice(len(w), size=60, replace=False)
    h = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (G.loss_and_grad(wp, *args, bf=False, **kw)[0] - G.loss_and_grad(wm, *args, bf=False, **kw)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_oracle_forward_closed_form():
    # with only the centre tap of the first layer's channel 0 -> 0 set and
    # identity-like later layers, the output is a known function of u
    H, W = 4, 4
    w = np.zeros(G.num_params())
    P = []
    o = 0
    for ci, co in G.LAYERS:
        P.append(o)
        o += 9 * co * ci + co
    # layer 0: W[4][0][0] = 1 (centre tap, feature u -> channel 0)
    w[P[0] + (4 * 32 + 0) * 8 + 0] = 1.0
    # layers 1, 2: centre tap channel 0 -> 0 weight 1
    for li in (1, 2):
        w[P[li] + (4 * 32 + 0) * 32 + 0] = 1.0
    # layer 3: centre tap channel 0 -> output 0 weight 1
    w[P[3] + (4 * 2 + 0) * 32 + 0] = 1.0
    rng = np.random.default_rng(1)
    un = rng.standard_normal(H * W * 2)
    f = rng.standard_normal(H * W * 2)
    X0 = G.features(un, f, f, 0.0055, H, W, 5e-3, 1e-2, 1e-3, 5.5e-3, 4.5e-3, rb=False)
    out, _ = G.forward(w, X0, bf=False)
    s = lambda z: z / (1 + np.exp(-z))  # noqa: E731
    expect = s(s(s(un.reshape(H, W, 2)[..., 0])))
    assert np.allclose(out[..., 0], expect, rtol=1e-14, atol=1e-14)
    assert np.allclose(out[..., 1], 0.0)


def test_base_is_exact_for_quadratic_trajectories():
    # y' = a + b t has the quadratic solution y0 + a t + b t^2 / 2, which BDF2
    # integrates exactly; f is linear in t, so its linear extrapolation is
    # exact too and the base must equal y(t_{n+1}) up to rounding
    rng = np.random.default_rng(4)
    y0, a, b = rng.standard_normal((3, 50))
    y = lambda t: y0 + a * t + 0.5 * b * t * t  # noqa: E731
    f = lambda t: a + b * t  # noqa: E731
    for dt, tn in ((5e-3, 0.3), (0.1, 1.7), (1.0, -2.0)):
        got = G.base(y(tn), y(tn - dt), f(tn), f(tn - dt), dt)
        assert np.allclose(got, y(tn + dt), rtol=1e-13, atol=1e-13)
    # a cubic term is not reproduced: the base is second order, not exact
    c = 0.7
    yc = lambda t: y(t) + c * t ** 3  # noqa: E731
    fc = lambda t: f(t) + 3 * c * t ** 2  # noqa: E731
    dt, tn = 0.1, 1.0
    err = G.base(yc(tn), yc(tn - dt), fc(tn), fc(tn - dt), dt) - yc(tn + dt)
    assert np.all(np.abs(err) > 1e-6)
